"""Bowtie filter table reader (the JSON format of bowtiefilterreader.cpp:51-97).

Format: {"filters": [{"name": str, "filterdata": [{"angle": rad, "weight": w}, ...]}]}.
Entries without a name, without data or with non-numeric fields are skipped,
as the reference does.  The packaged file holds the filters of the reference's
data/bowtiefilters/bowtiefilters.json used here, including the CT default
"Siemens Definition Flash W1 120kV" (beamsettingsmodel.cpp:1303).
"""
from __future__ import annotations

import json
from pathlib import Path

from .engine import BowtieFilter

PACKAGED = Path(__file__).resolve().parent / "data" / "bowtie_filters.json"
DEFAULT_CT_BOWTIE = "Siemens Definition Flash W1 120kV"


def read_filters(path: str | Path = PACKAGED) -> dict[str, BowtieFilter]:
    try:
        doc = json.loads(Path(path).read_text())
    except (OSError, json.JSONDecodeError):
        return {}
    if not isinstance(doc, dict) or not isinstance(doc.get("filters"), list):
        return {}
    out: dict[str, BowtieFilter] = {}
    for f in doc["filters"]:
        if not isinstance(f, dict):
            continue
        name = f.get("name")
        data = []
        for d in f.get("filterdata", []) or []:
            if isinstance(d, dict):
                a, w = d.get("angle"), d.get("weight")
                if isinstance(a, (int, float)) and isinstance(w, (int, float)) and not isinstance(a, bool):
                    data.append((float(a), float(w)))
        if data and isinstance(name, str) and name:
            out[name] = BowtieFilter(data)
    return out
