import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def _ensure_built():
    from opendxmc_amd import build as b
    if not b.PHYS.exists():
        b.build_physics_list()
    if not b.LIB.exists():
        b.build_library()
    if not b.ORACLE_LIB.exists():
        b.build_oracle()


_ensure_built()


def gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def oracle():
    import oracle as o
    o.lib()
    return o
