"""MI355X-native photon-transport engine for DXMClib's voxel dose path.

Host API mirrors the reference (World / AAVoxelGrid / Transport /
TransportProgress / beams); compute runs in libdxmc_hip.so (HIP, gfx950).
"""
from .engine import (AAVoxelGrid, BowtieFilter, CBCTBeam, CTAECFilter, CTSequentialBeam, CTSpiralBeam,
                     CTSpiralDualEnergyBeam, DXBeam, Material, PencilBeam, Transport, TransportBatch, TransportProgress,
                     Tube, World)

__all__ = ["AAVoxelGrid", "BowtieFilter", "CBCTBeam", "CTAECFilter", "CTSequentialBeam", "CTSpiralBeam",
           "CTSpiralDualEnergyBeam", "DXBeam", "Material", "PencilBeam", "Transport", "TransportBatch", "TransportProgress",
           "Tube", "World"]
