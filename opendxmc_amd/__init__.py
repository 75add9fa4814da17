"""MI355X-native photon-transport engine for DXMClib's voxel dose path.

Host API mirrors the reference (World / AAVoxelGrid / Transport /
TransportProgress / beams) and OpenDXMC's CT segmentation and dose table
pipelines; compute runs in libdxmc_hip.so (HIP, gfx950).
"""
from .engine import (AAVoxelGrid, BowtieFilter, CBCTBeam, CTAECFilter, CTOrganAECFilter, CTSequentialBeam, CTSpiralBeam,
                     CTSpiralDualEnergyBeam, DXBeam, Material, PencilBeam, Transport, TransportBatch, TransportProgress,
                     Tube, World)
from .pipelines import CTSegmentationPipeline, DoseTablePipeline

__all__ = ["AAVoxelGrid", "BowtieFilter", "CBCTBeam", "CTAECFilter", "CTOrganAECFilter", "CTSequentialBeam", "CTSpiralBeam",
           "CTSegmentationPipeline", "CTSpiralDualEnergyBeam", "DoseTablePipeline", "DXBeam", "Material", "PencilBeam", "Transport", "TransportBatch", "TransportProgress",
           "Tube", "World"]
