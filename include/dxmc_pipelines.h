/* Image pipelines on either side of the transport (part of libdxmc_hip.so):
 * CT number segmentation into material and density before a run, and the
 * organ dose table after it.  OpenDXMC runs both on the host
 * (ctsegmentationpipeline.cpp:61-166, dosetablepipeline.cpp:34-87); here each
 * is one pass of a HIP kernel over the voxels.  Status codes and
 * dxg_last_error() as in dxmc_hip.h.
 *
 * Every function exists twice: the plain name takes host arrays and stages
 * them through device memory; the _device variant takes device pointers
 * (hipMalloc'd memory or torch tensors on the current device) and a
 * hipStream_t (NULL = the default stream), and returns once its results are
 * complete.
 */
#ifndef DXMC_PIPELINES_H
#define DXMC_PIPELINES_H

#include "dxmc_beams.h"

#ifdef __cplusplus
extern "C" {
#endif

#define DXG_SEG_MAX_MATERIALS 128

/* getMeanCTNumbers (ctsegmentationpipeline.cpp:61-115) and the constants the
 * per-voxel map of ctsegmentationpipeline.cpp:134-155 reads.  All attenuation
 * values are averages over the normalised tube spectrum of
 * mu/rho = photoelectric + incoherent + coherent (cm2/g). */
typedef struct dxg_ct_segmentation {
    uint32_t n_materials;                            /* 1..128 */
    uint32_t reserved;
    double hu[DXG_SEG_MAX_MATERIALS];                /* 1000 <(mu_m - mu_w) / (mu_w - mu_a)>, mu = mu/rho x density */
    double separators[DXG_SEG_MAX_MATERIALS - 1];    /* (hu[i] + hu[i+1]) / 2 */
    double attenuation[DXG_SEG_MAX_MATERIALS];       /* <mu/rho> of each material */
    double water_attenuation, air_attenuation;       /* <mu/rho> of water and dry air */
    double water_density, air_density;               /* g/cm3 */
    /* density = max((slope hu / 1000 + offset) / attenuation[m], 0) with
     * slope = water_attenuation water_density - air_attenuation air_density and
     * offset = water_attenuation water_density (ctsegmentationpipeline.cpp:148) */
    double slope, offset;
} dxg_ct_segmentation;

/* CTSegmentationPipeline::updateImageData's tables for the acquisition tube
 * (Tube(kv) with setAlFiltration / setSnFiltration, :63-68; pass the tube with
 * its filtration) and the candidate materials in the order the separators are
 * taken (:120-127: dry air, adipose, soft tissue, skeletal muscle, cortical
 * bone with density 1.09).  water and air define the CT number scale. */
int dxg_ct_segmentation_tables(const dxg_tube* tube, const dxg_material* materials, const double* densities,
                               uint64_t n_materials, const dxg_material* water, double water_density,
                               const dxg_material* air, double air_density, dxg_ct_segmentation* out);

/* The per-voxel map (ctsegmentationpipeline.cpp:134-152): material[v] = the
 * first i with hu[v] < separators[i] (n_materials - 1 when none), density[v]
 * as above.  NaN CT numbers take the last material and a NaN density, as
 * std::max(NaN, 0.0) does. */
int dxg_ct_segment(const dxg_ct_segmentation* t, const double* hu, uint64_t n, uint8_t* material, double* density);
int dxg_ct_segment_device(const dxg_ct_segmentation* t, const double* hu, uint64_t n, uint8_t* material,
                          double* density, void* stream);

/* One row of DoseTablePipeline::updateImageData (dosetablepipeline.cpp:72-86)
 * for organ index i: number of voxels, volume (count x voxel volume), mass
 * (voxel volume x sum of density) and dose = energy / mass with energy the sum
 * of dose x voxel volume x density over the organ's voxels (in the dose
 * array's units).  Organs without voxels have n_voxels = 0 (the reference
 * skips their rows) and zeros elsewhere. */
typedef struct dxg_organ_dose {
    uint64_t n_voxels;
    double volume_cm3;
    double mass_g;
    double dose;
} dxg_organ_dose;

/* The whole table in one pass over the voxels: out[256] indexed by organ
 * (uint8).  spacing in cm.  Sums are formed in a fixed order for a given n and
 * device, except inside a group of 64 consecutive voxels made of more than
 * four runs of one organ (speckle rather than organ boundaries), whose voxels
 * are added by LDS atomics in the order the hardware serves them. */
int dxg_organ_dose_table(const uint8_t* organ, const double* dose, const double* density, uint64_t n,
                         const double spacing_cm[3], dxg_organ_dose out[256]);
int dxg_organ_dose_table_device(const uint8_t* organ, const double* dose, const double* density, uint64_t n,
                                const double spacing_cm[3], dxg_organ_dose out[256], void* stream);

/* Mean duration (ms) of the last launch of each kernel on this thread, from
 * HIP events around it on its stream: which = 0 segmentation, 1 dose table
 * (both of its kernels). */
double dxg_pipeline_last_ms(int which);

#ifdef __cplusplus
}
#endif

#endif /* DXMC_PIPELINES_H */
