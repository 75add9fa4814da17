/* Beam geometry and X-ray tube model of the MI355X transport engine
 * (part of libdxmc_hip.so).  These functions turn the settings of DXMClib's
 * beam classes -- as OpenDXMC sets them in beamsettingsmodel.cpp and stores
 * them in hdf5wrapper.cpp:548-585 -- into the exposure list and spectra that
 * dxg_transport_run() consumes (include/dxmc_hip.h).  All lengths in cm,
 * angles in radians, energies in keV.  Status codes as in dxmc_hip.h.
 *
 * Geometry convention (DXMClib's own is not in the reference tree):
 *   CT and CBCT sources turn about a rotation axis n through an isocentre c at
 *   radius R = SDD / 2.  At gantry angle a the beam direction is
 *   b(a) = e1 cos a + e2 sin a with e1 = normalise(y x n) (x x n if n || y),
 *   e2 = n x e1; the source sits at c - R b(a); the exposure cosines are
 *   c0 = n x b(a) and c1 = n, so that c0 x c1 = b(a) (beamactorcontainer.cpp:52).
 *   The fan half angle is asin(FOV / SDD) (the fan touches the scan field of
 *   view circle) and the slice half angle atan(collimation / SDD).
 */
#ifndef DXMC_BEAMS_H
#define DXMC_BEAMS_H

#include "dxmc_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Tube(voltage, anodeAngle) with addTubeFiltrationMaterial(Z, mm)
 * (beamsettingsmodel.cpp:257-347).  Model: Kramers thick-target
 * bremsstrahlung (kVp - E)/E photons per keV, anode self-absorption through
 * half the Kanaya-Okayama electron range divided by sin(anode angle), tungsten
 * K lines above 69.525 kV with Green-Cosslett (kVp/E_K - 1)^1.63 intensity, and
 * Beer-Lambert added filtration.  energy_step_kev is the bin width (0 = 0.5). */
typedef struct dxg_tube {
    double voltage_kv;
    double anode_angle_deg;
    const int32_t* filtration_Z;
    const double* filtration_mm;
    uint64_t n_filtration;
    double energy_step_kev;
} dxg_tube;

/* Tube::getSpecter: bin energies and photon weights normalised to sum 1.
 * Call with energy = weight = NULL to get the count in *n. */
int dxg_tube_spectrum(const dxg_tube* tube, double* energy_kev, double* weight, uint64_t cap, uint64_t* n);
/* Beam::tubeAlHalfValueLayer() (mm Al, air-kerma weighted) and
 * Beam::tubeMeanSpecterEnergy() (keV, photon weighted) (beamsettingsmodel.cpp:331-343). */
int dxg_tube_hvl_mm_al(const dxg_tube* tube, double* hvl_mm);
int dxg_tube_mean_energy(const dxg_tube* tube, double* mean_kev);

/* CTAECFilter(start, stop, weights) (datacontainer.cpp:59): tube current
 * weights along start->stop, linearly interpolated at the projection of the
 * exposure position and normalised to mean one.  n = 0 means no AEC. */
typedef struct dxg_aec {
    double start[3];
    double stop[3];
    const double* weights;
    uint64_t n;
} dxg_aec;

/* The organ AEC filter of CT spiral beams (organAECFilter(), settings at
 * beamsettingsmodel.cpp:349-437, stored at hdf5wrapper.cpp:576-580): tube
 * current modulated by gantry angle.  Angles in radians, measured as the
 * beam's start_angle (the gantry angle a of the exposure's tube).  Over the
 * arc from start_angle counter-clockwise to stop_angle the weight is
 * low_weight; over ramp_angle on each side outside the arc it rises linearly
 * to the high weight, which holds elsewhere.  The high weight is 1, or with
 * compensate_outside the value that keeps the mean weight over a full turn at
 * 1: (2 pi - low (L + r)) / (2 pi - L - r) for arc L and ramp r (1 when the
 * arc and ramps cover the whole turn).  use = 0 gives weight 1 everywhere.
 * DXMClib's own filter is not in the reference tree: this is the build's
 * definition from those settings. */
typedef struct dxg_organ_aec {
    int32_t use;
    int32_t compensate_outside;
    double start_angle;
    double stop_angle;
    double ramp_angle;
    double low_weight;
} dxg_organ_aec;

/* The filter's weight at gantry angle `angle` (rad), and its high weight
 * (maxWeight(), the "High weight" item of beamsettingsmodel.cpp:423-428). */
int dxg_organ_aec_weight(const dxg_organ_aec* f, double angle, double* weight);
int dxg_organ_aec_max_weight(const dxg_organ_aec* f, double* weight);

/* CTAECFilter from the water-equivalent diameter per slice
 * (DataContainer::calculateAECfilterFromWaterEquivalentDiameter,
 * datacontainer.cpp:42-100): for slice z of a dims[0] x dims[1] x dims[2] image
 * (index x + y nx + z nx ny, spacing in cm), the water-equivalent area is
 * A_w = (mean HU / 1000 + 1) nx ny dx dy for a CT image (image_is_ct = 1) or
 * sum(density) dx dy for a density image (g/cm3), the diameter
 * D_w = 2 sqrt(max(A_w, 0) / pi), and the weight exp(0.2 D_w) (mu of water
 * about 0.2 /cm).  weights receives dims[2] values; the filter runs from
 * (0, 0, -L) to (0, 0, L), L = dims[2] dz / 2, written to start and stop. */
int dxg_aec_from_wed(const uint64_t dims[3], const double spacing[3], const double* image, int image_is_ct, double* weights,
                     double start[3], double stop[3]);

/* PencilBeam<false>: position, direction, energy (20..150 keV clamp,
 * beamsettingsmodel.cpp:664), numberOfExposures, numberOfParticlesPerExposure.
 * Uses spectrum 0, no bowtie. */
typedef struct dxg_pencil_beam {
    double position[3];
    double direction[3];
    uint64_t n_exposures;
    uint64_t particles_per_exposure;
} dxg_pencil_beam;

/* DXBeam<false>: position, directionCosines, collimationHalfAngles,
 * numberOfExposures (dxmc_specialization.cpp:21-90 computes position and
 * cosines from rotation centre, SPD and angles). */
typedef struct dxg_dx_beam {
    double position[3];
    double cosines[6];
    double collimation_half_angles[2];
    uint64_t n_exposures;
    uint64_t particles_per_exposure;
} dxg_dx_beam;

/* CTSpiralBeam<false> / CTSpiralDualEnergyBeam<false> (beamsettingsmodel.cpp:1150-1837,
 * hdf5wrapper.cpp:548-585).  The gantry turns about start->stop; the table
 * advances pitch * collimation per turn.  Exposure i sits at angle
 * start_angle + i * step_angle; its weight is the AEC weight at its position
 * times the organ AEC weight at its tube's angle (times the tube's relative
 * mAs for dual source).  For dual energy, tube B is offset by
 * tube_b_offset_angle and exposures interleave A, B (beamactorcontainer.cpp:135-146),
 * tube A uses spectrum/bowtie 0 with weight relative_mas_a, tube B spectrum/bowtie 1
 * with weight relative_mas_b; fov_b is tube B's field of view.  has_bowtie = 0
 * leaves bowtie_index -1. */
typedef struct dxg_ct_spiral_beam {
    double start[3];
    double stop[3];
    double start_angle;
    double step_angle;
    double pitch;
    double collimation_cm;
    double sdd_cm;
    double fov_cm;
    uint64_t particles_per_exposure;
    int32_t dual_source;
    int32_t has_bowtie;
    double fov_b_cm;
    double tube_b_offset_angle;
    double relative_mas_a;
    double relative_mas_b;
    dxg_aec aec;
    dxg_organ_aec organ_aec;   /* applied to each tube at its own gantry angle */
} dxg_ct_spiral_beam;

/* CTSequentialBeam<false> (beamsettingsmodel.cpp:900-1150): number_of_slices
 * axial rotations at position + k * slice_spacing * scan_normal, each from
 * start_angle in step_angle steps over one full turn. */
typedef struct dxg_ct_sequential_beam {
    double position[3];
    double scan_normal[3];
    uint64_t number_of_slices;
    double slice_spacing_cm;
    double start_angle;
    double step_angle;
    double collimation_cm;
    double sdd_cm;
    double fov_cm;
    uint64_t particles_per_exposure;
    int32_t has_bowtie;
    int32_t pad;
} dxg_ct_sequential_beam;

/* CBCTBeam<false> (beamsettingsmodel.cpp:720-900): isocenter, rotationAxis,
 * SDD, start/stop/step angle, collimationHalfAngles. */
typedef struct dxg_cbct_beam {
    double isocenter[3];
    double rotation_axis[3];
    double start_angle;
    double stop_angle;
    double step_angle;
    double sdd_cm;
    double collimation_half_angles[2];
    uint64_t particles_per_exposure;
} dxg_cbct_beam;

/* Exposure builders: numberOfExposures() with out = NULL, else fill up to cap. */
int dxg_pencil_exposures(const dxg_pencil_beam* b, dxg_exposure* out, uint64_t cap, uint64_t* n);
int dxg_dx_exposures(const dxg_dx_beam* b, dxg_exposure* out, uint64_t cap, uint64_t* n);
int dxg_ct_spiral_exposures(const dxg_ct_spiral_beam* b, dxg_exposure* out, uint64_t cap, uint64_t* n);
int dxg_ct_sequential_exposures(const dxg_ct_sequential_beam* b, dxg_exposure* out, uint64_t cap, uint64_t* n);
int dxg_cbct_exposures(const dxg_cbct_beam* b, dxg_exposure* out, uint64_t cap, uint64_t* n);

/* ---- Beam calibration (the beams' DAPvalue / airKerma / CTDIw / CTDIvol
 * settings, beamsettingsmodel.cpp:574-583, 673-681, 1054-1071, 1276-1295,
 * applied through Transport's useBeamCalibration, simulationpipeline.cpp:165).
 * Each returns the factor f that dxg_beam.calibration takes: f x (dose summed
 * over the run's histories) is the dose in mGy for the stated output.  The
 * beam's exposures carry the particle counts of the run to be calibrated.
 *
 * DX / CBCT: the dose-area product (mGy cm2) of the whole beam.  The simulated
 * photons' air-KERMA-area product is sum over histories of w E mu_tr/rho(air)(E)
 * (per exposure: n x weight x spectrum mean), mu_tr from the physics list with
 * the Klein-Nishina energy-transfer fraction, dry air (C, N, O, Ar). */
int dxg_calibration_dap(const dxg_beam* beam, double dap_mgycm2, double* factor);
/* Pencil beam: the air KERMA (mGy) over a 1 cm2 cross-section of the beam. */
int dxg_calibration_air_kerma(const dxg_beam* beam, double air_kerma_mgy, double* factor);
/* CT: CTDIw (mGy) in a PMMA CTDI phantom (1.19 g/cm3, the given diameter,
 * 15 cm long) scored as dose to PMMA in five 100 mm x 1 cm chamber cylinders,
 * one on the axis and four 1 cm below the surface at 0, 90, 180, 270 degrees:
 * CTDI100 = (mean chamber dose) x 10 cm / collimation, CTDIw = centre / 3 +
 * 2/3 x mean of the four.  `rotation` holds the exposures of ONE axial
 * rotation about the z axis through the origin with the beam's spectra,
 * bowties and weights (n_particles per exposure each); it is simulated on the
 * device at correction level 1 with `seed`.  The factor is ctdi_w /
 * CTDIw(simulated) x (rotation particles per exposure / run particles per
 * exposure).  For a spiral pass ctdi_w = CTDIvol x pitch, for a sequential
 * scan CTDIvol x slice spacing / collimation.  ctdiw_sim (may be null)
 * receives the simulated CTDIw per rotation, in the rotation's dose units. */
int dxg_calibration_ctdi(const dxg_beam* rotation, double ctdi_w_mgy, double diameter_cm, double collimation_cm,
                         uint64_t run_particles_per_exposure, uint64_t seed, double* factor, double* ctdiw_sim);

#ifdef __cplusplus
}
#endif

#endif /* DXMC_BEAMS_H */
