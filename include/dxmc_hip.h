/* C-ABI of the MI355X voxel photon-transport engine (libdxmc_hip.so).
 *
 * This is the drop-in boundary for DXMClib's voxel dose path as OpenDXMC
 * drives it from SimulationPipeline::worker<C> (simulationpipeline.cpp:124-235).
 * Plain pointers and sizes only; no C++ or torch types cross it.  Every call
 * returns DXG_OK (0) or a negative status and never throws; the message of the
 * last failure on the calling thread is in dxg_last_error().  The C++ facade
 * under include/dxmc/ rebuilds World / AAVoxelGrid / Transport /
 * TransportProgress / DoseScore on top of these entry points.
 *
 * Units follow the reference: lengths in cm (datacontainer.cpp:247-252),
 * densities in g/cm3, energies in keV, angles in radians, dose in mGy
 * (datacontainer.hpp:109).  The voxel grid is centred on the origin
 * (datacontainer.cpp:174-180) and indexed i = x + y*nx + z*nx*ny
 * (otherphantomimportpipeline.cpp:44).
 */
#ifndef DXMC_HIP_H
#define DXMC_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    DXG_OK = 0,
    DXG_ERR_ARGUMENT = -1,     /* null pointer, bad size, bad enum */
    DXG_ERR_STATE = -2,        /* call out of order (e.g. run before build) */
    DXG_ERR_MATERIAL = -3,     /* Material::byWeight failure */
    DXG_ERR_DATA = -4,         /* material index >= number of materials */
    DXG_ERR_PHYSICS = -5,      /* physics list missing or corrupt */
    DXG_ERR_DEVICE = -6,       /* HIP runtime failure / no device */
    DXG_ERR_UNSUPPORTED = -7,  /* feature outside this build (e.g. correction level 3) */
    DXG_ERR_CANCELLED = -8     /* progress stop flag was set during the run */
};

typedef struct dxg_world dxg_world;
typedef struct dxg_progress dxg_progress;

/* Material<5>::byWeight(std::map<Z, mass fraction>)  (simulationpipeline.cpp:136).
 * Fractions are normalised to one. */
typedef struct dxg_material {
    const int32_t* Z;
    const double* mass_fraction;
    uint64_t n;
} dxg_material;

/* One exposure of a beam: Beam::exposure(i) with position(), directionCosines()
 * and collimationHalfAngles() (beamactorcontainer.cpp:114-150).  The beam axis
 * is cross(cosines[0..2], cosines[3..5]) (beamactorcontainer.cpp:52).  A
 * particle leaves `position` in the direction
 *   c0*sin(ax) + c1*sin(ay) + axis*sqrt(1 - sin^2 ax - sin^2 ay),
 * with ax, ay uniform in +-collimation_half_angles (beamactorcontainer.cpp:44-69);
 * its weight is `weight` times the bowtie weight at ax. */
typedef struct dxg_exposure {
    double position[3];
    double cosines[6];
    double collimation_half_angles[2];
    double weight;
    uint32_t spectrum_index;  /* into dxg_beam.spectra */
    int32_t bowtie_index;     /* into dxg_beam.bowties, or -1 */
    uint64_t n_particles;     /* numberOfParticlesPerExposure() */
} dxg_exposure;

/* Discrete photon spectrum: a particle takes energy[k] with probability
 * weight[k] / sum(weight).  A monoenergetic beam has n = 1. */
typedef struct dxg_spectrum {
    const double* energy_kev;
    const double* weight;
    uint64_t n;
} dxg_spectrum;

/* BowtieFilter(vector<pair<angle, weight>>) (bowtiefilterreader.cpp:92).
 * Linear interpolation in |angle|, clamped at the ends, normalised so the
 * central ray has weight one. */
typedef struct dxg_bowtie {
    const double* angle_rad;
    const double* weight;
    uint64_t n;
} dxg_bowtie;

/* Everything Transport::operator()(world, beam, ...) needs from a Beam. */
typedef struct dxg_beam {
    const dxg_exposure* exposures;
    uint64_t n_exposures;
    const dxg_spectrum* spectra;
    uint64_t n_spectra;
    const dxg_bowtie* bowties;
    uint64_t n_bowties;
    /* Beam calibration: mGy per (dose summed over the run's histories), from
     * dxg_calibration_* (include/dxmc_beams.h); 0 means uncalibrated (1). */
    double calibration;
} dxg_beam;

/* Device-side event counters of the last dxg_transport_run. */
typedef struct dxg_counters {
    uint64_t histories;
    uint64_t steps;            /* Woodcock steps, real + virtual */
    uint64_t real_events;      /* real interactions */
    uint64_t deposits;         /* scored deposits (tally updates) */
    double emitted_energy;     /* sum of w*E over emitted photons (keV) */
    double escaped_energy;     /* sum of w*E leaving the grid (keV) */
    double deposited_energy;   /* sum of scored w*dE (keV) */
    double kernel_ms;          /* summed duration of the transport kernels */
    uint64_t kernel_launches;
    uint64_t uncollided;       /* histories leaving the grid with no real interaction */
    uint64_t cell_crossings;   /* macro-cell boundary crossings (majorant changes, no voxel read) */
    /* per phase (refill, interact, outer-medium flights, step): wavefront
     * clock cycles, lanes doing the phase's work summed over the phase's
     * iterations, and iterations (libraries built with DXG_PHASE_CLOCKS;
     * zero otherwise) */
    uint64_t phase_cycles[4];
    uint64_t phase_lanes[4];
    uint64_t phase_iters[4];
    /* the launch settings the run ended with: dead / parked / outside lanes
     * that start a refill / interaction / outer-flight phase, histories a
     * wavefront claims per pool atomic, and 1 when they were chosen by the
     * run's own timing (dxg_transport_run tunes them on the first run of a
     * beam layout on a built world and keeps them; DXG_AUTOTUNE=0 or any
     * explicit DXG_REFILL / DXG_INTERACT / DXG_OUTERT / DXG_BATCH turns that off) */
    int32_t refill_threshold;
    int32_t interact_threshold;
    int32_t outer_threshold;
    int32_t autotuned;
    uint64_t batch;
} dxg_counters;

const char* dxg_last_error(void);

/* ---- World<AAVoxelGrid<5, CORRECTION, 255>> (simulationpipeline.cpp:127-131).
 * correction_level is the CORRECTION template argument: 0 = free electrons,
 * 1 = form factor / incoherent scattering function (the reference default,
 * simulationpipeline.hpp:62), 2 = impulse approximation with shell Doppler
 * broadening and K-shell fluorescence; other values give DXG_ERR_ARGUMENT. */
int dxg_world_create(dxg_world** out, int correction_level);
void dxg_world_destroy(dxg_world* w);

/* AAVoxelGrid::setData(dims, density, material, materials) (simulationpipeline.cpp:149).
 * Copies the arrays.  Fails with DXG_ERR_DATA when a material index is not
 * below n_mats (the readiness check of simulationpipeline.cpp:50-58) and with
 * DXG_ERR_MATERIAL when a material cannot be built (:136-140). */
int dxg_grid_set_data(dxg_world* w, const uint64_t dims[3], const double* density,
                      const uint8_t* material, const dxg_material* mats, uint64_t n_mats);

/* AAVoxelGrid::setSpacing(std::array<double,3> cm) (simulationpipeline.cpp:150). */
int dxg_grid_set_spacing(dxg_world* w, const double spacing_cm[3]);

/* World::build() (simulationpipeline.cpp:153): attenuation tables, Woodcock
 * majorant and sampling tables are built and the grid is made resident in
 * device memory.  Clears the scores. */
int dxg_world_build(dxg_world* w);

/* AAVoxelGrid::size() (simulationpipeline.cpp:174). */
uint64_t dxg_grid_size(const dxg_world* w);

/* Transport::operator()(world, beam, progress, useBeamCalibration)
 * (simulationpipeline.cpp:163-166).  Blocking.  Scores accumulate over calls.
 * `seed` keys the per-history random streams; histories of successive calls
 * on one world are numbered consecutively so their streams never overlap.
 * `progress` may be null.  When another thread calls
 * dxg_progress_request_stop() the run stops taking new histories and returns
 * DXG_ERR_CANCELLED (the reference returns without collecting dose,
 * simulationpipeline.cpp:169-170).  With apply_calibration set (the
 * reference's useBeamCalibration, simulationpipeline.cpp:165) every deposit of
 * the run is multiplied by beam->calibration (when > 0), so each beam's
 * contribution to the world's scores is in mGy for its stated output; the
 * energy counters stay uncalibrated. */
int dxg_transport_run(dxg_world* w, const dxg_beam* beam, dxg_progress* progress,
                      int apply_calibration, uint64_t seed);

/* ---- dxg_transport_run without allocation.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops).  Same arguments, scores, counters and
 * results as dxg_transport_run; it allocates and frees no device memory and
 * creates no HIP events (counters.kernel_ms is then the host's clock around
 * each launch and its wait).  The history pool, the event counters and the
 * beam's tables live in a fixed 1 MiB scratch array per device, which the
 * library holds until the process ends: the beam's exposure records (80 B
 * each), their first histories (8 B each) and its spectrum/bowtie tables must
 * fit (about 12,000 exposures with 1 keV spectra), else DXG_ERR_ARGUMENT.
 * Every call on one GPU shares that GPU's scratch, so calls that overlap on
 * one GPU (from several threads) race: serialise them. */
int dxg_transport_run_static(dxg_world* w, const dxg_beam* beam, dxg_progress* progress,
                             int apply_calibration, uint64_t seed);

/* Zero the energy scores and the history count. */
int dxg_world_clear_scores(dxg_world* w);

/* EnergyScore per voxel: sum of w*dE (keV), sum of (w*dE)^2 (keV^2), events. */
int dxg_read_energy_scores(const dxg_world* w, double* sum_e, double* sum_e2, uint64_t* n_events);

/* DoseScore per voxel: vgrid.doseScored(i).{dose(), variance(), numberOfEvents()}
 * (simulationpipeline.cpp:178,204,219).  dose = sum_e / (rho V) in mGy;
 * variance (mGy^2) of that sum over the H histories run so far,
 * H/(H-1) * (sum_e2 - sum_e^2 / H), converted like the dose.  Any output
 * pointer may be null. */
int dxg_read_dose_scores(const dxg_world* w, double* dose_mgy, double* variance_mgy2, uint64_t* n_events);

/* Histories run into the scores so far. */
uint64_t dxg_world_histories(const dxg_world* w);

int dxg_get_counters(const dxg_world* w, dxg_counters* out);

/* ---- TransportProgress (simulationpipeline.hpp:23,65; simulationpipeline.cpp:112-122,259-262).
 * All progress calls are thread-safe and may be made from any thread. */
int dxg_progress_create(dxg_progress** out);
void dxg_progress_destroy(dxg_progress* p);
void dxg_progress_request_stop(dxg_progress* p);        /* setStopSimulation() */
int dxg_progress_continue(const dxg_progress* p);       /* continueSimulation() */
void dxg_progress_get(const dxg_progress* p, uint64_t* n, uint64_t* total);  /* progress() */
/* message(): copies at most cap-1 bytes plus a terminating zero. */
void dxg_progress_message(const dxg_progress* p, char* buf, uint64_t cap);

/* ---- Material::attenuationValues(E) (ctsegmentationpipeline.cpp:73-98):
 * out[4*i + {0,1,2,3}] = photoelectric, incoherent, coherent, total (cm2/g). */
int dxg_material_attenuation(const dxg_material* m, const double* energy_kev, uint64_t n, double* out);

/* ---- Deterministic traversal through the transport kernels' own device
 * functions (AAVoxelGrid lookup, simulationpipeline.cpp:127; index order
 * x + y*nx + z*nx*ny, otherphantomimportpipeline.cpp:44): the refill path's
 * slab entry of each ray into the grid box (IEEE division), then per step
 * p = fma(steps[k], d, p) and the Woodcock step's voxel lookup
 * floor(fma(p, 1/spacing, -box_min/spacing)) per axis, all FP32.  out_index
 * is -1 outside the grid, out_material the material id or -1.  Needs a built
 * world; pos/dir are n_rays x 3 float, steps n_rays x n_steps float. */
int dxg_debug_traverse(const dxg_world* w, uint64_t n_rays, const float* pos, const float* dir,
                       uint64_t n_steps, const float* steps, int64_t* out_index, int32_t* out_material);

/* ---- Source particles through the transport kernels' own sampling code
 * (Exposure::sampleParticle, the refill path; geometry of
 * beamactorcontainer.cpp:44-69): beam histories 0..n-1 of `beam`, each keyed
 * by (seed, history index) as dxg_transport_run keys the first run on a fresh
 * world.  out[8*h + {0..7}] = ax, ay (collimation angles, rad), direction
 * (3), energy (keV), weight (exposure weight x bowtie weight at |ax|) and the
 * exposure index.  Needs a built world (its device); n at most the beam's
 * number of particles.  Scores are not touched. */
int dxg_debug_source(dxg_world* w, const dxg_beam* beam, uint64_t seed, uint64_t n, float* out);

/* ---- Compton scattering angle as the transport kernels sample it
 * (DXMClib's Klein-Nishina sampling in interactions<5, C>, absent from the
 * reference tree; caller simulationpipeline.cpp:159-170): cos(theta) at
 * energy_kev (1-200 keV) for each uniform u[i] in [0, 1), by the kernels' own
 * inverse-CDF table and FP32 interpolation, evaluated on the host (no GPU
 * needed).  The level-1 S(x) rejection is not applied. */
int dxg_debug_compton_cos(double energy_kev, uint64_t n, const float* u, float* cos_out);

/* The same for correction levels 1 and 2: cos(theta) of incoherent
 * scattering (Klein-Nishina x S(x)/Z) in material m, by the kernels' own
 * per-material table and FP32 interpolation, evaluated on the host. */
int dxg_debug_compton_cos_material(const dxg_material* m, double energy_kev, uint64_t n, const float* u, float* cos_out);

/* ---- Profiling aid (no counterpart in the reference): `iters` launches of a
 * plain float4 copy kernel over `bytes` bytes (multiple of 16) between two
 * fresh device buffers.  Under a hardware-counter profiler its FETCH_SIZE and
 * WRITE_SIZE per launch calibrate the counters against a known traffic of
 * `bytes` read and `bytes` written.  *ms receives the mean launch time. */
int dxg_debug_device_copy(uint64_t bytes, int iters, double* ms);

/* ---- Batched transport (no counterpart in the reference, whose Transport
 * runs one world and beam per call, simulationpipeline.cpp:159-170): many
 * (world, beam) instances of uneven size in ONE persistent launch.  Blocks
 * start on instances in proportion to their histories and move to instances
 * with histories left when theirs is exhausted.  The launch is captured into
 * a HIP graph on the first run and replayed by later runs of the same batch.
 * Each instance behaves exactly like dxg_transport_run(world, beam, NULL, 1,
 * seed) (calibration included): same RNG streams (keyed by seed and the world's history count), same
 * scores and counters, read through the world afterwards.  Worlds must be
 * built, live on one device, share one correction level, hold one instance of
 * the batch and outlive it.  No progress or cancellation inside a run. */
typedef struct dxg_batch dxg_batch;
int dxg_batch_create(dxg_batch** out);
void dxg_batch_destroy(dxg_batch* b);
/* The beam is copied. */
int dxg_batch_add(dxg_batch* b, dxg_world* w, const dxg_beam* beam, uint64_t seed);
uint64_t dxg_batch_size(const dxg_batch* b);
/* One launch over every instance; *kernel_ms (may be null) receives its device time. */
int dxg_batch_run(dxg_batch* b, double* kernel_ms);

#ifdef __cplusplus
}
#endif

#endif /* DXMC_HIP_H */
