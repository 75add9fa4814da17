// Types shared by the host world code and the HIP transport kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace dxg {

// One exposure in device form (FP32).  axis = c0 x c1 is precomputed.
struct DevExposure {
    float pos[3];
    float c0[3];
    float c1[3];
    float axis[3];
    float half_angle[2];
    float weight;
    int32_t spectrum;   // index into the spectrum table
    int32_t bowtie;     // index into the bowtie table or -1
    int32_t pad[3];
};
static_assert(sizeof(DevExposure) == 80, "exposure record layout");

struct DevCounters {
    unsigned long long steps;
    unsigned long long real_events;
    unsigned long long deposits;
    unsigned long long histories;
    double emitted;
    double escaped;
    double deposited;
    double pad;
};

// Everything one transport launch reads.  Passed by value as a kernel argument.
struct DevParams {
    // voxel grid, flat index x + nx (y + ny z)
    const uint8_t* material;
    const float* density;
    int nx, ny, nz, n_mat;
    float bmin[3];
    float bmax[3];
    float inv_spacing[3];
    int level;   // correction level 0 or 1

    // attenuation tables (see physics.hpp DeviceTables)
    const float* mu_tot;
    const float4* mu_parts;
    const float* majorant;
    const float* s_table;
    const float* ray_fwd;
    const float* ray_inv;

    // tallies: per voxel {sum w dE, sum (w dE)^2, events (u64 bits), unused}
    double* tally;

    // beam
    const DevExposure* exposures;
    const unsigned long long* ex_start;   // first beam-history of each exposure, n_ex entries
    int n_ex;
    const float* spec_energy;   // concatenated spectra
    const float* spec_prob;     // alias acceptance probability
    const int* spec_alias;      // alias index (relative to the spectrum)
    const int* spec_off;        // per spectrum: first bin
    const int* spec_n;          // per spectrum: number of bins
    const float* bow_angle;     // concatenated bowtie nodes (ascending |angle|)
    const float* bow_weight;
    const int* bow_off;
    const int* bow_n;

    unsigned long long h_begin;   // this launch covers beam histories [h_begin, h_end)
    unsigned long long h_end;
    unsigned long long seed;
    unsigned long long rng_offset;   // added to the beam-history index that keys the RNG
    unsigned long long* pool;        // history pool counter, zero at launch
    DevCounters* counters;
};

void launchTransport(const DevParams& p, int blocks, int threads, hipStream_t s);
void launchTraverse(const DevParams& p, uint64_t n_rays, const float* pos, const float* dir, uint64_t n_steps,
                    const float* steps, int64_t* out_index, int32_t* out_material, hipStream_t s);
void launchReset(unsigned long long* pool, DevCounters* counters, hipStream_t s);
int transportBlockSize();
const void* transportKernelSymbol();

}  // namespace dxg
