// Types shared by the host world code and the HIP transport kernels.
#pragma once

#include "dxmc_hip.h"

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

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
    unsigned long long uncollided;   // histories that left the grid with no real interaction
    unsigned long long crossings;    // macro-cell boundary crossings (no voxel read)
    // per phase (refill, interact, outer flights, step), DXG_PHASE_CLOCKS builds only:
    unsigned long long phase_cycles[4];   // wavefront clock cycles
    unsigned long long phase_lanes[4];    // lanes doing the phase's work, summed over iterations
    unsigned long long phase_iters[4];    // iterations
};

// Packed voxel word: the FP32 density with its 8 low mantissa bits replaced
// by the material id.  One 4-byte gather per Woodcock step yields both; the
// truncated density is below the true one by at most 2^-15 relative, so the
// majorant built from the true densities still bounds it.
__host__ __device__ inline uint32_t packVoxel(float density, uint32_t material)
{
    union {
        float f;
        uint32_t u;
    } v;
    v.f = density;
    return (v.u & 0xFFFFFF00u) | (material & 0xFFu);
}

// Small per-beam tables, staged into LDS by every block.  Word offsets into
// one uint32 blob:
//   spec_off[n_spec] spec_n[n_spec] bow_off[n_bow] bow_n[n_bow]
//   spec_energy[S] spec_prob[S] spec_alias[S] bow_angle[B] bow_weight[B]
struct BeamBlob {
    int n_spec, n_bow, n_bins, n_nodes;
    int o_spec_off, o_spec_n, o_bow_off, o_bow_n, o_energy, o_prob, o_alias, o_angle, o_weight;
    int words;
};

// The tally of voxel i is the 32-byte record tally[kTallyRecord * i + k]:
// k = 0 sum of w dE, 1 sum of (w dE)^2, 2 number of events (an FP64 integer,
// exact below 2^53), 3 unused.  The three values of one deposit sit in one
// aligned 32-byte block, so one atomic instruction can update all three from
// three lanes as a single memory request.
constexpr int kTallyRecord = 4;

// Klein-Nishina inverse-CDF table (physics.hpp kKnE0 .. kKnN): rows at
// E_r = 1 + 4 r keV, nodes at u_i = i / (kKnNDev - 1), entries {cos(u_i),
// cos(u_{i+1}) - cos(u_i)}.
constexpr int kKnRowsDev = 51;
constexpr int kKnNDev = 128;
constexpr float kKnInvDEDev = 0.25f;   // 1 / (4 keV)

// cos(theta) of a Compton scatter at energy e (keV, 1 <= e <= 201) from a
// uniform u in [0, 1): the inverse Klein-Nishina CDF, interpolated linearly in
// u within the rows either side of e and then linearly in e.  Shared by the
// transport kernels and the host-side table check, so both evaluate the same
// FP32 operations.
__host__ __device__ inline float knCos(const float2* tab, float e, float u)
{
#pragma clang fp contract(off)
    const float fe = fminf((e - 1.0f) * kKnInvDEDev, static_cast<float>(kKnRowsDev - 1));
    const int r0 = static_cast<int>(fe);
    const int r = r0 < kKnRowsDev - 2 ? r0 : kKnRowsDev - 2;
    const float te = fe - static_cast<float>(r);
    const float fu = u * static_cast<float>(kKnNDev - 1);
    const int i0 = static_cast<int>(fu);
    const int i = i0 < kKnNDev - 2 ? i0 : kKnNDev - 2;
    const float tu = fu - static_cast<float>(i);
    const float2 a = tab[r * kKnNDev + i];
    const float2 b = tab[(r + 1) * kKnNDev + i];
    const float ca = fmaf(tu, a.y, a.x);
    const float cb = fmaf(tu, b.y, b.x);
    return fmaf(te, cb - ca, ca);
}

// Klein-Nishina x S(x) tables (physics.hpp kKnsPerOctave, kKnsRows): rows at
// E_r = 2^(r / 12) keV.
constexpr int kKnsRowsDev = 93;
constexpr float kKnsPerOctaveDev = 12.0f;

// cos(theta) of an incoherent scatter (Klein-Nishina x S(x)/Z) on one
// material's table (physics.hpp buildComptonSTable), given the row
// coordinate 12 log2(e) and the node coordinate (1 - sqrt(1 - u)) (kKnNDev -
// 1) for energy e (1 <= e <= 200 keV) and uniform u in [0, 1): interpolated
// as in knCos.  The transport kernels compute the two coordinates with the
// hardware's log2 and square root (about 1 ulp), the host-side table check
// with the correctly rounded ones; the rest is this shared code.
__host__ __device__ inline float knsLookup(const float2* tab, float row, float node)
{
#pragma clang fp contract(off)
    const float fe = fminf(fmaxf(row, 0.0f), static_cast<float>(kKnsRowsDev - 1));
    const int r0 = static_cast<int>(fe);
    const int r = r0 < kKnsRowsDev - 2 ? r0 : kKnsRowsDev - 2;
    const float te = fe - static_cast<float>(r);
    const int i0 = static_cast<int>(node);
    const int i = i0 < kKnNDev - 2 ? i0 : kKnNDev - 2;
    const float tu = node - static_cast<float>(i);
    const float2 a = tab[r * kKnNDev + i];
    const float2 b = tab[(r + 1) * kKnNDev + i];
    const float ca = fmaf(tu, a.y, a.x);
    const float cb = fmaf(tu, b.y, b.x);
    return fmaf(te, cb - ca, ca);
}

// Everything one transport launch reads.  Passed by value as a kernel argument.
struct DevParams {
    // voxel grid of packed words, flat index x + nx (y + ny z)
    const uint32_t* voxels;
    int nx, ny, nz, n_mat;
    int lut_n;   // energy bins of the attenuation and majorant tables staged in LDS (the beam's highest energy, <= kLutN)
    float bmin[3];
    float bmax[3];
    float inv_spacing[3];
    float vox_off[3];   // -bmin * inv_spacing: voxel coordinate = fma(p, inv_spacing, vox_off) on sampling paths
    int level;   // correction level 0 or 1
    // outer medium (global-majorant tracking): every voxel outside the box
    // [outer_lo, outer_hi] holds the packed word outer_word; outer_on is 0 when
    // the box is the whole grid
    int outer_on;
    uint32_t outer_word;
    float outer_lo[3], outer_hi[3];
    int index24;   // nx < 2^24 and ny * nz < 2^24: voxel indices from 24-bit multiplies

    // attenuation tables (see physics.hpp DeviceTables)
    const float* mu_tot;
    const float4* mu_parts;
    const float* majorant;
    const float* ray_fwd;
    const float* ray_inv;
    const float2* kn_inv;   // Klein-Nishina inverse CDF, kKnRowsDev x kKnNDev (knCos)
    const float2* kns_inv;  // per material, Klein-Nishina x S(x) inverse CDF, kKnsRowsDev x kKnNDev each (knsLookup; levels 1, 2)
    // correction level 2 (see physics.hpp DeviceTables)
    const float4* ia_shell;
    const float* ia_inv;
    const float* fl_cum;
    const float* fl_line;

    // macro-cell majorants (null: the global majorant alone); cell c of
    // cell_len cm per axis bounds mu * rho by cell_ratio[c] * majorant_j
    const float* cell_ratio;
    int cnx, cny, cnz;
    float cell_len[3];
    float cell_inv[3];
    // the same bounds as bytes, cell_ratio[c] <= cell_q[c] * cell_qscale,
    // staged in LDS after the beam blob when cell_lds (a read per crossing
    // from LDS instead of a dependent global load)
    const uint8_t* cell_q;
    float cell_qscale;
    int cell_lds;

    // tally: one record of kTallyRecord doubles per voxel, in tally_replicas
    // copies tally_stride doubles apart; work-group b adds to copy b % replicas
    double* tally;
    unsigned long long tally_stride;
    int tally_replicas;
    int tally_merge;      // wavefront deposits on one record merge in LDS before the atomics (small grids)
    float dose_scale;   // beam calibration applied to every deposit (1: none)

    // beam
    const DevExposure* exposures;
    const unsigned long long* ex_start;   // first beam-history of each exposure, n_ex entries
    int n_ex;
    unsigned long long ppe;               // particles per exposure when uniform, else 0
    double inv_ppe;                       // 1 / ppe (exposure index without 64-bit division)
    const uint32_t* blob;                 // BeamBlob words
    BeamBlob bb;

    unsigned long long h_begin;   // this launch covers beam histories [h_begin, h_end)
    unsigned long long h_end;
    unsigned long long seed;
    unsigned long long rng_offset;   // added to the beam-history index that keys the RNG
    unsigned long long* pool;        // history pool counter, zero at launch
    DevCounters* counters;
    int refill_threshold;            // dead lanes that trigger a wavefront refill
    int interact_threshold;          // parked lanes that trigger the interaction phase
    int outer_threshold;             // lanes outside the box that trigger the outer-flight phase
    int lds_lut;                     // attenuation table staged in LDS
    unsigned long long batch;        // histories a wavefront claims per pool atomic (>= 64)
};

// LDS bytes a transport block needs for these params.
size_t transportLdsBytes(const DevParams& p);
void launchTransport(const DevParams& p, int blocks, int threads, hipStream_t s);
void launchTraverse(const DevParams& p, uint64_t n_rays, const float* pos, const float* dir, uint64_t n_steps,
                    const float* steps, int64_t* out_index, int32_t* out_material, hipStream_t s);
void launchReset(unsigned long long* pool, DevCounters* counters, hipStream_t s);
// Adds tally copies 1..replicas-1 into copy 0 (in that order) and zeroes them.
void launchTallyFold(double* tally, unsigned long long stride, int replicas, hipStream_t s);
// Source particles of beam histories [0, n): 8 floats each (ax, ay, direction,
// E, w, exposure index), from the transport kernels' sampling code.
void launchSource(const DevParams& p, uint64_t n, float* out, hipStream_t s);
int transportBlockSize();
void launchCopy(const void* src, void* dst, uint64_t bytes, hipStream_t s);
// Resident blocks per CU of the transport kernel at the given LDS size.
int transportBlocksPerCU(bool lds_lut, int level, bool cells, size_t lds_bytes);
// Many instances in one persistent launch (pool and counters of every
// instance reset first); params and start_inst are device arrays.
void launchTransportBatch(const DevParams* params, int n_inst, const int* start_inst, bool lds_lut, int level, bool cells,
                          size_t lds, int blocks, hipStream_t s);
int transportBatchBlocksPerCU(bool lds_lut, int level, bool cells, size_t lds_bytes);

// Device-resident state of a built world, for kernels outside the transport
// (the dose denoiser reads the tally and the voxel words in place).
struct WorldView {
    const uint32_t* voxels;    // packed words (packVoxel)
    const double* tally;       // one record of kTallyRecord doubles per voxel
    int nx, ny, nz;
    double spacing[3];         // cm
    uint64_t histories;
    int device;
    std::vector<float> embedding;   // per material: total mass attenuation at 60 keV over water's
    const uint8_t* material_host;   // material index per voxel, grid order
    uint64_t generation;            // changes whenever the world is (re)built
};
int worldView(const dxg_world* w, WorldView& out);

// The allocation-free run's scratch array (transport.hip): its size, and its
// address on `device`, which must be the current device (hipGetSymbolAddress
// on the first call per device; a hipError_t).
size_t staticScratchBytes();
int staticScratch(int device, void** out);

// Run preparation shared by dxg_transport_run and batched runs: validates the
// beam, uploads its tables to the world and fills p (all but the history
// range); total receives the beam's history count.
int prepareRun(dxg_world* w, const dxg_beam* beam, uint64_t seed, DevParams& p, uint64_t& total);
// Folds the world's tally copies into one after a run (no-op with one copy);
// enqueued on the world's stream.
int foldTally(dxg_world* w);
// After a run outside dxg_transport_run: account n histories and set the counters.
void addHistories(dxg_world* w, uint64_t n, const dxg_counters& c);
int worldDevice(const dxg_world* w);

}  // namespace dxg
