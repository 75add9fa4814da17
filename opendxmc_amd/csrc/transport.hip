// HIP/CDNA4 kernels of the voxel photon-transport engine.
//
// transport_kernel is a persistent history-pool kernel.  Every lane carries
// one photon in registers (FP32 state) and is in one of these states:
//   DEAD      needs a new history,
//   TRACK     Woodcock tracking,
//   INTERACT  parked at a real collision, waiting for its interaction,
//   COMPTON / RAYLEIGH  parked with that process chosen, its rejection
//             sampling not yet accepted,
//   OUTER     waiting for an exact flight through the outer medium,
//   DONE      the history pool is exhausted.
// Each loop iteration a wavefront runs exactly ONE phase, chosen from ballots
// of the lane states, with the lanes of the other states masked off:
//   * refill   once refill_threshold lanes are DEAD (or nothing else can run):
//              they take consecutive ids from the wavefront's private batch of
//              pool histories (one global atomic per `batch` histories) and
//              sample their source particle together;
//   * interact once interact_threshold lanes are parked (or no lane tracks
//              and fewer than outer_threshold wait for outer flights):
//              photoelectric / Compton / Rayleigh for all parked lanes at
//              once, one rejection-sampling trial per lane and phase;
//   * outer    once outer_threshold lanes wait outside the box of the
//              non-outer voxels (or no lane tracks): exact flights through the
//              uniform outer medium, to a collision, the box or the grid's edge;
//   * step     otherwise: every TRACK lane advances one Woodcock step -- one
//              4-byte gather of the packed voxel word (density + material id),
//              one LDS read of the attenuation at the photon's energy bin.
// The expensive branches therefore run with many lanes active instead of
// stalling a full wavefront for a few lanes.  The per-history random streams
// are consumed in the same order in every phase schedule, so results do not
// depend on thresholds, launch geometry or lane placement.
//
// Deposits go straight to the FP64 per-voxel tally with hardware float
// atomics; event counters are summed across the wavefront before one atomic
// per wavefront.  LDS holds, per block: the Woodcock majorant table, the
// per-material total attenuation table (when it fits), and the beam's
// spectrum alias tables and bowtie nodes.
//
// Algorithms (DXMClib's AAVoxelGrid::transport and interactions<5, C>, absent
// from the reference tree; the caller is simulationpipeline.cpp:159-170):
//   Woodcock tracking with a per-energy-bin majorant; photoelectric
//   absorption scored as KERMA; Klein-Nishina (level 0) or Klein-Nishina x
//   S(x)/Z of the material (levels 1, 2) by its inverse CDF, tabulated in
//   FP64 on 128 nodes per row (rows every 4 keV, or 12 per octave) and
//   interpolated bilinearly; Thomson by the exact inverse of its cubic CDF
//   (level 1: x^2 drawn from F^2(x), accepted with (1+cos^2)/2).
#include "device.hpp"

#include <algorithm>
#include <mutex>

namespace dxg {

namespace {

constexpr float kLutE0f = 1.0f;
constexpr float kLutInvDEf = 4.0f;    // 1 / 0.25 keV
constexpr int kLutNd = 797;
constexpr float kXMaxf = 16.0f;
constexpr int kSfNd = 1024;
constexpr int kRayInvNd = 1024;
constexpr float kInvHc = 1.0f / 12.398419843320026f;
constexpr float kInvMeC2 = 1.0f / 510.99895f;
constexpr float kCutoff = 1.0f;
// Work-group size and the occupancy target of the transport kernels
// (compile-time).  Six wavefronts per SIMD is what the LDS tables allow
// (about 26 KB per work-group of 256); the target keeps the macro-cell
// instances at 80 VGPRs (81-82 otherwise: five wavefronts).
#ifndef DXG_TRANSPORT_BLOCK
#define DXG_TRANSPORT_BLOCK 256
#endif
#ifndef DXG_TRANSPORT_WAVES_PER_EU
#define DXG_TRANSPORT_WAVES_PER_EU 6
#endif
constexpr int kBlock = DXG_TRANSPORT_BLOCK;
// Resident work-groups per CU at most: the occupancy target.  Instances
// that fit in 72 VGPRs and small LDS tables would allow seven; seven
// measured slower (C2 10.67 against 10.50 ms per 1e8 histories, the 64
// batched instances 28.5 against 26.3 ms).
constexpr int kMaxBlocksPerCU = DXG_TRANSPORT_WAVES_PER_EU * 4 * 64 / DXG_TRANSPORT_BLOCK;   // 4 SIMDs per CU
static_assert(kMaxBlocksPerCU >= 1, "work-groups per CU");
#define DXG_TRANSPORT_OCCUPANCY __attribute__((amdgpu_waves_per_eu(DXG_TRANSPORT_WAVES_PER_EU)))
constexpr int kMaxCross = 8;    // macro-cell crossings a lane walks per iteration
constexpr int kIaShellsd = 5;    // impulse-approximation shells per material
constexpr int kIaInvNd = 2048;   // unit Compton profile inverse-CDF nodes
constexpr int kFlMaxd = 12;      // fluorescing elements per material

// kCompton / kRayleigh: parked with that process chosen, its rejection
// sampling still to accept (parked like kInteract)
enum LaneState : int { kDead = 0, kTrack = 1, kInteract = 2, kDone = 3, kOuter = 4, kCompton = 5, kRayleigh = 6 };

// The hardware's approximations (v_log_f32, v_exp_f32, v_rcp_f32, v_sqrt_f32;
// about 1 ulp) for sampling.  The library calls expand to ~10-20 instructions
// each for denormal handling and correct rounding, which sampling does not
// need: a flight length -ln(u)/mu cost 22 instructions of a ~70-instruction
// Woodcock step.  Arguments are normal, positive floats (u >= 2^-23).
constexpr float kLn2f = 0.693147181f;
constexpr float kLog2ef = 1.442695041f;
__device__ inline float hwLog(float x) { return __builtin_amdgcn_logf(x) * kLn2f; }
__device__ inline float hwExp(float x) { return __builtin_amdgcn_exp2f(x * kLog2ef); }
__device__ inline float hwRcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ inline float hwSqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

struct Rng {
    uint32_t s0, s1, s2, s3;
};

__device__ inline uint64_t splitmix(uint64_t& x)
{
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__device__ inline void rngSeed(Rng& r, uint64_t seed, uint64_t history)
{
    uint64_t x = seed ^ (history * 0xD1B54A32D192ED03ull);
    const uint64_t a = splitmix(x), b = splitmix(x);
    r.s0 = static_cast<uint32_t>(a);
    r.s1 = static_cast<uint32_t>(a >> 32);
    r.s2 = static_cast<uint32_t>(b);
    r.s3 = static_cast<uint32_t>(b >> 32);
    if ((r.s0 | r.s1 | r.s2 | r.s3) == 0)
        r.s0 = 1;
}

// xoshiro128+; uniform in (0, 1] on a grid of 2^-23: the top 23 bits become
// the mantissa of a float in [1, 2) (shift, or) and 2 - f lands in (0, 1]
// (one subtract) -- three instructions where a convert-and-scale takes four.
__device__ inline float rngU(Rng& r)
{
    const uint32_t out = r.s0 + r.s3;
    const uint32_t t = r.s1 << 9;
    r.s2 ^= r.s0;
    r.s3 ^= r.s1;
    r.s1 ^= r.s2;
    r.s0 ^= r.s3;
    r.s2 ^= t;
    r.s3 = (r.s3 << 11) | (r.s3 >> 21);
    return 2.0f - __uint_as_float((out >> 9) | 0x3F800000u);
}

// Slab entry of a ray into the grid box, with IEEE division (three per
// history, at its start): the entry point is reproducible bit for bit on any
// IEEE FP32 machine with a fused multiply-add.
__device__ inline bool boxEntry(const DevParams& P, float& px, float& py, float& pz, float dx, float dy, float dz)
{
    float t0 = 0.0f, t1 = 3.0e38f;
    {
        const float id = 1.0f / dx;
        const float ta = (P.bmin[0] - px) * id, tb = (P.bmax[0] - px) * id;
        t0 = fmaxf(t0, fminf(ta, tb));
        t1 = fminf(t1, fmaxf(ta, tb));
    }
    {
        const float id = 1.0f / dy;
        const float ta = (P.bmin[1] - py) * id, tb = (P.bmax[1] - py) * id;
        t0 = fmaxf(t0, fminf(ta, tb));
        t1 = fminf(t1, fmaxf(ta, tb));
    }
    {
        const float id = 1.0f / dz;
        const float ta = (P.bmin[2] - pz) * id, tb = (P.bmax[2] - pz) * id;
        t0 = fmaxf(t0, fminf(ta, tb));
        t1 = fminf(t1, fmaxf(ta, tb));
    }
    if (!(t0 <= t1))
        return false;
    px = __fmaf_rn(t0, dx, px);
    py = __fmaf_rn(t0, dy, py);
    pz = __fmaf_rn(t0, dz, pz);
    return true;
}

// Moves a point by s along d: one fused multiply-add per axis.
__device__ inline void advance(float& px, float& py, float& pz, float s, float dx, float dy, float dz)
{
    px = __fmaf_rn(s, dx, px);
    py = __fmaf_rn(s, dy, py);
    pz = __fmaf_rn(s, dz, pz);
}

// floor(x) converted to int by the hardware's floor-convert (one instruction
// for floor + cvt; saturates outside the int range).
__device__ inline int floorToInt(float x)
{
    int r;
    asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(r) : "v"(x));
    return r;
}

// Flat voxel index (x + y nx + z nx ny, the AAVoxelGrid order) of a point, or
// -1 outside the grid.  The voxel coordinate per axis is
//   floor(fma(p, inv_spacing, vox_off)),  vox_off = -box_min * inv_spacing,
// rounded once in FP32: a point within an ulp of a voxel face belongs to the
// voxel this rounding picks, the same on every IEEE FP32 machine with a fused
// multiply-add.  The only voxel lookup of the transport kernels, and the one
// the traversal check runs.
__device__ inline int voxelIndex(const DevParams& P, float px, float py, float pz)
{
    const int ix = floorToInt(__fmaf_rn(px, P.inv_spacing[0], P.vox_off[0]));
    const int iy = floorToInt(__fmaf_rn(py, P.inv_spacing[1], P.vox_off[1]));
    const int iz = floorToInt(__fmaf_rn(pz, P.inv_spacing[2], P.vox_off[2]));
    if (static_cast<unsigned>(ix) >= static_cast<unsigned>(P.nx) || static_cast<unsigned>(iy) >= static_cast<unsigned>(P.ny)
        || static_cast<unsigned>(iz) >= static_cast<unsigned>(P.nz))
        return -1;
    if (P.index24)   // full-rate v_mad_u32_u24 (the 32-bit multiply is quarter rate); products stay below 2^31
        return static_cast<int>(__umul24(static_cast<unsigned>(P.nx), static_cast<unsigned>(iy) + __umul24(static_cast<unsigned>(P.ny), static_cast<unsigned>(iz))))
               + ix;
    return ix + P.nx * (iy + P.ny * iz);
}

// Material id and FP32 density of a packed voxel word (density with its 8 low
// mantissa bits replaced by the material id).
__device__ inline int wordMaterial(uint32_t word) { return static_cast<int>(word & 0xFFu); }
__device__ inline float wordDensity(uint32_t word) { return __uint_as_float(word & 0xFFFFFF00u); }

// Whether a point lies in the closed box [lo, hi].
__device__ inline bool inBox(const float* lo, const float* hi, float px, float py, float pz)
{
    return px >= lo[0] && px <= hi[0] && py >= lo[1] && py <= hi[1] && pz >= lo[2] && pz <= hi[2];
}

// Narrows [t0, t1] to the part of the ray p + t d inside the slab [lo, hi] of
// one axis (id = 1 / d).
__device__ inline void clipSlab(float p, float d, float id, float lo, float hi, float& t0, float& t1)
{
    if (d != 0.0f) {
        const float ta = (lo - p) * id, tb = (hi - p) * id;
        t0 = fmaxf(t0, fminf(ta, tb));
        t1 = fminf(t1, fmaxf(ta, tb));
    } else if (p < lo || p > hi) {
        t0 = __builtin_inff();
    }
}

// Sine and cosine of 2 pi u (u in revolutions, |u| <= 256): the hardware's
// v_sin_f32 / v_cos_f32 take revolutions, so no scaling by 2 pi and back.
__device__ inline void sinCosRev(float u, float& s, float& c)
{
    s = __builtin_amdgcn_sinf(u);
    c = __builtin_amdgcn_cosf(u);
}

// Turns the direction by the polar cosine cost and the azimuth 2 pi u.
__device__ inline void rotate(float& dx, float& dy, float& dz, float cost, float u)
{
    const float sint = hwSqrt(fmaxf(0.0f, 1.0f - cost * cost));
    float sp, cp;
    sinCosRev(u, sp, cp);
    float nx, ny, nz;
    if (fabsf(dz) > 0.99999f) {
        nx = sint * cp;
        ny = sint * sp;
        nz = dz > 0.0f ? cost : -cost;
    } else {
        const float t = hwSqrt(1.0f - dz * dz);
        const float it = hwRcp(t);
        nx = dx * cost + sint * (dx * dz * cp - dy * sp) * it;
        ny = dy * cost + sint * (dy * dz * cp + dx * sp) * it;
        nz = dz * cost - t * sint * cp;
    }
    const float inv = rsqrtf(nx * nx + ny * ny + nz * nz);
    dx = nx * inv;
    dy = ny * inv;
    dz = nz * inv;
}

// LDS bytes per wavefront for the merge of tallyFlush: 64 hash slots (the
// lane that owns each) and 64 lanes' record updates {w dE, (w dE)^2, events}.
constexpr int kMergeWaveBytes = 64 * 4 + 64 * 3 * 8;

// Scores the deposits of one interaction phase: each lane with deps > 0 made
// deps (1 or 2) deposits e1 [, e2] at its voxel vi, and its record gains
// (sum w dE, sum (w dE)^2, events) = (d1 + d2, d1^2 + d2^2, deps) with
// d = e * dose_scale.  Wavefront-collective: every lane of the wavefront calls
// it.  FP64 atomics are executed beyond L2 at a fixed rate of requests, and
// one atomic instruction's lanes that fall in one 64-byte block travel as one
// request; so a record update is spread over three consecutive lanes of one
// instruction (one value each, all in the record's aligned 32 bytes) -- one
// request per update where three separate atomics made three.  With s_merge
// (this wavefront's kMergeWaveBytes of LDS) updates of one record merge first.

__device__ inline void tallyFlush(const DevParams& P, double* tally, int deps, int vi, float e1, float e2, unsigned char* s_merge)
{
    const unsigned long long m0 = __ballot(deps > 0);
    if (m0 == 0)
        return;
    const int lane = static_cast<int>(__lane_id());
    const double scale = static_cast<double>(P.dose_scale);
    if (s_merge) {
        // Merge first (small grids, where many lanes deposit on few records:
        // FP64 atomics on one record are applied one after another).  Each
        // depositing lane claims the hash slot of its voxel; a lane whose slot
        // went to another lane with the same voxel adds its update to that
        // lane's in LDS and drops out; a lane that lost its slot to another
        // voxel flushes its own.  One wavefront's LDS operations execute in
        // program order, so the updates are in place before the adds and the
        // adds before the reads.
        int* own = reinterpret_cast<int*>(s_merge);
        double* acc = reinterpret_cast<double*>(s_merge + 64 * 4);
        const bool has = deps > 0;
        const int hslot = static_cast<int>((static_cast<unsigned>(vi) * 2654435761u) >> 26);
        if (has)
            __hip_atomic_store(own + hslot, lane, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
        const int w = has ? __hip_atomic_load(own + hslot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT) : lane;
        const int vw = __builtin_amdgcn_ds_bpermute(w * 4, vi);
        const bool dup = has && w != lane && vw == vi;
        const double d1 = static_cast<double>(e1) * scale, d2 = static_cast<double>(e2) * scale;
        const double u0 = d1 + d2, u1 = d1 * d1 + d2 * d2, u2 = static_cast<double>(deps);
        if (has && !dup) {
            acc[3 * lane] = u0;
            acc[3 * lane + 1] = u1;
            acc[3 * lane + 2] = u2;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        if (dup) {
            atomicAdd(acc + 3 * w, u0);
            atomicAdd(acc + 3 * w + 1, u1);
            atomicAdd(acc + 3 * w + 2, u2);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        const unsigned long long m = __ballot(has && !dup);
        const int k = __popcll(m);
        const int rank = __popcll(m & ((1ull << lane) - 1ull));
        const int dest = (has && !dup) ? rank : k + (lane - rank);
        const int src_of = __builtin_amdgcn_ds_permute(dest * 4, lane);
        for (int base = 0; base < 3 * k; base += 64) {
            const int slot = base + lane;
            const int q = slot / 3;
            const int c = slot - 3 * q;
            const int src = __builtin_amdgcn_ds_bpermute((q & 63) * 4, src_of);
            const int idx = __builtin_amdgcn_ds_bpermute(src * 4, vi);
            if (slot < 3 * k)
                atomicAdd(tally + static_cast<size_t>(idx) * kTallyRecord + c, acc[3 * src + c]);
        }
        return;
    }
    const unsigned long long m = m0;
    const int k = __popcll(m);
    const int rank = __popcll(m & ((1ull << lane) - 1ull));
    // lane q < k learns the lane of update q: updating lanes go to their rank,
    // the others to k and up, a permutation, so every destination is written once
    const int dest = deps > 0 ? rank : k + (lane - rank);
    const int src_of = __builtin_amdgcn_ds_permute(dest * 4, lane);
    for (int base = 0; base < 3 * k; base += 64) {   // wavefront-uniform rounds of 64 slots
        const int slot = base + lane;                  // slot = 3 * update + value
        const int q = slot / 3;
        const int c = slot - 3 * q;
        const int src = __builtin_amdgcn_ds_bpermute((q & 63) * 4, src_of);
        const int idx = __builtin_amdgcn_ds_bpermute(src * 4, vi);
        const int n = __builtin_amdgcn_ds_bpermute(src * 4, deps);
        const double d1 = static_cast<double>(__int_as_float(__builtin_amdgcn_ds_bpermute(src * 4, __float_as_int(e1)))) * scale;
        const double d2 = static_cast<double>(__int_as_float(__builtin_amdgcn_ds_bpermute(src * 4, __float_as_int(e2)))) * scale;
        if (slot < 3 * k) {
            const double v = c == 0 ? d1 + d2 : (c == 1 ? d1 * d1 + d2 * d2 : static_cast<double>(n));
            atomicAdd(tally + static_cast<size_t>(idx) * kTallyRecord + c, v);
        }
    }
}

__device__ inline float lerpTab(const float* tab, int j, float t)
{
    const float a = tab[j];
    const float b = tab[j + 1];
    return __fmaf_rn(t, b - a, a);
}

__device__ inline float asFloat(uint32_t u) { return __uint_as_float(u); }

// Macro-cell coordinate of position p on one axis; a photon on a cell face
// moving backwards belongs to the cell behind the face.
__device__ inline int cellCoord(float p, float bmin, float inv, int n, float d)
{
    const float f = (p - bmin) * inv;
    const float fl = floorf(f);
    int c = static_cast<int>(fl);
    if (d < 0.0f && f == fl)
        c -= 1;
    return min(max(c, 0), n - 1);
}

// Path length to the exit face of cell c along one axis (+inf when the
// direction is parallel to the faces; never negative); inv_d = 1 / d.
__device__ inline float exitDistance(float p, float d, float inv_d, float bmin, float len, int c)
{
    if (d == 0.0f)
        return __builtin_inff();
    const float face = bmin + static_cast<float>(c + (d > 0.0f ? 1 : 0)) * len;
    return fmaxf((face - p) * inv_d, 0.0f);
}

// Energy bin of the attenuation tables and the bin's majorant.  Macro-cell
// tracking scales the majorant per cell, so CELLS keeps it as is; the global
// tracker keeps its reciprocal (the mean free path between tentative
// collisions), so a step multiplies instead of taking a quarter-rate
// reciprocal: the reciprocal is taken once per energy change.
template <bool CELLS>
__device__ inline void energyBin(float E, const float* s_maj, int& ej, float& et, float& maj)
{
    const float f = (E - kLutE0f) * kLutInvDEf;
    ej = min(static_cast<int>(f), kLutNd - 2);
    et = f - static_cast<float>(ej);
    maj = CELLS ? s_maj[ej] : hwRcp(s_maj[ej]);
}

// Exposure::sampleParticle for beam history h (rng already seeded): the
// exposure holding h, a direction uniform in the collimation half angles
// (ax, ay) about the exposure's frame, the energy from the spectrum's alias
// table and the weight times the bowtie weight at |ax| (both from the beam
// blob, staged in LDS).  The photon starts at the exposure's source position.
__device__ inline int sampleSource(const DevParams& P, const uint32_t* s_blob, unsigned long long h, Rng& rng, float& px, float& py,
                                   float& pz, float& dx, float& dy, float& dz, float& E, float& w, float& ax, float& ay)
{
    // exposure containing beam history h
    int ei;
    if (P.ppe) {   // h / ppe from the double reciprocal, corrected by one either way
        long long q = static_cast<long long>(static_cast<double>(h) * P.inv_ppe);
        const unsigned long long lo = static_cast<unsigned long long>(q) * P.ppe;
        if (lo > h)
            --q;
        else if (lo + P.ppe <= h)
            ++q;
        ei = static_cast<int>(min(static_cast<unsigned long long>(q), static_cast<unsigned long long>(P.n_ex - 1)));
    } else {
        int lo = 0, hi = P.n_ex - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (P.ex_start[mid] <= h)
                lo = mid;
            else
                hi = mid - 1;
        }
        ei = lo;
    }
    const DevExposure& ex = P.exposures[ei];
    ax = ex.half_angle[0] * (2.0f * rngU(rng) - 1.0f);
    ay = ex.half_angle[1] * (2.0f * rngU(rng) - 1.0f);
    const float sx = __sinf(ax), sy = __sinf(ay);
    const float sz = hwSqrt(fmaxf(0.0f, 1.0f - sx * sx - sy * sy));
    dx = ex.c0[0] * sx + ex.c1[0] * sy + ex.axis[0] * sz;
    dy = ex.c0[1] * sx + ex.c1[1] * sy + ex.axis[1] * sz;
    dz = ex.c0[2] * sx + ex.c1[2] * sy + ex.axis[2] * sz;
    const float inv = rsqrtf(dx * dx + dy * dy + dz * dz);
    dx *= inv;
    dy *= inv;
    dz *= inv;
    px = ex.pos[0];
    py = ex.pos[1];
    pz = ex.pos[2];
    {   // spectrum: Walker alias with one uniform, from LDS
        const int off = static_cast<int>(s_blob[P.bb.o_spec_off + ex.spectrum]);
        const int nb = static_cast<int>(s_blob[P.bb.o_spec_n + ex.spectrum]);
        const float f = rngU(rng) * static_cast<float>(nb);
        int k = min(static_cast<int>(f), nb - 1);
        const float frac = f - static_cast<float>(k);
        if (frac >= asFloat(s_blob[P.bb.o_prob + off + k]))
            k = static_cast<int>(s_blob[P.bb.o_alias + off + k]);
        E = asFloat(s_blob[P.bb.o_energy + off + k]);
    }
    w = ex.weight;
    if (ex.bowtie >= 0) {   // piecewise linear in |angle|, from LDS
        const int off = static_cast<int>(s_blob[P.bb.o_bow_off + ex.bowtie]);
        const int nb = static_cast<int>(s_blob[P.bb.o_bow_n + ex.bowtie]);
        const float* ang = reinterpret_cast<const float*>(s_blob + P.bb.o_angle) + off;
        const float* wt = reinterpret_cast<const float*>(s_blob + P.bb.o_weight) + off;
        const float a = fabsf(ax);
        float bw;
        if (a <= ang[0]) {
            bw = wt[0];
        } else if (a >= ang[nb - 1]) {
            bw = wt[nb - 1];
        } else {
            int k = 0;
            while (ang[k + 1] < a)
                ++k;
            const float t = (a - ang[k]) / (ang[k + 1] - ang[k]);
            bw = wt[k] + t * (wt[k + 1] - wt[k]);
        }
        w *= bw;
    }
    return ei;
}

// Outcome of a flight through the outer medium (outerFlight).
enum FlightEnd { kFlightCollide = 0, kFlightBox = 1, kFlightEscape = 2 };

// A photon in the outer medium (outside the box of the non-outer voxels,
// where every voxel holds outer_word): its flight is sampled exactly with
// that medium's attenuation and ends at a real collision in the outer medium
// (vi receives the voxel), at the box's face (the photon continues there with
// Woodcock steps: flights are memoryless, so the rest is drawn anew) or at the
// grid's edge.  One random number.
__device__ inline FlightEnd outerFlight(const DevParams& P, const float* mu_tab, int lstride, int ej, float et, Rng& rng, float& px,
                                        float& py, float& pz, float dx, float dy, float dz, int& vi)
{
    const int ma = wordMaterial(P.outer_word);
    const float mu_a = lerpTab(mu_tab + ma * lstride, ej, et) * wordDensity(P.outer_word);
    const float s = -hwLog(rngU(rng)) * hwRcp(mu_a);
    const float ix = hwRcp(dx), iy = hwRcp(dy), iz = hwRcp(dz);
    float t0 = 0.0f, t1 = __builtin_inff();
    clipSlab(px, dx, ix, P.outer_lo[0], P.outer_hi[0], t0, t1);
    clipSlab(py, dy, iy, P.outer_lo[1], P.outer_hi[1], t0, t1);
    clipSlab(pz, dz, iz, P.outer_lo[2], P.outer_hi[2], t0, t1);
    // an empty box (a uniform grid: lo above hi) is never entered
    const float t_box = t0 <= t1 && P.outer_lo[0] <= P.outer_hi[0] ? t0 : __builtin_inff();
    float t_edge = 0.0f, t_far = __builtin_inff();   // the grid: where the ray leaves it
    clipSlab(px, dx, ix, P.bmin[0], P.bmax[0], t_edge, t_far);
    clipSlab(py, dy, iy, P.bmin[1], P.bmax[1], t_edge, t_far);
    clipSlab(pz, dz, iz, P.bmin[2], P.bmax[2], t_edge, t_far);
    if (s < fminf(t_box, t_far)) {
        advance(px, py, pz, s, dx, dy, dz);
        vi = voxelIndex(P, px, py, pz);
        return vi < 0 ? kFlightEscape : kFlightCollide;   // vi < 0: rounding at the grid's face
    }
    if (t_box < t_far) {
        advance(px, py, pz, t_box, dx, dy, dz);
        return kFlightBox;
    }
    return kFlightEscape;
}

// Photons of a work-group sampled at their source and carried to the grid's
// box (sampleSource, then boxEntry, then with global tracking the first
// flight through the outer medium) by one wavefront for 64 histories at once,
// with every lane busy, and taken by the dead lanes of any wavefront of the
// group: a refill then costs its lanes three LDS reads instead of the
// sampling, run for the few lanes that happen to be dead.  The draws are
// those the taking lane would have made, in the same order.
//   head counts entries taken, tail entries published; both only grow.  A
//   wavefront samples only while it holds busy and the ring is empty, so it
//   never overwrites an entry that is published and not yet taken.  A taker
//   reads its entries, then claims them with a compare-and-swap on head; when
//   that fails (another wavefront took entries first, after which the ring
//   may have been written again) it reads again.  done is set, before busy
//   clears, by the sampler that finds the instance's pool used up.
constexpr int kRing = 64;
struct SourceQueue {
    unsigned head, tail, busy, done;
    unsigned long long cur, end;   // pool histories claimed by the group and not yet sampled (sampler only)
    float4 ring[3][kRing];         // px py pz dx | dy dz E w | RNG state after sampling
};

__device__ inline void queueReset(SourceQueue& Q)
{
    Q.head = Q.tail = Q.busy = Q.done = 0;
    Q.cur = Q.end = 0;
}

__device__ inline unsigned ldsAcquire(const unsigned* a)
{
    return __builtin_amdgcn_readfirstlane(__hip_atomic_load(a, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP));
}

}  // namespace

// LDS words before the beam blob: the majorant and, when LDS_LUT, the
// attenuation table, lut_n bins each.
template <bool LDS_LUT>
__device__ inline int blobOffset(const DevParams& P)
{
    return P.lut_n + (LDS_LUT ? P.n_mat * P.lut_n : 0);
}

// Stages one instance's tables into LDS: the majorant, the attenuation table
// when LDS_LUT (rows of lut_n bins: the beam's energies need no more), and the
// beam blob.  Returns the attenuation table to read; its row stride is
// lutStride<LDS_LUT>(P).
// Byte offset of the tally-merge area in dynamic LDS: after the beam blob and
// the macro-cell bytes, 8-byte aligned; kMergeWaveBytes per wavefront.
template <bool LDS_LUT>
__device__ inline int mergeOffset(const DevParams& P)
{
    const int cells = P.cell_ratio && P.cell_lds ? (P.cnx * P.cny * P.cnz + 3) / 4 : 0;
    return (4 * (blobOffset<LDS_LUT>(P) + P.bb.words + cells) + 7) & ~7;
}

template <bool LDS_LUT>
__device__ inline int lutStride(const DevParams& P)
{
    return LDS_LUT ? P.lut_n : kLutNd;
}

template <bool LDS_LUT>
__device__ inline const float* stageTables(const DevParams& P, uint32_t* lds)
{
    float* s_maj = reinterpret_cast<float*>(lds);
    float* s_mu = s_maj + P.lut_n;
    uint32_t* s_blob = lds + blobOffset<LDS_LUT>(P);
    for (int i = threadIdx.x; i < P.lut_n; i += blockDim.x)
        s_maj[i] = P.majorant[i];
    if (LDS_LUT)   // rows of lut_n bins (of kLutNd in memory)
        for (int i = threadIdx.x; i < P.n_mat * P.lut_n; i += blockDim.x) {
            const int m = i / P.lut_n;
            s_mu[i] = P.mu_tot[m * kLutNd + (i - m * P.lut_n)];
        }
    for (int i = threadIdx.x; i < P.bb.words; i += blockDim.x)
        s_blob[i] = P.blob[i];
    if (P.cell_ratio && P.cell_lds) {
        const int words = (P.cnx * P.cny * P.cnz + 3) / 4;
        const uint32_t* src = reinterpret_cast<const uint32_t*>(P.cell_q);
        for (int i = threadIdx.x; i < words; i += blockDim.x)
            s_blob[P.bb.words + i] = src[i];
    }
    return LDS_LUT ? s_mu : P.mu_tot;
}

// The history loop of one wavefront for one instance: claims batches of
// [P.h_begin, P.h_end) from P.pool until none is left and every lane is done,
// then adds the wavefront's counters and energy sums to P.counters.
template <bool LDS_LUT, int LEVEL, bool CELLS>
__device__ __forceinline__ void runHistories(const DevParams& P, const float* s_maj, const float* mu_tab, const uint32_t* s_blob,
                                             double* e_wave, SourceQueue& Q, unsigned char* s_merge)
{
    const int lane = static_cast<int>(__lane_id());
    const int lstride = lutStride<LDS_LUT>(P);   // row stride of mu_tab
    float px = 0, py = 0, pz = 0, dx = 0, dy = 0, dz = 1, E = 0, w = 0;
    float maj = 1, et = 0;   // majorant at the photon's energy (its reciprocal without CELLS: energyBin)
    int ej = 0, vi = 0;
    int state = kDead;
    float tau = -1.0f;   // optical depth left in the current flight (macro cells; < 0: draw a new one)
    int cellc = 0;       // macro cell of the photon: cx | cy << 10 | cz << 20
    int mat = 0;         // material of the voxel at the photon (read with its word in the step)
    bool outer = false;  // global tracking: the photon is in the outer medium, outside the box of the rest
    const uint8_t* s_cellq = reinterpret_cast<const uint8_t*>(s_blob + P.bb.words);
    // majorant ratio of macro cell c (LDS bytes or the global float table)
    auto cellRatio = [&](int c) {
        return P.cell_lds ? static_cast<float>(s_cellq[c]) * P.cell_qscale : P.cell_ratio[c];
    };
    // flat macro-cell index (cell counts are below 1024 per axis: full-rate 24-bit multiplies)
    auto cellIndex = [](const DevParams& Q, int cx, int cy, int cz) {
        return static_cast<int>(__umul24(static_cast<unsigned>(Q.cnx), static_cast<unsigned>(cy) + __umul24(static_cast<unsigned>(Q.cny), static_cast<unsigned>(cz))))
               + cx;
    };
    Rng rng { 1, 2, 3, 4 };
    bool collided = false;
    double* const tally = P.tally + (P.tally_replicas > 1 ? static_cast<size_t>(blockIdx.x % P.tally_replicas) * P.tally_stride : 0);

    // wavefront-uniform event counters (popcounts of ballots, kept in SGPRs)
    // (32 bits: a wavefront's share of one launch of at most ~1e8 histories
    // stays far below 2^32 events; 32-bit counters leave SGPRs to the loop)
    unsigned n_steps = 0, n_real = 0, n_dep = 0, n_hist = 0, n_cross = 0;
#ifdef DXG_PHASE_CLOCKS
    // per phase (refill, interact, outer flights, step): wavefront clock
    // cycles, working lanes and iterations; diagnostics builds only
    unsigned long long phase_cycles[4] = { 0, 0, 0, 0 }, phase_lanes[4] = { 0, 0, 0, 0 }, phase_iters[4] = { 0, 0, 0, 0 };
    unsigned long long t_iter = 0;
#define DXG_PHASE_END(k, lanes)                                        \
    do {                                                               \
        phase_cycles[k] += __builtin_amdgcn_s_memtime() - t_iter;      \
        phase_lanes[k] += static_cast<unsigned long long>(lanes);      \
        phase_iters[k] += 1;                                           \
    } while (0)
#else
#define DXG_PHASE_END(k, lanes)
#endif

    for (;;) {
#ifdef DXG_PHASE_CLOCKS
        t_iter = __builtin_amdgcn_s_memtime();
#endif
        const unsigned long long m_dead = __ballot(state == kDead);
        const unsigned long long m_int = __ballot(state == kInteract || state >= kCompton);
        const unsigned long long m_track = __ballot(state == kTrack);
        const unsigned long long m_outer = __ballot(state == kOuter);
        const int n_dead = __popcll(m_dead), n_int = __popcll(m_int), n_track = __popcll(m_track), n_outer = __popcll(m_outer);
        if (n_dead + n_int + n_track + n_outer == 0)
            break;

        // ------------------------------------------------------------ refill
        if (n_dead > 0 && (n_dead >= P.refill_threshold || (n_track == 0 && n_int == 0 && n_outer == 0))) {
            const int rank = __popcll(m_dead & ((1ull << lane) - 1ull));
            const bool dead = state == kDead;
            // take up to n_dead photons from the group's queue
            int k = 0;
            for (;;) {
                const unsigned head = ldsAcquire(&Q.head);
                const unsigned tail = ldsAcquire(&Q.tail);
                k = min(n_dead, static_cast<int>(tail - head));
                if (k == 0)
                    break;
                const bool take = dead && rank < k;
                float4 a = make_float4(0.0f, 0.0f, 0.0f, 0.0f), b = a, c = a;
                if (take) {
                    const int slot = static_cast<int>((head + static_cast<unsigned>(rank)) & (kRing - 1));
                    a = Q.ring[0][slot];
                    b = Q.ring[1][slot];
                    c = Q.ring[2][slot];
                }
                int ok = 0;
                if (lane == 0) {
                    unsigned expected = head;
                    ok = __hip_atomic_compare_exchange_strong(&Q.head, &expected, head + static_cast<unsigned>(k), __ATOMIC_ACQ_REL,
                                                              __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                }
                if (__builtin_amdgcn_readfirstlane(ok)) {
                    if (take) {
                        px = a.x;
                        py = a.y;
                        pz = a.z;
                        dx = a.w;
                        dy = b.x;
                        dz = b.y;
                        E = b.z;
                        w = fabsf(b.w);
                        rng.s0 = __float_as_uint(c.x);
                        rng.s1 = __float_as_uint(c.y);
                        rng.s2 = __float_as_uint(c.z);
                        rng.s3 = __float_as_uint(c.w);
                        collided = false;
                        energyBin<CELLS>(E, s_maj, ej, et, maj);
                        // at the box's face or entered straight into it, else at a
                        // collision in the outer medium (negative weight)
                        outer = b.w < 0.0f;
                        state = kTrack;
                        if (outer) {
                            vi = voxelIndex(P, px, py, pz);
                            mat = wordMaterial(P.outer_word);
                            state = kInteract;
                        }
                        tau = -1.0f;
                        if (CELLS)
                            cellc = cellCoord(px, P.bmin[0], P.cell_inv[0], P.cnx, dx)
                                    | (cellCoord(py, P.bmin[1], P.cell_inv[1], P.cny, dy) << 10)
                                    | (cellCoord(pz, P.bmin[2], P.cell_inv[2], P.cnz, dz) << 20);
                    }
                    break;
                }
            }
            bool sampled = false;
            if (k < n_dead) {
                // the queue ran short: sample the next 64 histories unless
                // another wavefront is at it or the pool is used up
                int got = 0;
                if (lane == 0) {
                    unsigned expected = 0;
                    got = __hip_atomic_load(&Q.done, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) == 0
                          && __hip_atomic_compare_exchange_strong(&Q.busy, &expected, 1u, __ATOMIC_ACQUIRE, __ATOMIC_RELAXED,
                                                                  __HIP_MEMORY_SCOPE_WORKGROUP);
                }
                if (__builtin_amdgcn_readfirstlane(got)) {
                    const unsigned tail = ldsAcquire(&Q.tail);
                    if (ldsAcquire(&Q.head) == tail) {   // every published entry taken: the ring may be written
                        sampled = true;
                        int n = 0;
                        unsigned long long first = 0;
                        if (lane == 0) {
                            const unsigned long long total = P.h_end - P.h_begin;
                            unsigned long long qc = Q.cur, qe = Q.end;
                            if (qc >= qe) {
                                const unsigned long long base = atomicAdd(P.pool, static_cast<unsigned long long>(P.batch));
                                qc = base;
                                qe = min(base + static_cast<unsigned long long>(P.batch), total);
                            }
                            n = qe > qc ? static_cast<int>(min(qe - qc, 64ull)) : 0;
                            first = qc;
                            Q.cur = qc + static_cast<unsigned long long>(n);
                            Q.end = qe;
                            if (n == 0)
                                __hip_atomic_store(&Q.done, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
                        }
                        n = __builtin_amdgcn_readfirstlane(n);
                        first = (static_cast<unsigned long long>(__builtin_amdgcn_readfirstlane(static_cast<unsigned>(first >> 32))) << 32)
                                | __builtin_amdgcn_readfirstlane(static_cast<unsigned>(first));
                        bool hit = false;
                        float sx = 0, sy = 0, sz = 0, sdx = 0, sdy = 0, sdz = 1, sE = 0, sw = 0;
                        Rng sr { 1, 2, 3, 4 };
                        if (lane < n) {
                            const unsigned long long h = P.h_begin + first + static_cast<unsigned long long>(lane);
                            rngSeed(sr, P.seed, h + P.rng_offset);
                            float ax, ay;
                            (void)sampleSource(P, s_blob, h, sr, sx, sy, sz, sdx, sdy, sdz, sE, sw, ax, ay);
                            const double we = static_cast<double>(sw) * static_cast<double>(sE);
                            atomicAdd(&e_wave[0], we);
                            hit = boxEntry(P, sx, sy, sz, sdx, sdy, sdz);
                            if (hit && !CELLS && P.outer_on && !inBox(P.outer_lo, P.outer_hi, sx, sy, sz)) {
                                // the first flight through the outer medium, here with
                                // every lane busy: the photon is queued at the box's face
                                // or at a collision in the outer medium (negative weight)
                                int sej, svi;
                                float set, smaj;
                                energyBin<CELLS>(sE, s_maj, sej, set, smaj);
                                const FlightEnd fe = outerFlight(P, mu_tab, lstride, sej, set, sr, sx, sy, sz, sdx, sdy, sdz, svi);
                                if (fe == kFlightEscape) {
                                    hit = false;
                                    atomicAdd(&e_wave[3], 1.0);   // uncollided exit
                                } else if (fe == kFlightCollide) {
                                    sw = -sw;
                                }
                            }
                            if (!hit)
                                atomicAdd(&e_wave[1], we);
                        }
                        const unsigned long long m_hit = __ballot(hit);
                        if (hit) {
                            const int slot = static_cast<int>((tail + static_cast<unsigned>(__popcll(m_hit & ((1ull << lane) - 1ull)))) & (kRing - 1));
                            Q.ring[0][slot] = make_float4(sx, sy, sz, sdx);
                            Q.ring[1][slot] = make_float4(sdy, sdz, sE, sw);
                            Q.ring[2][slot] = make_float4(__uint_as_float(sr.s0), __uint_as_float(sr.s1), __uint_as_float(sr.s2), __uint_as_float(sr.s3));
                        }
                        n_hist += static_cast<unsigned>(n);
                        if (lane == 0)   // publish, after the ring writes
                            __hip_atomic_fetch_add(&Q.tail, static_cast<unsigned>(__popcll(m_hit)), __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
                    }
                    if (lane == 0)
                        __hip_atomic_store(&Q.busy, 0u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
                }
                // the pool used up and every entry taken: the lanes still dead are done
                if (ldsAcquire(&Q.done) != 0 && ldsAcquire(&Q.head) == ldsAcquire(&Q.tail) && dead && rank >= k)
                    state = kDone;
            }
            DXG_PHASE_END(0, n_dead);
            if (k > 0 || sampled)
                continue;
            if (n_track + n_int + n_outer == 0) {   // nothing to do but wait for another wavefront's sampling
                __builtin_amdgcn_s_sleep(2);
                continue;
            }
            // nothing taken: the other phases run in this iteration
#ifdef DXG_PHASE_CLOCKS
            t_iter = __builtin_amdgcn_s_memtime();
#endif
        }

        // -------------------------------------------------------- interact
        // One attempt per parked lane and phase: a lane whose Compton or
        // Rayleigh rejection sampling rejects stays parked with its process
        // chosen (kCompton / kRayleigh) and tries again in the next phase, so
        // the phase runs each lane's trial once instead of looping until the
        // wavefront's slowest lane accepts.  The trials draw the same numbers
        // in the same order as a loop would.
        // (with no lane tracking, waiting outer flights go first once enough
        // wait: lanes parked for another rejection trial would otherwise run
        // the phase for a few lanes again and again)
        if (n_int > 0 && (n_int >= P.interact_threshold || (n_track == 0 && n_outer < P.outer_threshold))) {
            int deps = 0;
            float e1 = 0.0f, e2 = 0.0f;   // this phase's deposits at vi (tallyFlush)
            double e_scored = 0.0;
            const bool fresh = state == kInteract;
            n_real += __popcll(__ballot(fresh));
            if (fresh || state >= kCompton) {
                const int m = mat;
                int proc;   // 0 photoelectric, 1 Compton, 2 Rayleigh
                if (fresh) {
                    collided = true;
                    tau = -1.0f;   // the photon stays in its macro cell; its next flight is drawn anew
                    const float4 a = P.mu_parts[m * kLutNd + ej];
                    const float4 b = P.mu_parts[m * kLutNd + ej + 1];
                    const float c_pe = __fmaf_rn(et, b.x - a.x, a.x);
                    const float c_inc = __fmaf_rn(et, b.y - a.y, a.y);
                    const float c_tot = __fmaf_rn(et, b.z - a.z, a.z);
                    const float r = rngU(rng) * c_tot;
                    proc = r < c_pe ? 0 : (r < c_inc ? 1 : 2);
                } else {
                    proc = state == kCompton ? 1 : 2;
                }
                state = outer ? kOuter : kTrack;   // the photon stays where it is
                float scatter_cost = 2.0f;   // cos of the scattering angle; 2: the photon does not scatter
                if (proc == 0) {
                    float e_out = 0.0f;   // carried away by a K fluorescence photon (level 2)
                    if (LEVEL == 2) {
                        const float* cum = P.fl_cum + (static_cast<size_t>(m) * kLutNd + ej) * kFlMaxd;
                        const float u = rngU(rng);
                        int k = 0;
                        for (; k < kFlMaxd; ++k)
                            if (u < __fmaf_rn(et, cum[kFlMaxd + k] - cum[k], cum[k]))
                                break;
                        if (k < kFlMaxd) {
                            const float* line = P.fl_line + (m * kFlMaxd + k) * 8;
                            if (rngU(rng) < line[0]) {
                                const float ul = rngU(rng);
                                const float el = ul < line[4] ? line[1] : (ul < line[5] ? line[2] : line[3]);
                                if (el >= kCutoff && el < E)
                                    e_out = el;
                            }
                        }
                    }
                    const float de = w * (E - e_out);
                    e1 = de;
                    deps = 1;
                    e_scored = de;
                    if (e_out > 0.0f) {
                        // the fluorescence photon leaves the absorption point isotropically
                        E = e_out;
                        const float ct = 2.0f * rngU(rng) - 1.0f;
                        const float st = hwSqrt(fmaxf(0.0f, 1.0f - ct * ct));
                        float sp, cp;
                        sinCosRev(rngU(rng), sp, cp);
                        dx = st * cp;
                        dy = st * sp;
                        dz = ct;
                        energyBin<CELLS>(E, s_maj, ej, et, maj);
                    } else {
                        state = kDead;
                    }
                } else if (proc == 1) {
                    // Compton: the angle from an inverse-CDF table with no
                    // rejection, Klein-Nishina at level 0 and Klein-Nishina x
                    // S(x)/Z of the voxel's material at levels 1 and 2; level 2
                    // adds the impulse approximation (one trial): a shell drawn
                    // by electron fraction, p_z from its Compton profile, the
                    // Doppler-shifted E'
                    const float k = E * kInvMeC2;
                    float cost;
                    if (LEVEL == 0) {
                        cost = knCos(P.kn_inv, E, rngU(rng));
                    } else {
                        const float node = (1.0f - hwSqrt(1.0f - rngU(rng))) * static_cast<float>(kKnNDev - 1);
                        cost = knsLookup(P.kns_inv + m * (kKnsRowsDev * kKnNDev), __builtin_amdgcn_logf(E) * kKnsPerOctaveDev, node);
                    }
                    float enew = E * hwRcp(1.0f + k * (1.0f - cost));
                    bool ok = true;
                    if (LEVEL == 2 && ok) {
                        const float us = rngU(rng);
                        const float4* sh = P.ia_shell + m * kIaShellsd;
                        float4 s = sh[0];
                        for (int q = 1; q < kIaShellsd && us > s.x; ++q)
                            s = sh[q];
                        ok = E > s.y;   // else the shell cannot be ionised: another trial
                        if (ok) {
                            const float fy = rngU(rng) * static_cast<float>(kIaInvNd - 1);
                            const int kk = min(static_cast<int>(fy), kIaInvNd - 2);
                            const float pz = lerpTab(P.ia_inv, kk, fy - static_cast<float>(kk)) * s.z;
                            // E'/mc^2 from p_z |q| = E E'(1-cos)/mc^2 - (E - E') (in m_e c units):
                            // b = k (h + p_z sqrt(Q)) / (K^2 - p_z^2), with the discriminant
                            // written as p_z^2 Q so that FP32 does not cancel it away
                            const float K = 1.0f + k * (1.0f - cost);
                            const float p2 = pz * pz;
                            const float h = K - p2 * cost;
                            const float den = K * K - p2;
                            const float Q = (K - cost) * (K - cost) + (1.0f - cost * cost) * (1.0f - p2);
                            ok = den > 0.0f;
                            if (ok) {
                                enew = k * (h + pz * hwSqrt(Q)) * hwRcp(den) * (1.0f / kInvMeC2);
                                ok = enew > 0.0f && E - enew >= s.y;
                            }
                        }
                    }
                    if (!ok) {
                        state = kCompton;   // parked: the next phase draws the next trial
                    } else {
                        const float de = w * (E - enew);
                        e1 = de;
                        deps = 1;
                        e_scored = de;
                        E = enew;
                        scatter_cost = cost;
                        if (E < kCutoff) {
                            const float rest = w * E;
                            e2 = rest;
                            deps = 2;
                            e_scored += rest;
                            state = kDead;
                        } else {
                            energyBin<CELLS>(E, s_maj, ej, et, maj);
                        }
                    }
                } else {
                    // Rayleigh, one trial (level 0 always accepts)
                    float cost;
                    bool ok = true;
                    if (LEVEL == 0) {
                        const float q = 4.0f - 8.0f * rngU(rng);
                        const float aa = -0.5f * q + hwSqrt(0.25f * q * q + 1.0f);
                        const float u = cbrtf(aa);
                        cost = fminf(1.0f, fmaxf(-1.0f, u - hwRcp(u)));
                    } else {
                        const float xm = E * kInvHc;
                        const float umax = xm * xm;
                        const float fx = xm * (static_cast<float>(kSfNd - 1) / kXMaxf);
                        float amax = 1.0f;
                        if (fx < static_cast<float>(kSfNd - 1)) {
                            const int kk = static_cast<int>(fx);
                            amax = lerpTab(P.ray_fwd + m * kSfNd, kk, fx - static_cast<float>(kk));
                        }
                        const float fy = rngU(rng) * amax * static_cast<float>(kRayInvNd - 1);
                        const int kk = min(static_cast<int>(fy), kRayInvNd - 2);
                        const float u = fminf(umax, lerpTab(P.ray_inv + m * kRayInvNd, kk, fy - static_cast<float>(kk)) * (kXMaxf * kXMaxf));
                        cost = 1.0f - 2.0f * u * hwRcp(umax);
                        ok = 2.0f * rngU(rng) <= 1.0f + cost * cost;
                    }
                    if (ok)
                        scatter_cost = cost;
                    else
                        state = kRayleigh;   // parked: the next phase draws the next trial
                }
                // Compton and Rayleigh lanes turn together (the azimuth is the
                // last draw of both processes)
                if (scatter_cost <= 1.0f)
                    rotate(dx, dy, dz, scatter_cost, rngU(rng));
            }
            tallyFlush(P, tally, deps, vi, e1, e2, s_merge);
            if (deps > 0)
                atomicAdd(&e_wave[2], e_scored);
            n_dep += __popcll(__ballot(deps > 0)) + __popcll(__ballot(deps > 1));
            DXG_PHASE_END(1, n_int);
            continue;
        }

        // ------------------------------------------------------------ step
        if (CELLS) {
            // Woodcock tracking with macro-cell majorants: a flight carries an
            // optical depth tau through cells of majorant ratio[c] * maj.  In
            // one iteration a tracking lane walks up to kMaxCross cells
            // (Amanatides-Woo: the next face per axis advances by one cell
            // length over |d|), subtracting mu_c * segment from tau, and ends
            // the flight at a tentative collision inside the cell where tau
            // runs out; that collision is real with probability mu(voxel) / mu_c.
            bool tentative = false;
            int crossed = 0;
            if (state == kTrack) {
                if (tau < 0.0f)
                    tau = -hwLog(rngU(rng));
                int cx = cellc & 1023, cy = (cellc >> 10) & 1023, cz = cellc >> 20;
                float mu_c = cellRatio(cellIndex(P, cx, cy, cz)) * maj;
                const float idx = hwRcp(dx), idy = hwRcp(dy), idz = hwRcp(dz);
                float tmx = exitDistance(px, dx, idx, P.bmin[0], P.cell_len[0], cx);
                float tmy = exitDistance(py, dy, idy, P.bmin[1], P.cell_len[1], cy);
                float tmz = exitDistance(pz, dz, idz, P.bmin[2], P.cell_len[2], cz);
                float t0 = 0.0f;   // path walked in this iteration up to the current cell's entry face
                float s = 0.0f;    // path to move the photon by at the end of the iteration
                bool out = false;
                for (;;) {
                    const float t1 = fminf(tmx, fminf(tmy, tmz));
                    const float seg = t1 - t0;
                    if (tau < mu_c * seg) {
                        tentative = true;
                        s = t0 + tau * hwRcp(mu_c);
                        tau = -1.0f;
                        break;
                    }
                    if (crossed == kMaxCross) {   // the rest of the flight goes on next iteration
                        s = t0;
                        break;
                    }
                    tau -= mu_c * seg;
                    t0 = t1;
                    ++crossed;
                    if (tmx == t1) {
                        cx += dx > 0.0f ? 1 : -1;
                        tmx += P.cell_len[0] * fabsf(idx);   // one cell further (dx != 0 here: tmx was finite)
                    } else if (tmy == t1) {
                        cy += dy > 0.0f ? 1 : -1;
                        tmy += P.cell_len[1] * fabsf(idy);   // one cell further (dy != 0 here: tmy was finite)
                    } else {
                        cz += dz > 0.0f ? 1 : -1;
                        tmz += P.cell_len[2] * fabsf(idz);   // one cell further (dz != 0 here: tmz was finite)
                    }
                    if (static_cast<unsigned>(cx) >= static_cast<unsigned>(P.cnx) || static_cast<unsigned>(cy) >= static_cast<unsigned>(P.cny)
                        || static_cast<unsigned>(cz) >= static_cast<unsigned>(P.cnz)) {
                        out = true;
                        s = t0;
                        break;
                    }
                    mu_c = cellRatio(cellIndex(P, cx, cy, cz)) * maj;
                }
                advance(px, py, pz, s, dx, dy, dz);
                cellc = cx | (cy << 10) | (cz << 20);
                if (tentative)
                    vi = voxelIndex(P, px, py, pz);
                if (out || (tentative && vi < 0)) {
                    atomicAdd(&e_wave[1], static_cast<double>(w) * static_cast<double>(E));
                    if (!collided)
                        atomicAdd(&e_wave[3], 1.0);   // uncollided exit
                    state = kDead;
                    tentative = false;
                } else if (tentative) {
                    const uint32_t word = P.voxels[vi];
                    const int m = wordMaterial(word);
                    const float rho = wordDensity(word);
                    const float mu = lerpTab(mu_tab + m * lstride, ej, et) * rho;
                    mat = m;
                    if (!(rngU(rng) * mu_c > mu))
                        state = kInteract;
                }
            }
            n_steps += __popcll(__ballot(tentative));
            for (int b = 0; b < 4; ++b)   // crossed <= kMaxCross < 16: sum it bit by bit
                n_cross += static_cast<unsigned>(__popcll(__ballot((crossed >> b) & 1))) << b;
            DXG_PHASE_END(3, n_track);
            continue;
        }
        // ----------------------------------------------- outer-medium flights
        // Lanes outside the box of the non-outer voxels (global tracking) wait
        // like parked collisions and fly together once outer_threshold of them
        // are waiting or no lane tracks, so a step never runs this code for a
        // few lanes beside the Woodcock steps of the rest.
        if (n_outer > 0 && (n_outer >= P.outer_threshold || n_track == 0)) {
            if (state == kOuter) {
                const FlightEnd end = outerFlight(P, mu_tab, lstride, ej, et, rng, px, py, pz, dx, dy, dz, vi);
                if (end == kFlightCollide) {
                    mat = wordMaterial(P.outer_word);
                    state = kInteract;
                } else if (end == kFlightBox) {
                    outer = false;
                    state = kTrack;
                } else {
                    atomicAdd(&e_wave[1], static_cast<double>(w) * static_cast<double>(E));
                    if (!collided)
                        atomicAdd(&e_wave[3], 1.0);   // uncollided exit
                    state = kDead;
                }
            }
            DXG_PHASE_END(2, n_outer);
            continue;
        }

        if (state == kTrack) {
            const float mfp = maj;   // the global instance keeps 1 / majorant
            const float s = -hwLog(rngU(rng)) * mfp;
            advance(px, py, pz, s, dx, dy, dz);
            vi = voxelIndex(P, px, py, pz);
            if (vi < 0) {
                atomicAdd(&e_wave[1], static_cast<double>(w) * static_cast<double>(E));
                if (!collided)
                    atomicAdd(&e_wave[3], 1.0);   // uncollided exit
                state = kDead;
            } else {
                const uint32_t word = P.voxels[vi];
                const int m = wordMaterial(word);
                const float rho = wordDensity(word);
                const float mu = lerpTab(mu_tab + m * lstride, ej, et) * rho;
                mat = m;
                if (P.outer_on && word == P.outer_word)   // outer medium: inside the box or out of it
                    outer = !inBox(P.outer_lo, P.outer_hi, px, py, pz);
                if (!(rngU(rng) > mu * mfp))
                    state = kInteract;   // real collision; virtual ones keep tracking
                else if (outer)
                    state = kOuter;      // left the box: the next flight is exact
            }
        }
        n_steps += n_track;
        DXG_PHASE_END(3, n_track);
    }
#undef DXG_PHASE_END

    // ---- one atomic per wavefront and counter
    const double e_emit = e_wave[0], e_esc = e_wave[1], e_dep = e_wave[2];
    const unsigned long long n_unc = static_cast<unsigned long long>(e_wave[3]);
    if (lane == 0) {
        atomicAdd(&P.counters->steps, static_cast<unsigned long long>(n_steps));
        atomicAdd(&P.counters->real_events, static_cast<unsigned long long>(n_real));
        atomicAdd(&P.counters->deposits, static_cast<unsigned long long>(n_dep));
        atomicAdd(&P.counters->histories, static_cast<unsigned long long>(n_hist));
        atomicAdd(&P.counters->uncollided, n_unc);
        atomicAdd(&P.counters->crossings, static_cast<unsigned long long>(n_cross));
#ifdef DXG_PHASE_CLOCKS
        for (int k = 0; k < 4; ++k) {
            atomicAdd(&P.counters->phase_cycles[k], phase_cycles[k]);
            atomicAdd(&P.counters->phase_lanes[k], phase_lanes[k]);
            atomicAdd(&P.counters->phase_iters[k], phase_iters[k]);
        }
#endif
        atomicAdd(&P.counters->emitted, e_emit);
        atomicAdd(&P.counters->escaped, e_esc);
        atomicAdd(&P.counters->deposited, e_dep);
    }
}

template <bool LDS_LUT, int LEVEL, bool CELLS>
__global__ void __launch_bounds__(kBlock) DXG_TRANSPORT_OCCUPANCY transport_kernel(const DevParams P)
{
    extern __shared__ uint32_t lds[];
    const float* mu_tab = stageTables<LDS_LUT>(P, lds);
    // per-wavefront FP64 sums in LDS: energy emitted, escaped, deposited; uncollided exits
    __shared__ double s_energy[kBlock / 64][4];
    __shared__ SourceQueue s_queue;
    const int lane = static_cast<int>(__lane_id());
    const int wave = static_cast<int>(threadIdx.x) >> 6;
    if (lane < 4)
        s_energy[wave][lane] = 0.0;
    if (threadIdx.x == 0)
        queueReset(s_queue);
    __syncthreads();
    const float* s_maj = reinterpret_cast<const float*>(lds);
    const uint32_t* s_blob = lds + blobOffset<LDS_LUT>(P);
    unsigned char* const s_merge = P.tally_merge ? reinterpret_cast<unsigned char*>(lds) + mergeOffset<LDS_LUT>(P) + wave * kMergeWaveBytes : nullptr;
    runHistories<LDS_LUT, LEVEL, CELLS>(P, s_maj, mu_tab, s_blob, s_energy[wave], s_queue, s_merge);
}

// Many instances (world + beam) in one persistent launch.  A block starts on
// the instance start_inst[blockIdx.x]; when that instance's pool is empty and
// all of the block's wavefronts are done, the block moves to an instance with
// histories left (scanning from a block-dependent offset), re-stages its LDS
// tables and continues, until no instance has work.
template <bool LDS_LUT, int LEVEL, bool CELLS>
__global__ void __launch_bounds__(kBlock) DXG_TRANSPORT_OCCUPANCY transport_batch_kernel(const DevParams* __restrict__ params, int n_inst,
                                                                 const int* __restrict__ start_inst)
{
    extern __shared__ uint32_t lds[];
    __shared__ double s_energy[kBlock / 64][4];
    __shared__ SourceQueue s_queue;
    __shared__ int s_next;
    const int lane = static_cast<int>(__lane_id());
    const int wave = static_cast<int>(threadIdx.x) >> 6;
    int inst = start_inst[blockIdx.x];
    for (;;) {
        inst = __builtin_amdgcn_readfirstlane(inst);
        const DevParams& P = params[inst];
        const float* mu_tab = stageTables<LDS_LUT>(P, lds);
        if (lane < 4)
            s_energy[wave][lane] = 0.0;
        if (threadIdx.x == 0)
            queueReset(s_queue);
        __syncthreads();
        const float* s_maj = reinterpret_cast<const float*>(lds);
        const uint32_t* s_blob = lds + blobOffset<LDS_LUT>(P);
        unsigned char* const s_merge = P.tally_merge ? reinterpret_cast<unsigned char*>(lds) + mergeOffset<LDS_LUT>(P) + wave * kMergeWaveBytes : nullptr;
        runHistories<LDS_LUT, LEVEL, CELLS>(P, s_maj, mu_tab, s_blob, s_energy[wave], s_queue, s_merge);
        __syncthreads();   // every wavefront is done with this instance's tables
        if (threadIdx.x == 0) {
            int next = -1;
            for (int k = 0; k < n_inst && next < 0; ++k) {
                const int c = (static_cast<int>(blockIdx.x) + k) % n_inst;
                const DevParams& Q = params[c];
                const unsigned long long claimed = __hip_atomic_load(Q.pool, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (claimed < Q.h_end - Q.h_begin)
                    next = c;
            }
            s_next = next;
        }
        __syncthreads();
        inst = s_next;
        if (inst < 0)
            break;
    }
}

__global__ void batch_reset_kernel(const DevParams* __restrict__ params, int n_inst)
{
    const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n_inst)
        return;
    *params[i].pool = 0;
    DevCounters* c = params[i].counters;
    c->steps = 0;
    c->real_events = 0;
    c->deposits = 0;
    c->histories = 0;
    c->emitted = 0;
    c->escaped = 0;
    c->deposited = 0;
    c->uncollided = 0;
    c->crossings = 0;
    for (int k = 0; k < 4; ++k)
        c->phase_cycles[k] = c->phase_lanes[k] = c->phase_iters[k] = 0;
}

// The traversal of the transport kernels replayed on given rays: the slab
// entry of the refill path (boxEntry), then per step advance() and the voxel
// lookup of the Woodcock step (voxelIndex, wordMaterial) -- the same device
// functions, so the indices are the ones tracking visits.
__global__ void traverse_kernel(const DevParams P, uint64_t n_rays, const float* __restrict__ pos, const float* __restrict__ dir,
                                uint64_t n_steps, const float* __restrict__ steps, int64_t* __restrict__ out_index,
                                int32_t* __restrict__ out_material)
{
    const uint64_t r = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= n_rays)
        return;
    float px = pos[3 * r], py = pos[3 * r + 1], pz = pos[3 * r + 2];
    const float dx = dir[3 * r], dy = dir[3 * r + 1], dz = dir[3 * r + 2];
    const bool hit = boxEntry(P, px, py, pz, dx, dy, dz);
    for (uint64_t k = 0; k < n_steps; ++k) {
        int64_t vi = -1;
        int32_t mm = -1;
        if (hit) {
            advance(px, py, pz, steps[r * n_steps + k], dx, dy, dz);
            const int v = voxelIndex(P, px, py, pz);
            if (v >= 0) {
                vi = v;
                mm = wordMaterial(P.voxels[v]);
            }
        }
        out_index[r * n_steps + k] = vi;
        out_material[r * n_steps + k] = mm;
    }
}

// Exposure::sampleParticle replayed for beam histories [0, n) through the
// transport kernels' own sampleSource, with the beam blob staged in LDS as
// they stage it.  out[8 h + {0..7}] = ax, ay, dx, dy, dz, E, w, exposure index.
__global__ void source_kernel(const DevParams P, uint64_t n, float* __restrict__ out)
{
    extern __shared__ uint32_t lds[];
    for (int i = threadIdx.x; i < P.bb.words; i += blockDim.x)
        lds[i] = P.blob[i];
    __syncthreads();
    const uint64_t h = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (h >= n)
        return;
    Rng rng;
    rngSeed(rng, P.seed, h + P.rng_offset);
    float px, py, pz, dx, dy, dz, E, w, ax, ay;
    const int ei = sampleSource(P, lds, h, rng, px, py, pz, dx, dy, dz, E, w, ax, ay);
    float* o = out + 8 * h;
    o[0] = ax;
    o[1] = ay;
    o[2] = dx;
    o[3] = dy;
    o[4] = dz;
    o[5] = E;
    o[6] = w;
    o[7] = static_cast<float>(ei);
}

__global__ void tally_fold_kernel(double* __restrict__ t, unsigned long long stride, int replicas)
{
    for (unsigned long long i = static_cast<unsigned long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < stride;
         i += static_cast<unsigned long long>(gridDim.x) * blockDim.x) {
        double s = t[i];
        for (int r = 1; r < replicas; ++r) {
            double* q = t + static_cast<size_t>(r) * stride + i;
            s += *q;
            *q = 0.0;
        }
        t[i] = s;
    }
}

__global__ void reset_kernel(unsigned long long* pool, DevCounters* c)
{
    if (threadIdx.x == 0) {
        *pool = 0;
        c->steps = 0;
        c->real_events = 0;
        c->deposits = 0;
        c->histories = 0;
        c->emitted = 0;
        c->escaped = 0;
        c->deposited = 0;
        c->uncollided = 0;
        c->crossings = 0;
        for (int k = 0; k < 4; ++k)
            c->phase_cycles[k] = c->phase_lanes[k] = c->phase_iters[k] = 0;
    }
}

__global__ void copy_kernel(const float4* __restrict__ src, float4* __restrict__ dst, uint64_t n)
{
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
        dst[i] = src[i];
}

size_t transportLdsBytes(const DevParams& p)
{
    const size_t cells = p.cell_ratio && p.cell_lds ? (static_cast<size_t>(p.cnx) * p.cny * p.cnz + 3) / 4 : 0;
    const size_t bytes = sizeof(uint32_t) * (static_cast<size_t>(p.lut_n) + (p.lds_lut ? static_cast<size_t>(p.n_mat) * p.lut_n : 0)
                                             + static_cast<size_t>(p.bb.words) + cells);
    return p.tally_merge ? ((bytes + 7) & ~static_cast<size_t>(7)) + static_cast<size_t>(kBlock / 64) * kMergeWaveBytes : bytes;
}

namespace {

// The kernel instance for (LDS table placement, correction level, macro cells).
const void* transportInstance(bool lds_lut, int level, bool cells)
{
    using K = void (*)(DevParams);
    static const K table[2][2][3] = {
        { { &transport_kernel<false, 0, false>, &transport_kernel<false, 1, false>, &transport_kernel<false, 2, false> },
          { &transport_kernel<true, 0, false>, &transport_kernel<true, 1, false>, &transport_kernel<true, 2, false> } },
        { { &transport_kernel<false, 0, true>, &transport_kernel<false, 1, true>, &transport_kernel<false, 2, true> },
          { &transport_kernel<true, 0, true>, &transport_kernel<true, 1, true>, &transport_kernel<true, 2, true> } },
    };
    return reinterpret_cast<const void*>(table[cells ? 1 : 0][lds_lut ? 1 : 0][level < 0 ? 0 : (level > 2 ? 2 : level)]);
}

}  // namespace

void launchTransport(const DevParams& p, int blocks, int threads, hipStream_t s)
{
    const size_t lds = transportLdsBytes(p);
    void* args[] = { const_cast<DevParams*>(&p) };
    (void)hipLaunchKernel(transportInstance(p.lds_lut != 0, p.level, p.cell_ratio != nullptr), dim3(static_cast<unsigned>(blocks)),
                          dim3(static_cast<unsigned>(threads)), args, lds, s);
}

namespace {

const void* batchInstance(bool lds_lut, int level, bool cells)
{
    using K = void (*)(const DevParams*, int, const int*);
    static const K table[2][2][3] = {
        { { &transport_batch_kernel<false, 0, false>, &transport_batch_kernel<false, 1, false>, &transport_batch_kernel<false, 2, false> },
          { &transport_batch_kernel<true, 0, false>, &transport_batch_kernel<true, 1, false>, &transport_batch_kernel<true, 2, false> } },
        { { &transport_batch_kernel<false, 0, true>, &transport_batch_kernel<false, 1, true>, &transport_batch_kernel<false, 2, true> },
          { &transport_batch_kernel<true, 0, true>, &transport_batch_kernel<true, 1, true>, &transport_batch_kernel<true, 2, true> } },
    };
    return reinterpret_cast<const void*>(table[cells ? 1 : 0][lds_lut ? 1 : 0][level < 0 ? 0 : (level > 2 ? 2 : level)]);
}

}  // namespace

void launchTransportBatch(const DevParams* params, int n_inst, const int* start_inst, bool lds_lut, int level, bool cells,
                          size_t lds, int blocks, hipStream_t s)
{
    hipLaunchKernelGGL(batch_reset_kernel, dim3(static_cast<unsigned>((n_inst + 63) / 64)), dim3(64), 0, s, params, n_inst);
    void* args[] = { const_cast<DevParams**>(&params), &n_inst, const_cast<int**>(&start_inst) };
    (void)hipLaunchKernel(batchInstance(lds_lut, level, cells), dim3(static_cast<unsigned>(blocks)), dim3(kBlock), args, lds, s);
}

int transportBatchBlocksPerCU(bool lds_lut, int level, bool cells, size_t lds_bytes)
{
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, batchInstance(lds_lut, level, cells), kBlock, lds_bytes) != hipSuccess)
        return 1;
    return std::clamp(n, 1, kMaxBlocksPerCU);
}

int transportBlocksPerCU(bool lds_lut, int level, bool cells, size_t lds_bytes)
{
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, transportInstance(lds_lut, level, cells), kBlock, lds_bytes) != hipSuccess)
        return 1;
    return std::clamp(n, 1, kMaxBlocksPerCU);
}

void launchTraverse(const DevParams& p, uint64_t n_rays, const float* pos, const float* dir, uint64_t n_steps,
                    const float* steps, int64_t* out_index, int32_t* out_material, hipStream_t s)
{
    const unsigned blocks = static_cast<unsigned>((n_rays + 255) / 256);
    hipLaunchKernelGGL(traverse_kernel, dim3(blocks), dim3(256), 0, s, p, n_rays, pos, dir, n_steps, steps, out_index, out_material);
}

void launchSource(const DevParams& p, uint64_t n, float* out, hipStream_t s)
{
    const unsigned blocks = static_cast<unsigned>((n + 255) / 256);
    hipLaunchKernelGGL(source_kernel, dim3(blocks), dim3(256), sizeof(uint32_t) * static_cast<size_t>(p.bb.words), s, p, n, out);
}

void launchTallyFold(double* tally, unsigned long long stride, int replicas, hipStream_t s)
{
    if (replicas <= 1)
        return;
    const unsigned long long want = (stride + 255) / 256;
    const unsigned blocks = static_cast<unsigned>(std::min<unsigned long long>(want, 8192));
    hipLaunchKernelGGL(tally_fold_kernel, dim3(blocks), dim3(256), 0, s, tally, stride, replicas);
}

void launchReset(unsigned long long* pool, DevCounters* counters, hipStream_t s)
{
    hipLaunchKernelGGL(reset_kernel, dim3(1), dim3(64), 0, s, pool, counters);
}

void launchCopy(const void* src, void* dst, uint64_t bytes, hipStream_t s)
{
    const uint64_t n = bytes / sizeof(float4);
    hipLaunchKernelGGL(copy_kernel, dim3(4096), dim3(256), 0, s, static_cast<const float4*>(src), static_cast<float4*>(dst), n);
}

int transportBlockSize() { return kBlock; }

// ---------------------------------------------------------------------------
// Scratch of the allocation-free run (dxg_transport_run_static): one array per
// device, held until the process ends.  Host code never passes `scratch`
// itself (in host code the name is a stand-in, not a device address): it asks
// staticScratch() for the array's address on the current device, taken once
// per device with hipGetSymbolAddress.
constexpr size_t kStaticScratchBytes = size_t(1) << 20;
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxDevices] = {};   // scratch's address per device, once taken
}  // namespace

size_t staticScratchBytes() { return kStaticScratchBytes; }

int staticScratch(int device, void** out)
{
    if (device < 0 || device >= kMaxDevices)
        return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[device]) {
        int cur = -1;
        hipError_t e = hipGetDevice(&cur);
        if (e != hipSuccess)
            return e;
        if (cur != device)
            return hipErrorInvalidDevice;   // the device must be current
        void* ptr = nullptr;
        e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess)
            return e;
        g_scratch_ptr[device] = ptr;
    }
    *out = g_scratch_ptr[device];
    return hipSuccess;
}

}  // namespace dxg
