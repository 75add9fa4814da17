// HIP/CDNA4 kernels of the voxel photon-transport engine.
//
// transport_kernel is a persistent history-pool kernel.  Every lane carries
// one photon in registers (FP32 state).  When a wavefront has enough dead
// lanes (absorbed, escaped, missed the grid) -- refill_threshold of them, or
// all -- it claims exactly that many new histories from a global pool counter
// with ONE atomic and hands them to the dead lanes in lane order, so the
// expensive source-sampling path is paid once for many lanes.  Each loop
// iteration every live lane advances one Woodcock step: one 4-byte gather of
// the packed voxel word (density + material id), one LDS read of the
// material's attenuation at the photon's energy bin, one compare.  Deposits go
// straight to the FP64 per-voxel tally with hardware float atomics; event
// counters are reduced across the wavefront before one atomic per wavefront.
//
// LDS holds, per block: the Woodcock majorant table, the per-material total
// attenuation table (when it fits), and the beam's spectrum alias tables and
// bowtie nodes.
//
// Algorithms (DXMClib's AAVoxelGrid::transport and interactions<5, C>, absent
// from the reference tree; the caller is simulationpipeline.cpp:159-170):
//   Woodcock tracking with a per-energy-bin majorant; photoelectric
//   absorption scored as KERMA; Klein-Nishina by Butcher & Messel (level 1:
//   rejection on S(x)/Z); Thomson by the exact inverse of its cubic CDF
//   (level 1: x^2 drawn from F^2(x), accepted with (1+cos^2)/2).
#include "device.hpp"

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
constexpr float kTwoPi = 6.283185307179586f;
constexpr int kBlock = 256;

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

// xoshiro128+; uniform in (0, 1]
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
    return static_cast<float>((out >> 8) + 1u) * (1.0f / 16777216.0f);
}

// Slab entry of a ray into the grid box.  Identical operation order to the
// oracle's FP32 traversal (bit-exact check).
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

// Flat voxel index of a point, or -1 outside the grid.
__device__ inline int voxelOf(const DevParams& P, float px, float py, float pz)
{
    const int ix = static_cast<int>(floorf((px - P.bmin[0]) * P.inv_spacing[0]));
    const int iy = static_cast<int>(floorf((py - P.bmin[1]) * P.inv_spacing[1]));
    const int iz = static_cast<int>(floorf((pz - P.bmin[2]) * P.inv_spacing[2]));
    if (static_cast<unsigned>(ix) >= static_cast<unsigned>(P.nx) || static_cast<unsigned>(iy) >= static_cast<unsigned>(P.ny)
        || static_cast<unsigned>(iz) >= static_cast<unsigned>(P.nz))
        return -1;
    return ix + P.nx * (iy + P.ny * iz);
}

__device__ inline void rotate(float& dx, float& dy, float& dz, float cost, float phi)
{
    const float sint = sqrtf(fmaxf(0.0f, 1.0f - cost * cost));
    float sp, cp;
    __sincosf(phi, &sp, &cp);
    float nx, ny, nz;
    if (fabsf(dz) > 0.99999f) {
        nx = sint * cp;
        ny = sint * sp;
        nz = dz > 0.0f ? cost : -cost;
    } else {
        const float t = sqrtf(1.0f - dz * dz);
        const float it = 1.0f / t;
        nx = dx * cost + sint * (dx * dz * cp - dy * sp) * it;
        ny = dy * cost + sint * (dy * dz * cp + dx * sp) * it;
        nz = dz * cost - t * sint * cp;
    }
    const float inv = rsqrtf(nx * nx + ny * ny + nz * nz);
    dx = nx * inv;
    dy = ny * inv;
    dz = nz * inv;
}

__device__ inline void tallyAdd(const DevParams& P, int idx, float e)
{
    const double de = static_cast<double>(e);
    double* t = P.tally + 4 * static_cast<size_t>(idx);
    atomicAdd(t, de);
    atomicAdd(t + 1, de * de);
    atomicAdd(reinterpret_cast<unsigned long long*>(t + 2), 1ull);
}

template <typename T>
__device__ inline T waveSum(T v)
{
    for (int o = 32; o > 0; o >>= 1)
        v += __shfl_xor(v, o, 64);
    return v;
}

__device__ inline float lerpTab(const float* tab, int j, float t)
{
    const float a = tab[j];
    const float b = tab[j + 1];
    return __fmaf_rn(t, b - a, a);
}

__device__ inline float asFloat(uint32_t u) { return __uint_as_float(u); }

}  // namespace

template <bool LDS_LUT>
__global__ void __launch_bounds__(kBlock) transport_kernel(const DevParams P)
{
    extern __shared__ uint32_t lds[];
    float* s_maj = reinterpret_cast<float*>(lds);
    float* s_mu = s_maj + kLutNd;
    uint32_t* s_blob = lds + kLutNd + (LDS_LUT ? P.n_mat * kLutNd : 0);
    for (int i = threadIdx.x; i < kLutNd; i += blockDim.x)
        s_maj[i] = P.majorant[i];
    if (LDS_LUT)
        for (int i = threadIdx.x; i < P.n_mat * kLutNd; i += blockDim.x)
            s_mu[i] = P.mu_tot[i];
    for (int i = threadIdx.x; i < P.bb.words; i += blockDim.x)
        s_blob[i] = P.blob[i];
    __syncthreads();
    const float* mu_tab = LDS_LUT ? s_mu : P.mu_tot;

    const int lane = static_cast<int>(__lane_id());
    float px = 0, py = 0, pz = 0, dx = 0, dy = 0, dz = 1, E = 0, w = 0;
    float maj = 1, inv_maj = 1, et = 0;
    int ej = 0;
    Rng rng { 1, 2, 3, 4 };
    bool alive = false, done = false, collided = false;

    unsigned long long n_steps = 0, n_real = 0, n_dep = 0, n_hist = 0, n_unc = 0;
    double e_emit = 0, e_esc = 0, e_dep = 0;
    // wavefront-private batch of pool histories [cur, end), claimed P.batch at a time
    unsigned long long cur = 0, end = 0;

    for (;;) {
        // ---- refill: once enough lanes are dead, they take the next histories together
        const bool need = !alive && !done;
        const unsigned long long need_mask = __ballot(need);
        const int n_need = __popcll(need_mask);
        if (n_need > 0 && (n_need >= P.refill_threshold || __ballot(alive) == 0)) {
            const unsigned long long rank = __popcll(need_mask & ((1ull << lane) - 1ull));
            const unsigned long long avail = end - cur;
            unsigned long long mine;
            if (avail >= static_cast<unsigned long long>(n_need)) {
                mine = cur + rank;
                cur += n_need;
            } else {
                // the rest of this batch first, then a fresh batch from the pool
                unsigned long long base = 0;
                if (lane == __ffsll(static_cast<long long>(need_mask)) - 1)
                    base = atomicAdd(P.pool, static_cast<unsigned long long>(P.batch));
                base = __shfl(base, __ffsll(static_cast<long long>(need_mask)) - 1, 64);
                mine = rank < avail ? cur + rank : base + (rank - avail);
                cur = base + (n_need - avail);
                end = base + P.batch;
            }
            if (need) {
                const unsigned long long h = P.h_begin + mine;
                if (h >= P.h_end) {
                    done = true;
                } else {
                    ++n_hist;
                    collided = false;
                    rngSeed(rng, P.seed, h + P.rng_offset);
                    // exposure containing beam history h
                    int ei;
                    if (P.ppe) {
                        ei = static_cast<int>(min(h / P.ppe, static_cast<unsigned long long>(P.n_ex - 1)));
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
                    // Exposure::sampleParticle
                    const float ax = ex.half_angle[0] * (2.0f * rngU(rng) - 1.0f);
                    const float ay = ex.half_angle[1] * (2.0f * rngU(rng) - 1.0f);
                    const float sx = __sinf(ax), sy = __sinf(ay);
                    const float sz = sqrtf(fmaxf(0.0f, 1.0f - sx * sx - sy * sy));
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
                    e_emit += static_cast<double>(w) * static_cast<double>(E);
                    alive = boxEntry(P, px, py, pz, dx, dy, dz);
                    if (alive) {
                        const float f = (E - kLutE0f) * kLutInvDEf;
                        ej = min(static_cast<int>(f), kLutNd - 2);
                        et = f - static_cast<float>(ej);
                        maj = s_maj[ej];
                        inv_maj = 1.0f / maj;
                    } else {
                        e_esc += static_cast<double>(w) * static_cast<double>(E);
                    }
                }
            }
        }
        if (__ballot(alive) == 0) {
            if (__ballot(!done) == 0)
                break;
            continue;
        }
        if (!alive)
            continue;

        // ---- one Woodcock step
        const float s = -__logf(rngU(rng)) * inv_maj;
        px = __fmaf_rn(s, dx, px);
        py = __fmaf_rn(s, dy, py);
        pz = __fmaf_rn(s, dz, pz);
        ++n_steps;
        const int vi = voxelOf(P, px, py, pz);
        if (vi < 0) {
            e_esc += static_cast<double>(w) * static_cast<double>(E);
            n_unc += collided ? 0ull : 1ull;
            alive = false;
            continue;
        }
        const uint32_t word = P.voxels[vi];
        const int m = static_cast<int>(word & 0xFFu);
        const float rho = asFloat(word & 0xFFFFFF00u);
        const float mu = lerpTab(mu_tab + m * kLutNd, ej, et) * rho;
        if (rngU(rng) * maj > mu)
            continue;   // virtual collision

        // ---- real interaction
        ++n_real;
        collided = true;
        const float4 a = P.mu_parts[m * kLutNd + ej];
        const float4 b = P.mu_parts[m * kLutNd + ej + 1];
        const float c_pe = __fmaf_rn(et, b.x - a.x, a.x);
        const float c_inc = __fmaf_rn(et, b.y - a.y, a.y);
        const float c_tot = __fmaf_rn(et, b.z - a.z, a.z);
        const float r = rngU(rng) * c_tot;
        if (r < c_pe) {
            const float de = w * E;
            tallyAdd(P, vi, de);
            ++n_dep;
            e_dep += static_cast<double>(de);
            alive = false;
            continue;
        }
        float cost;
        if (r < c_inc) {
            // Compton: Klein-Nishina (Butcher & Messel), level 1 rejects on S(x)/Z
            const float k = E * kInvMeC2;
            const float e0 = 1.0f / (1.0f + 2.0f * k);
            const float a1 = __logf(1.0f + 2.0f * k);
            const float a2 = 0.5f * (1.0f - e0 * e0);
            float eps;
            for (;;) {
                const float r1 = rngU(rng), r2 = rngU(rng), r3 = rngU(rng);
                if (r1 * (a1 + a2) < a1)
                    eps = __expf(-a1 * r2);
                else
                    eps = sqrtf(e0 * e0 + (1.0f - e0 * e0) * r2);
                const float t = (1.0f - eps) / (k * eps);
                const float sin2 = t * (2.0f - t);
                const float g = 1.0f - eps * sin2 / (1.0f + eps * eps);
                if (r3 > g)
                    continue;
                cost = 1.0f - t;
                if (P.level == 0)
                    break;
                const float x = E * kInvHc * sqrtf(fmaxf(0.0f, 0.5f * (1.0f - cost)));
                const float fx = x * (static_cast<float>(kSfNd - 1) / kXMaxf);
                float sv = 1.0f;
                if (fx < static_cast<float>(kSfNd - 1)) {
                    const int kk = static_cast<int>(fx);
                    sv = lerpTab(P.s_table + m * kSfNd, kk, fx - static_cast<float>(kk));
                }
                if (rngU(rng) <= sv)
                    break;
            }
            const float enew = E * eps;
            const float de = w * (E - enew);
            tallyAdd(P, vi, de);
            ++n_dep;
            e_dep += static_cast<double>(de);
            E = enew;
            rotate(dx, dy, dz, cost, kTwoPi * rngU(rng));
            if (E < kCutoff) {
                const float rest = w * E;
                tallyAdd(P, vi, rest);
                ++n_dep;
                e_dep += static_cast<double>(rest);
                alive = false;
                continue;
            }
            const float f = (E - kLutE0f) * kLutInvDEf;
            ej = min(static_cast<int>(f), kLutNd - 2);
            et = f - static_cast<float>(ej);
            maj = s_maj[ej];
            inv_maj = 1.0f / maj;
        } else {
            // Rayleigh
            if (P.level == 0) {
                const float q = 4.0f - 8.0f * rngU(rng);
                const float aa = -0.5f * q + sqrtf(0.25f * q * q + 1.0f);
                const float u = cbrtf(aa);
                cost = fminf(1.0f, fmaxf(-1.0f, u - 1.0f / u));
            } else {
                const float xm = E * kInvHc;
                const float umax = xm * xm;
                const float fx = xm * (static_cast<float>(kSfNd - 1) / kXMaxf);
                float amax = 1.0f;
                if (fx < static_cast<float>(kSfNd - 1)) {
                    const int kk = static_cast<int>(fx);
                    amax = lerpTab(P.ray_fwd + m * kSfNd, kk, fx - static_cast<float>(kk));
                }
                for (;;) {
                    const float fy = rngU(rng) * amax * static_cast<float>(kRayInvNd - 1);
                    const int kk = min(static_cast<int>(fy), kRayInvNd - 2);
                    const float u = fminf(umax, lerpTab(P.ray_inv + m * kRayInvNd, kk, fy - static_cast<float>(kk)) * (kXMaxf * kXMaxf));
                    cost = 1.0f - 2.0f * u / umax;
                    if (2.0f * rngU(rng) <= 1.0f + cost * cost)
                        break;
                }
            }
            rotate(dx, dy, dz, cost, kTwoPi * rngU(rng));
        }
    }

    // ---- wavefront reduction of the counters, one atomic per wavefront
    n_steps = waveSum(n_steps);
    n_real = waveSum(n_real);
    n_dep = waveSum(n_dep);
    n_hist = waveSum(n_hist);
    n_unc = waveSum(n_unc);
    e_emit = waveSum(e_emit);
    e_esc = waveSum(e_esc);
    e_dep = waveSum(e_dep);
    if (lane == 0) {
        atomicAdd(&P.counters->steps, n_steps);
        atomicAdd(&P.counters->real_events, n_real);
        atomicAdd(&P.counters->deposits, n_dep);
        atomicAdd(&P.counters->histories, n_hist);
        atomicAdd(&P.counters->uncollided, n_unc);
        atomicAdd(&P.counters->emitted, e_emit);
        atomicAdd(&P.counters->escaped, e_esc);
        atomicAdd(&P.counters->deposited, e_dep);
    }
}

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
            const float s = steps[r * n_steps + k];
            px = __fmaf_rn(s, dx, px);
            py = __fmaf_rn(s, dy, py);
            pz = __fmaf_rn(s, dz, pz);
            const int v = voxelOf(P, px, py, pz);
            if (v >= 0) {
                vi = v;
                mm = static_cast<int32_t>(P.voxels[v] & 0xFFu);
            }
        }
        out_index[r * n_steps + k] = vi;
        out_material[r * n_steps + k] = mm;
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
    }
}

size_t transportLdsBytes(const DevParams& p)
{
    return sizeof(uint32_t) * (static_cast<size_t>(kLutNd) + (p.lds_lut ? static_cast<size_t>(p.n_mat) * kLutNd : 0)
                               + static_cast<size_t>(p.bb.words));
}

void launchTransport(const DevParams& p, int blocks, int threads, hipStream_t s)
{
    const size_t lds = transportLdsBytes(p);
    if (p.lds_lut)
        hipLaunchKernelGGL(transport_kernel<true>, dim3(static_cast<unsigned>(blocks)), dim3(static_cast<unsigned>(threads)), lds, s, p);
    else
        hipLaunchKernelGGL(transport_kernel<false>, dim3(static_cast<unsigned>(blocks)), dim3(static_cast<unsigned>(threads)), lds, s, p);
}

int transportBlocksPerCU(bool lds_lut, size_t lds_bytes)
{
    int n = 0;
    const void* fn = lds_lut ? reinterpret_cast<const void*>(&transport_kernel<true>) : reinterpret_cast<const void*>(&transport_kernel<false>);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fn, kBlock, lds_bytes) != hipSuccess)
        return 1;
    return n > 0 ? n : 1;
}

void launchTraverse(const DevParams& p, uint64_t n_rays, const float* pos, const float* dir, uint64_t n_steps,
                    const float* steps, int64_t* out_index, int32_t* out_material, hipStream_t s)
{
    const unsigned blocks = static_cast<unsigned>((n_rays + 255) / 256);
    hipLaunchKernelGGL(traverse_kernel, dim3(blocks), dim3(256), 0, s, p, n_rays, pos, dir, n_steps, steps, out_index, out_material);
}

void launchReset(unsigned long long* pool, DevCounters* counters, hipStream_t s)
{
    hipLaunchKernelGGL(reset_kernel, dim3(1), dim3(64), 0, s, pool, counters);
}

int transportBlockSize() { return kBlock; }

}  // namespace dxg

namespace dxg {

__global__ void copy_kernel(const float4* __restrict__ src, float4* __restrict__ dst, uint64_t n)
{
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
        dst[i] = src[i];
}

void launchCopy(const void* src, void* dst, uint64_t bytes, hipStream_t s)
{
    const uint64_t n = bytes / sizeof(float4);
    hipLaunchKernelGGL(copy_kernel, dim3(4096), dim3(256), 0, s, static_cast<const float4*>(src), static_cast<float4*>(dst), n);
}

}  // namespace dxg
