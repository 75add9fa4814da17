// CT number segmentation and the organ dose table (C-ABI in
// include/dxmc_pipelines.h).  Both are single passes over the voxels and
// bound by HBM bandwidth: 17 bytes per voxel each (segmentation reads an 8-B
// CT number and writes a 1-B material index and an 8-B density; the table
// reads a 1-B organ index, an 8-B dose and an 8-B density).
//
// Segmentation: every thread maps voxel pairs (16-B loads and stores, each
// wavefront instruction covering 1 KB of contiguous memory), four pairs in
// flight per thread.  The material and density tables travel as kernel
// arguments.
//
// Organ dose table: each wavefront owns a contiguous run of voxels and walks
// it in groups of 64 (one voxel per lane, coalesced).  Organ indices come in
// long runs, so a wavefront keeps per-lane accumulators for one "current"
// organ and only folds them into its LDS table (256 organs x energy, density
// sum and count) when the organ changes; a group holding several organs is
// split by ballot, one organ at a time, and wave-reduced (lane 0 writes the
// wavefront's own LDS slice, in program order).  A group of more than four
// runs (speckle, not organ boundaries) is added by LDS atomics on the
// wavefront's slice instead.  The work-group adds its four slices in a fixed order
// and a second kernel adds the work-groups' partials in a fixed order.
#include "dxmc_pipelines.h"

#include "physics.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

namespace {

int fail(int code, const std::string& msg) { return dxg::setError(code, msg); }

#define HIP_TRY(expr)                                                                     \
    do {                                                                                  \
        const hipError_t e_ = (expr);                                                     \
        if (e_ != hipSuccess)                                                             \
            return fail(DXG_ERR_DEVICE, std::string(#expr ": ") + hipGetErrorString(e_)); \
    } while (0)

thread_local double g_last_ms[2] = { 0.0, 0.0 };

int computeUnits()
{
    int dev = 0, n = 0;
    if (hipGetDevice(&dev) == hipSuccess && hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess
        && n > 0)
        return n;
    return 256;
}

// The caller's pointer must lie in a device allocation holding `bytes` from it.
int checkDeviceRange(const void* p, size_t bytes, const char* what)
{
    if (bytes == 0)
        return DXG_OK;
    if (!p)
        return fail(DXG_ERR_ARGUMENT, std::string(what) + ": null pointer");
    hipPointerAttribute_t attr {};
    if (hipPointerGetAttributes(&attr, p) != hipSuccess || attr.type != hipMemoryTypeDevice) {
        (void)hipGetLastError();
        return fail(DXG_ERR_ARGUMENT, std::string(what) + ": not device memory");
    }
    hipDeviceptr_t base = nullptr;
    size_t size = 0;
    if (hipMemGetAddressRange(&base, &size, const_cast<void*>(p)) != hipSuccess) {
        (void)hipGetLastError();
        return fail(DXG_ERR_ARGUMENT, std::string(what) + ": unknown device allocation");
    }
    const auto b = reinterpret_cast<uintptr_t>(base), q = reinterpret_cast<uintptr_t>(p);
    if (q < b || q - b + bytes > size)
        return fail(DXG_ERR_ARGUMENT, std::string(what) + ": buffer smaller than the voxel count");
    return DXG_OK;
}

// HIP events around the launches of one call, on the caller's stream.
struct Timer {
    hipEvent_t a = nullptr, b = nullptr;
    ~Timer()
    {
        if (a)
            (void)hipEventDestroy(a);
        if (b)
            (void)hipEventDestroy(b);
    }
    int start(hipStream_t s)
    {
        HIP_TRY(hipEventCreate(&a));
        HIP_TRY(hipEventCreate(&b));
        HIP_TRY(hipEventRecord(a, s));
        return DXG_OK;
    }
    int stop(hipStream_t s, double& ms)
    {
        HIP_TRY(hipEventRecord(b, s));
        HIP_TRY(hipEventSynchronize(b));
        float f = 0.0f;
        HIP_TRY(hipEventElapsedTime(&f, a, b));
        ms = f;
        return DXG_OK;
    }
};

template <typename T>
struct DeviceBuffer {
    T* p = nullptr;
    ~DeviceBuffer()
    {
        if (p)
            (void)hipFree(p);
    }
    int alloc(size_t n)
    {
        HIP_TRY(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
        return DXG_OK;
    }
};

// ---------------------------------------------------------------- segmentation

constexpr int kSegThreads = 256;
constexpr int kSegPairs = 2;    // voxel pairs per thread (full-grid kernel)
constexpr int kSegUnroll = 4;   // voxels per thread, sizing the any-alignment kernel's grid

struct SegArgs {
    double att[DXG_SEG_MAX_MATERIALS];
    double sep[DXG_SEG_MAX_MATERIALS - 1];
    double slope, offset;
    uint32_t nsep;
};

// ctsegmentationpipeline.cpp:136-152, with the reference's evaluation order:
// ((slope * hu) / 1000 + offset) / attenuation, then std::max(d, 0.0), which
// keeps NaN and -0.0.
__device__ inline void segmentOne(const SegArgs& a, double h, uint8_t& m, double& d)
{
#pragma clang fp contract(off)
    uint32_t idx = a.nsep;
    for (int i = static_cast<int>(a.nsep) - 1; i >= 0; --i)
        if (h < a.sep[i])
            idx = static_cast<uint32_t>(i);
    const double v = (a.slope * h / 1000.0 + a.offset) / a.att[idx];
    m = static_cast<uint8_t>(idx);
    d = v < 0.0 ? 0.0 : v;
}

// Voxel pairs 2p, 2p+1 (16-B aligned CT numbers and densities, 2-B aligned
// materials), kSegPairs pairs per thread 256 apart, one launch covering the
// volume; the data is touched once, so loads and stores bypass the caches'
// retention (nontemporal).  An odd last voxel goes to thread 0 of block 0.
__global__ __launch_bounds__(kSegThreads) void segment_pairs_kernel(SegArgs a, const double2* __restrict__ hu, uint64_t n,
                                                                    uchar2* __restrict__ mat, double2* __restrict__ dens)
{
    const uint64_t pairs = n >> 1;
    const uint64_t p0 = static_cast<uint64_t>(blockIdx.x) * (kSegThreads * kSegPairs) + threadIdx.x;
    double2 h[kSegPairs];
#pragma unroll
    for (int u = 0; u < kSegPairs; ++u) {
        const uint64_t p = p0 + u * kSegThreads;
        if (p < pairs) {
            h[u].x = __builtin_nontemporal_load(&hu[p].x);
            h[u].y = __builtin_nontemporal_load(&hu[p].y);
        }
    }
#pragma unroll
    for (int u = 0; u < kSegPairs; ++u) {
        const uint64_t p = p0 + u * kSegThreads;
        if (p < pairs) {
            uint8_t m0, m1;
            double d0, d1;
            segmentOne(a, h[u].x, m0, d0);
            segmentOne(a, h[u].y, m1, d1);
            mat[p] = make_uchar2(m0, m1);
            __builtin_nontemporal_store(d0, &dens[p].x);
            __builtin_nontemporal_store(d1, &dens[p].y);
        }
    }
    if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0) {
        const double* h1 = reinterpret_cast<const double*>(hu);
        uint8_t m;
        double d;
        segmentOne(a, h1[n - 1], m, d);
        reinterpret_cast<uint8_t*>(mat)[n - 1] = m;
        reinterpret_cast<double*>(dens)[n - 1] = d;
    }
}

// Any alignment: one voxel per thread and step.
__global__ __launch_bounds__(kSegThreads) void segment_kernel(SegArgs a, const double* __restrict__ hu, uint64_t n,
                                                              uint8_t* __restrict__ mat, double* __restrict__ dens)
{
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kSegThreads;
    for (uint64_t v = static_cast<uint64_t>(blockIdx.x) * kSegThreads + threadIdx.x; v < n; v += stride) {
        uint8_t m;
        double d;
        segmentOne(a, hu[v], m, d);
        mat[v] = m;
        dens[v] = d;
    }
}

// ---------------------------------------------------------------- organ table

constexpr int kOrgans = 256;
constexpr int kTableWaves = 4;
constexpr int kTableThreads = kTableWaves * 64;
constexpr int kTableGroups = 4;   // groups of 64 voxels loaded per step
constexpr int kTableSplits = 4;   // a group of more runs than this is added by LDS atomics

struct TableSlices {
    double e[kTableWaves][kOrgans];
    double d[kTableWaves][kOrgans];
    unsigned long long n[kTableWaves][kOrgans];
};

// One DPP move of a double inside each row of 16 lanes.
template <int CTRL>
__device__ inline double dppMove(double x)
{
    const unsigned long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_update_dpp(0, static_cast<int>(b & 0xffffffffull), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, static_cast<int>(b >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double((static_cast<unsigned long long>(static_cast<unsigned>(hi)) << 32) | static_cast<unsigned>(lo));
}

__device__ inline double readLane(double x, int lane)
{
    const unsigned long long b = __double_as_longlong(x);
    const unsigned lo = __builtin_amdgcn_readlane(static_cast<int>(b & 0xffffffffull), lane);
    const unsigned hi = __builtin_amdgcn_readlane(static_cast<int>(b >> 32), lane);
    return __longlong_as_double((static_cast<unsigned long long>(hi) << 32) | lo);
}

// Sum over the wavefront, the same value in every lane and a fixed order:
// rotations by 8, 4, 2, 1 inside each row of 16 lanes (DPP row_ror) leave the
// row's sum in each of its lanes; the four row sums are then added in order.
__device__ inline double waveSum(double x)
{
    x += dppMove<0x128>(x);
    x += dppMove<0x124>(x);
    x += dppMove<0x122>(x);
    x += dppMove<0x121>(x);
    return (readLane(x, 0) + readLane(x, 16)) + (readLane(x, 32) + readLane(x, 48));
}

struct TableWave {
    double* se;
    double* sd;
    unsigned long long* sn;
    int lane;
    int cur = -1;               // organ of the per-lane accumulators (wavefront-uniform)
    double ae = 0.0, ad = 0.0;  // per-lane energy and density sums of `cur`
    unsigned long long an = 0;  // voxels of `cur` (wavefront-uniform)

    __device__ void add(int organ, double e, double d, unsigned long long count)
    {
        const double te = waveSum(e), td = waveSum(d);
        if (lane == 0) {
            se[organ] += te;
            sd[organ] += td;
            sn[organ] += count;
        }
    }
    __device__ void flush()
    {
        if (cur >= 0)
            add(cur, ae, ad, an);
        ae = ad = 0.0;
        an = 0;
    }
    // One group of 64 voxels: lanes past the end carry organ kOrgans.
    __device__ void group(int o, double e, double d)
    {
        const unsigned long long valid = __ballot(o < kOrgans);
        if (valid == 0)
            return;
        const int o0 = __builtin_amdgcn_readfirstlane(o);
        const unsigned long long same = __ballot(o == o0);
        if (same == valid) {
            if (o0 != cur) {
                flush();
                cur = o0;
            }
            ae += e;
            ad += d;
            an += __popcll(valid);
            return;
        }
        // runs of one organ along the group: a few (organ boundaries) are
        // split off by ballot; many (speckle) go to LDS atomics
        const int prev = __shfl_up(o, 1, 64);
        const unsigned long long starts = __ballot(lane > 0 && o < kOrgans && o != prev);
        if (__popcll(starts) >= kTableSplits) {
            if (o < kOrgans) {
                atomicAdd(&se[o], e);
                atomicAdd(&sd[o], d);
                atomicAdd(&sn[o], 1ull);
            }
            return;
        }
        // the organ of the last voxel continues into the next group
        const int last = 63 - __clzll(valid);
        const int ol = __builtin_amdgcn_readlane(o, last);
        if (ol != cur) {
            flush();
            cur = ol;
        }
        unsigned long long rem = valid;   // at most kTableSplits organs
        while (rem) {
            const int l = __ffsll(static_cast<long long>(rem)) - 1;
            const int oi = __builtin_amdgcn_readlane(o, l);
            const bool mine = o == oi;
            const unsigned long long m = __ballot(mine);
            const double me = mine ? e : 0.0, md = mine ? d : 0.0;
            if (oi == cur) {
                ae += me;
                ad += md;
                an += __popcll(m);
            } else {
                add(oi, me, md, __popcll(m));
            }
            rem &= ~m;
        }
    }
};

// Per work-group partial sums of energy (dose x voxel volume x density,
// dosetablepipeline.cpp:63-68), density and voxel count, stored organ-major
// ([organ][block]) for the finishing kernel.
__global__ __launch_bounds__(kTableThreads) void organ_table_kernel(const uint8_t* __restrict__ organ, const double* __restrict__ dose,
                                                                    const double* __restrict__ density, uint64_t n, double voxel_volume,
                                                                    uint64_t chunk, double* __restrict__ part_e,
                                                                    double* __restrict__ part_d, unsigned long long* __restrict__ part_n)
{
#pragma clang fp contract(off)
    __shared__ TableSlices s;
    for (int i = threadIdx.x; i < kTableWaves * kOrgans; i += kTableThreads) {
        (&s.e[0][0])[i] = 0.0;
        (&s.d[0][0])[i] = 0.0;
        (&s.n[0][0])[i] = 0;
    }
    __syncthreads();
    const int w = threadIdx.x / 64;
    TableWave t { s.e[w], s.d[w], s.n[w], static_cast<int>(threadIdx.x % 64) };
    const uint64_t start = (static_cast<uint64_t>(blockIdx.x) * kTableWaves + w) * chunk;
    const uint64_t end = std::min(n, start + chunk);
    for (uint64_t base = start; base < end; base += 64 * kTableGroups) {
        int o[kTableGroups];
        double e[kTableGroups], d[kTableGroups];
#pragma unroll
        for (int g = 0; g < kTableGroups; ++g) {
            const uint64_t v = base + g * 64 + t.lane;
            if (v < end) {   // read once: nontemporal
                const double rho = __builtin_nontemporal_load(&density[v]);
                o[g] = __builtin_nontemporal_load(&organ[v]);
                e[g] = __builtin_nontemporal_load(&dose[v]) * (voxel_volume * rho);
                d[g] = rho;
            } else {
                o[g] = kOrgans;
                e[g] = d[g] = 0.0;
            }
        }
#pragma unroll
        for (int g = 0; g < kTableGroups; ++g)
            t.group(o[g], e[g], d[g]);
    }
    t.flush();
    __syncthreads();
    const int k = threadIdx.x;   // organ
    const size_t out = static_cast<size_t>(k) * gridDim.x + blockIdx.x;
    part_e[out] = ((s.e[0][k] + s.e[1][k]) + s.e[2][k]) + s.e[3][k];
    part_d[out] = ((s.d[0][k] + s.d[1][k]) + s.d[2][k]) + s.d[3][k];
    part_n[out] = s.n[0][k] + s.n[1][k] + s.n[2][k] + s.n[3][k];
}

// One work-group per organ adds the partials in a fixed order (strided
// per-thread sums, then a tree in LDS) and forms the row
// (dosetablepipeline.cpp:71-85).
__global__ __launch_bounds__(kTableThreads) void organ_table_finish_kernel(const double* __restrict__ part_e, const double* __restrict__ part_d,
                                                                           const unsigned long long* __restrict__ part_n, int blocks,
                                                                           double voxel_volume, dxg_organ_dose* __restrict__ out)
{
#pragma clang fp contract(off)
    __shared__ double se[kTableThreads], sd[kTableThreads];
    __shared__ unsigned long long sn[kTableThreads];
    const int k = blockIdx.x, j = threadIdx.x;
    const size_t row = static_cast<size_t>(k) * blocks;
    double e = 0.0, d = 0.0;
    unsigned long long c = 0;
    for (int b = j; b < blocks; b += kTableThreads) {
        e += part_e[row + b];
        d += part_d[row + b];
        c += part_n[row + b];
    }
    se[j] = e;
    sd[j] = d;
    sn[j] = c;
    __syncthreads();
    for (int h = kTableThreads / 2; h > 0; h >>= 1) {
        if (j < h) {
            se[j] += se[j + h];
            sd[j] += sd[j + h];
            sn[j] += sn[j + h];
        }
        __syncthreads();
    }
    if (j == 0) {
        dxg_organ_dose r { 0, 0.0, 0.0, 0.0 };
        if (sn[0] > 0) {
            const double mass = voxel_volume * sd[0];
            r.n_voxels = sn[0];
            r.volume_cm3 = static_cast<double>(sn[0]) * voxel_volume;
            r.mass_g = mass;
            r.dose = se[0] / mass;
        }
        out[k] = r;
    }
}

bool validTables(const dxg_ct_segmentation* t)
{
    return t && t->n_materials >= 1 && t->n_materials <= DXG_SEG_MAX_MATERIALS;
}

SegArgs segArgs(const dxg_ct_segmentation& t)
{
    SegArgs a {};
    for (uint32_t i = 0; i < t.n_materials; ++i)
        a.att[i] = t.attenuation[i];
    for (uint32_t i = 0; i + 1 < t.n_materials; ++i)
        a.sep[i] = t.separators[i];
    a.slope = t.slope;
    a.offset = t.offset;
    a.nsep = t.n_materials - 1;
    return a;
}

int runSegmentation(const dxg_ct_segmentation& t, const double* hu, uint64_t n, uint8_t* mat, double* dens, hipStream_t s)
{
    const SegArgs a = segArgs(t);
    const int cus = computeUnits();
    const bool vec = (reinterpret_cast<uintptr_t>(hu) % 16 == 0) && (reinterpret_cast<uintptr_t>(dens) % 16 == 0)
        && (reinterpret_cast<uintptr_t>(mat) % 2 == 0);
    const uint64_t per_block = vec ? kSegThreads * kSegPairs : kSegThreads * kSegUnroll;
    const uint64_t want = std::max<uint64_t>((vec ? n / 2 : n) + per_block - 1, per_block) / per_block;
    if (vec && want > 0x7fffffffull)
        return fail(DXG_ERR_ARGUMENT, "ct segmentation: too many voxels for one launch");
    const int blocks = static_cast<int>(vec ? want : std::min<uint64_t>(want, static_cast<uint64_t>(cus) * 8));
    Timer tm;
    if (int rc = tm.start(s); rc != DXG_OK)
        return rc;
    if (vec)
        segment_pairs_kernel<<<blocks, kSegThreads, 0, s>>>(a, reinterpret_cast<const double2*>(hu), n, reinterpret_cast<uchar2*>(mat),
                                                            reinterpret_cast<double2*>(dens));
    else
        segment_kernel<<<blocks, kSegThreads, 0, s>>>(a, hu, n, mat, dens);
    HIP_TRY(hipGetLastError());
    return tm.stop(s, g_last_ms[0]);
}

// Partial sums and result rows of the dose table: one device buffer per
// calling thread and device, grown on demand and kept until the process ends
// (1536 work-groups need 9.4 MB).  A call finishes its work on the buffer
// before it returns, so calls from one thread never overlap on it.
int tableWorkspace(size_t bytes, char*& out)
{
    struct Workspace {
        char* p = nullptr;
        size_t cap = 0;
    };
    thread_local std::vector<Workspace> per_device;
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (static_cast<size_t>(dev) >= per_device.size())
        per_device.resize(dev + 1);
    Workspace& w = per_device[dev];
    if (w.cap < bytes) {
        if (w.p)
            HIP_TRY(hipFree(w.p));
        w.p = nullptr;
        w.cap = 0;
        HIP_TRY(hipMalloc(&w.p, bytes));
        w.cap = bytes;
    }
    out = w.p;
    return DXG_OK;
}

int runTable(const uint8_t* organ, const double* dose, const double* density, uint64_t n, const double spacing[3],
             dxg_organ_dose* out_host, hipStream_t s)
{
    const double vv = spacing[0] * spacing[1] * spacing[2];
    const int cus = computeUnits();
    const uint64_t per_block = kTableWaves * 64ull * 16;   // at least 16 groups per wavefront
    const int blocks = static_cast<int>(std::clamp<uint64_t>((n + per_block - 1) / per_block, 1, static_cast<uint64_t>(cus) * 6));
    const uint64_t waves = static_cast<uint64_t>(blocks) * kTableWaves;
    const uint64_t chunk = ((n + waves - 1) / waves + 63) / 64 * 64;
    const size_t np = static_cast<size_t>(blocks) * kOrgans;
    char* ws = nullptr;
    if (int rc = tableWorkspace(3 * np * sizeof(double) + kOrgans * sizeof(dxg_organ_dose), ws); rc != DXG_OK)
        return rc;
    double* pe = reinterpret_cast<double*>(ws);
    double* pd = pe + np;
    auto* pn = reinterpret_cast<unsigned long long*>(pd + np);
    auto* rows = reinterpret_cast<dxg_organ_dose*>(pn + np);
    Timer tm;
    if (int rc = tm.start(s); rc != DXG_OK)
        return rc;
    organ_table_kernel<<<blocks, kTableThreads, 0, s>>>(organ, dose, density, n, vv, chunk, pe, pd, pn);
    HIP_TRY(hipGetLastError());
    organ_table_finish_kernel<<<kOrgans, kTableThreads, 0, s>>>(pe, pd, pn, blocks, vv, rows);
    HIP_TRY(hipGetLastError());
    if (int rc = tm.stop(s, g_last_ms[1]); rc != DXG_OK)
        return rc;
    HIP_TRY(hipMemcpyAsync(out_host, rows, kOrgans * sizeof(dxg_organ_dose), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return DXG_OK;
}

bool validSpacing(const double* sp)
{
    return sp && sp[0] > 0.0 && sp[1] > 0.0 && sp[2] > 0.0 && std::isfinite(sp[0] * sp[1] * sp[2]);
}

// Spectrum average <f(E)> = sum_i w_i f(E_i) of a material's total mu/rho,
// and the per-bin values (cm2/g).
int spectrumAttenuation(const dxg_material* m, const std::vector<double>& e, std::vector<double>& mu)
{
    std::vector<double> parts(4 * e.size());
    if (int rc = dxg_material_attenuation(m, e.data(), e.size(), parts.data()); rc != DXG_OK)
        return rc;
    mu.resize(e.size());
    for (size_t i = 0; i < e.size(); ++i)
        mu[i] = parts[4 * i] + parts[4 * i + 1] + parts[4 * i + 2];
    return DXG_OK;
}

}  // namespace

extern "C" {

int dxg_ct_segmentation_tables(const dxg_tube* tube, const dxg_material* materials, const double* densities, uint64_t n_materials,
                               const dxg_material* water, double water_density, const dxg_material* air, double air_density,
                               dxg_ct_segmentation* out)
{
    if (!tube || !materials || !densities || !water || !air || !out)
        return fail(DXG_ERR_ARGUMENT, "dxg_ct_segmentation_tables: null argument");
    if (n_materials < 1 || n_materials > DXG_SEG_MAX_MATERIALS)
        return fail(DXG_ERR_ARGUMENT, "dxg_ct_segmentation_tables: between 1 and " + std::to_string(DXG_SEG_MAX_MATERIALS) + " materials");
    if (!(water_density > 0.0) || !(air_density > 0.0))
        return fail(DXG_ERR_ARGUMENT, "dxg_ct_segmentation_tables: densities must be positive");
    for (uint64_t i = 0; i < n_materials; ++i)
        if (!(densities[i] > 0.0))
            return fail(DXG_ERR_ARGUMENT, "dxg_ct_segmentation_tables: densities must be positive");
    uint64_t nb = 0;
    if (int rc = dxg_tube_spectrum(tube, nullptr, nullptr, 0, &nb); rc != DXG_OK)
        return rc;
    std::vector<double> e(nb), w(nb);
    if (int rc = dxg_tube_spectrum(tube, e.data(), w.data(), nb, &nb); rc != DXG_OK)
        return rc;
    std::vector<double> mw, ma, mm;
    if (int rc = spectrumAttenuation(water, e, mw); rc != DXG_OK)
        return rc;
    if (int rc = spectrumAttenuation(air, e, ma); rc != DXG_OK)
        return rc;

    dxg_ct_segmentation t {};
    t.n_materials = static_cast<uint32_t>(n_materials);
    // getMeanCTNumbers (ctsegmentationpipeline.cpp:78-105): the CT number is
    // the spectrum average of the per-energy ratio; the attenuations are
    // spectrum averages of mu/rho.
    for (uint64_t m = 0; m < n_materials; ++m) {
        if (int rc = spectrumAttenuation(&materials[m], e, mm); rc != DXG_OK)
            return rc;
        double hu = 0.0, att = 0.0;
        for (size_t i = 0; i < nb; ++i) {
            const double uw = mw[i] * water_density;
            const double ua = ma[i] * air_density;
            const double um = mm[i] * densities[m];
            hu += w[i] * (um - uw) / (uw - ua);
            att += w[i] * mm[i];
        }
        t.hu[m] = 1000.0 * hu;
        t.attenuation[m] = att;
    }
    double aw = 0.0, aa = 0.0;
    for (size_t i = 0; i < nb; ++i) {
        aw += w[i] * mw[i];
        aa += w[i] * ma[i];
    }
    t.water_attenuation = aw;
    t.air_attenuation = aa;
    t.water_density = water_density;
    t.air_density = air_density;
    for (uint64_t m = 0; m + 1 < n_materials; ++m)
        t.separators[m] = (t.hu[m] + t.hu[m + 1]) / 2.0;
    t.slope = aw * water_density - aa * air_density;
    t.offset = aw * water_density;
    *out = t;
    return DXG_OK;
}

int dxg_ct_segment_device(const dxg_ct_segmentation* t, const double* hu, uint64_t n, uint8_t* material, double* density, void* stream)
{
    if (!validTables(t))
        return fail(DXG_ERR_ARGUMENT, "dxg_ct_segment_device: invalid tables");
    if (n == 0)
        return DXG_OK;
    if (int rc = checkDeviceRange(hu, n * sizeof(double), "dxg_ct_segment_device: hu"); rc != DXG_OK)
        return rc;
    if (int rc = checkDeviceRange(material, n, "dxg_ct_segment_device: material"); rc != DXG_OK)
        return rc;
    if (int rc = checkDeviceRange(density, n * sizeof(double), "dxg_ct_segment_device: density"); rc != DXG_OK)
        return rc;
    return runSegmentation(*t, hu, n, material, density, static_cast<hipStream_t>(stream));
}

int dxg_ct_segment(const dxg_ct_segmentation* t, const double* hu, uint64_t n, uint8_t* material, double* density)
{
    if (!validTables(t))
        return fail(DXG_ERR_ARGUMENT, "dxg_ct_segment: invalid tables");
    if (n == 0)
        return DXG_OK;
    if (!hu || !material || !density)
        return fail(DXG_ERR_ARGUMENT, "dxg_ct_segment: null argument");
    DeviceBuffer<double> dh, dd;
    DeviceBuffer<uint8_t> dm;
    if (int rc = dh.alloc(n); rc != DXG_OK)
        return rc;
    if (int rc = dd.alloc(n); rc != DXG_OK)
        return rc;
    if (int rc = dm.alloc(n); rc != DXG_OK)
        return rc;
    HIP_TRY(hipMemcpy(dh.p, hu, n * sizeof(double), hipMemcpyHostToDevice));
    if (int rc = runSegmentation(*t, dh.p, n, dm.p, dd.p, nullptr); rc != DXG_OK)
        return rc;
    HIP_TRY(hipMemcpy(material, dm.p, n, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(density, dd.p, n * sizeof(double), hipMemcpyDeviceToHost));
    return DXG_OK;
}

int dxg_organ_dose_table_device(const uint8_t* organ, const double* dose, const double* density, uint64_t n, const double spacing_cm[3],
                                dxg_organ_dose out[256], void* stream)
{
    if (!out || !validSpacing(spacing_cm))
        return fail(DXG_ERR_ARGUMENT, "dxg_organ_dose_table_device: null output or invalid spacing");
    if (n == 0) {
        std::fill(out, out + kOrgans, dxg_organ_dose { 0, 0.0, 0.0, 0.0 });
        return DXG_OK;
    }
    if (int rc = checkDeviceRange(organ, n, "dxg_organ_dose_table_device: organ"); rc != DXG_OK)
        return rc;
    if (int rc = checkDeviceRange(dose, n * sizeof(double), "dxg_organ_dose_table_device: dose"); rc != DXG_OK)
        return rc;
    if (int rc = checkDeviceRange(density, n * sizeof(double), "dxg_organ_dose_table_device: density"); rc != DXG_OK)
        return rc;
    return runTable(organ, dose, density, n, spacing_cm, out, static_cast<hipStream_t>(stream));
}

int dxg_organ_dose_table(const uint8_t* organ, const double* dose, const double* density, uint64_t n, const double spacing_cm[3],
                         dxg_organ_dose out[256])
{
    if (!out || !validSpacing(spacing_cm))
        return fail(DXG_ERR_ARGUMENT, "dxg_organ_dose_table: null output or invalid spacing");
    if (n == 0) {
        std::fill(out, out + kOrgans, dxg_organ_dose { 0, 0.0, 0.0, 0.0 });
        return DXG_OK;
    }
    if (!organ || !dose || !density)
        return fail(DXG_ERR_ARGUMENT, "dxg_organ_dose_table: null argument");
    DeviceBuffer<uint8_t> dorg;
    DeviceBuffer<double> ddose, ddens;
    if (int rc = dorg.alloc(n); rc != DXG_OK)
        return rc;
    if (int rc = ddose.alloc(n); rc != DXG_OK)
        return rc;
    if (int rc = ddens.alloc(n); rc != DXG_OK)
        return rc;
    HIP_TRY(hipMemcpy(dorg.p, organ, n, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(ddose.p, dose, n * sizeof(double), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(ddens.p, density, n * sizeof(double), hipMemcpyHostToDevice));
    return runTable(dorg.p, ddose.p, ddens.p, n, spacing_cm, out, nullptr);
}

double dxg_pipeline_last_ms(int which)
{
    return which == 0 || which == 1 ? g_last_ms[which] : 0.0;
}

}  // extern "C"
