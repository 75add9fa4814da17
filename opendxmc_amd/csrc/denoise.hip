// Dose denoiser: BF16 implicit-GEMM 3-D convolutions on CDNA4 matrix cores,
// FP32 accumulation, hand-written backward pass and Adam (C-ABI in
// include/dxmc_denoise.h).
//
// Tensors are channel-last BF16 with 32 channels: voxel v = ((n*D + z)*H + y)*W + x
// occupies 64 bytes.  Every layer is a 3x3x3 convolution 32 -> 32 (the input
// features and the one-channel output are zero padded to 32), so one kernel
// serves all of them.
//
// Matrix-core operand layout (v_mfma_f32_16x16x32_bf16, D = A*B, A 16x32,
// B 32x16): lane l holds A[l%16][8*(l/16)+e] and B[8*(l/16)+e][l%16] in element
// e of its 8-element operands and receives D[4*(l/16)+i][l%16] in accumulator i.
// dxg_mfma_probe checks exactly this on the device.
//
// Convolution (forward, and input gradient with flipped transposed weights):
// A = weights (16 output channels x 32 input channels of one tap), B = 32
// input channels x 16 voxels along x.  A wavefront owns 64 consecutive voxels
// of one row and both output-channel halves: per tap 2 weight fragments from
// LDS, 4 activation fragments (16 B per lane, coalesced 1 KB per instruction)
// from memory, 8 MFMAs; 27 taps accumulate in 32 FP32 registers.
//
// Weight gradient: dW[co][tap][ci] = sum_v g[v][co] x[v+off(tap)][ci] reduces
// over voxels, so K runs over 32 voxels along x.  A work-group transposes one
// 4-row x 32-voxel tile of g and its 18-row halo of x into channel-major LDS;
// each wavefront keeps the partial sums of 7 taps in registers and reads the
// x-shifted operand with two aligned 16-B reads and a byte funnel shift.
#include "dxmc_denoise.h"

#include "device.hpp"
#include "physics.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <random>
#include <string>
#include <vector>

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned u32x2 __attribute__((ext_vector_type(2)));

constexpr int kC = DXG_DENOISE_CHANNELS;            // 32
constexpr int kTaps = 27;
constexpr int kLayerW = kC * kTaps * kC;             // 27648 weights per layer
constexpr int kLayerP = kLayerW + kC;                // + 32 biases
constexpr int kLayers = DXG_DENOISE_LAYERS;          // 18
constexpr int kBlocks = (kLayers - 2) / 2;           // 8 residual blocks
constexpr int kXtLen = 56;                           // halo row length in LDS (112 B: conflict-free 16-B reads)
constexpr int kGtLen = 40;                           // gradient row length in LDS (80 B)

int fail(int code, const std::string& msg) { return dxg::setError(code, msg); }

#define HIP_TRY(expr)                                                                     \
    do {                                                                                  \
        const hipError_t e_ = (expr);                                                     \
        if (e_ != hipSuccess)                                                             \
            return fail(DXG_ERR_DEVICE, std::string(#expr ": ") + hipGetErrorString(e_)); \
    } while (0)

// ---------------------------------------------------------------- BF16 helpers

__host__ __device__ inline uint16_t toBf16(float f)
{
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u)
        return static_cast<uint16_t>((u >> 16) | 0x40u);   // NaN stays NaN
    u += 0x7fffu + ((u >> 16) & 1u);                       // round to nearest even
    return static_cast<uint16_t>(u >> 16);
}

__host__ __device__ inline float fromBf16(uint32_t h)
{
    const uint32_t u = h << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

// Two floats to BF16 (round to nearest even) in one v_cvt_pk_bf16_f32; the
// same rounding as toBf16 (a NaN comes out quiet).
__device__ inline uint32_t pack2(float lo, float hi)
{
    typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
    const bf16x2 v = { static_cast<__bf16>(lo), static_cast<__bf16>(hi) };
    return __builtin_bit_cast(uint32_t, v);
}

__device__ inline bf16x8 asOperand(u32x4 v) { return __builtin_bit_cast(bf16x8, v); }

__device__ inline u32x4 load16(const uint16_t* p) { return *reinterpret_cast<const u32x4*>(p); }

// A row of 64 voxels x 32 channels of zeros with one voxel of zeros on each
// side (never written): the convolution's stand-in for input rows and halo
// voxels outside the volume.  Kernels get the address of its voxel 0 (one
// voxel in), so the halo voxels -1 and 64 of the zero row are zeros too.
__device__ __align__(256) uint16_t kZeroRow[66 * 32];

// ---------------------------------------------------------------- kernels

__global__ __launch_bounds__(64) void mfma_probe_kernel(const float* a, const float* b, float* d)
{
    const int l = threadIdx.x, lr = l & 15, lq = l >> 4;
    u32x4 av, bv;
    for (int e = 0; e < 4; ++e) {
        const int k = 8 * lq + 2 * e;
        av[e] = pack2(a[lr * 32 + k], a[lr * 32 + k + 1]);
        bv[e] = pack2(b[k * 16 + lr], b[(k + 1) * 16 + lr]);
    }
    const f32x4 acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(asOperand(av), asOperand(bv), f32x4 { 0, 0, 0, 0 }, 0, 0, 0);
    for (int i = 0; i < 4; ++i)
        d[(4 * lq + i) * 16 + lr] = acc[i];
}

struct ConvArgs {
    const uint16_t* x;       // input tensor
    const uint16_t* w;       // [27][32 out][32 in] BF16
    const float* bias;       // 32 or null
    const uint16_t* res;     // residual added before the activation, or null (may alias y)
    const uint16_t* gate;    // output multiplied by (gate > 0), or null
    uint16_t* y;             // BF16 output or null
    float* y32;              // FP32 output or null
    int y32_channels;        // 1 or 32
    int relu;
    int nb, nz, ny, nx;
    int wave_tiles;          // tiles (or row pairs) to compute: nb*nz*ny*(nx/64), or the length of `tiles`
    int chunk;               // work-group tiles (4 wave tiles) per work-group
    const int* tiles;        // null: every tile; else the (ascending) tile indices to compute (pair entries: conv3d_pair_kernel)
    int in4;                 // the network's first layer: x is the 4-channel compact input (8 bytes per voxel), conv3d_first_kernel
    const uint16_t* zrow;    // voxel 0 of kZeroRow on the device (set by launchConv)
};

// Weight fragment (tap, out channel co, 16-B chunk q) lives at chunk
// position q ^ ((co >> 2) & 3) of its 64-B row: 16 lanes reading 16
// consecutive rows then touch 16 distinct bank groups.
__device__ inline int wSlot(int tap, int co, int q) { return (tap * 32 + co) * 4 + (q ^ ((co >> 2) & 3)); }

// DPP row moves (16-lane rows) on the four dwords of an operand fragment;
// lanes whose source falls outside their row keep `old`.
constexpr int kRowShl1 = 0x101, kRowShl15 = 0x10F, kRowShr1 = 0x111, kRowShr15 = 0x11F;

template <int CTRL>
__device__ inline u32x4 dpp4(u32x4 old, u32x4 src)
{
    u32x4 r;
#pragma unroll
    for (int e = 0; e < 4; ++e)
        r[e] = static_cast<unsigned>(__builtin_amdgcn_update_dpp(static_cast<int>(old[e]), static_cast<int>(src[e]), CTRL, 0xF, 0xF, false));
    return r;
}

// Wavefronts per convolution work-group; the layer's weights (55 KB of LDS)
// allow two work-groups per CU.  Four (two wavefronts per SIMD) measured
// faster than six (three per SIMD): 539 against 451 TFLOP/s over the 18
// layers of a 128x256x256 volume.
#ifndef DXG_CONV_WAVES
#define DXG_CONV_WAVES 4
#endif
constexpr int kConvWaves = DXG_CONV_WAVES;
// Output rows per wavefront tile of conv3d_rows_kernel.
#ifndef DXG_CONV_ROWS
#define DXG_CONV_ROWS 2
#endif
constexpr int kConvRows = DXG_CONV_ROWS;
// Input-row buffers of conv3d_rows_kernel: rows load kConvRing - 1 rows ahead of their MFMAs.
#ifndef DXG_CONV_RING
#define DXG_CONV_RING 3
#endif
constexpr int kConvRing = DXG_CONV_RING;
// One wavefront tile: 64 voxels of one x row.
struct Tile {
    int x0, y, z, n;
    bool valid;
};
// A row-pair tile with what its input rows need: the address of voxel x0 of
// row y, plane z; bit q of `rows` set when input row q lies inside the
// volume; whether the halo voxels x0 - 1 and x0 + 64 exist.
struct RowsTile : Tile {
    const uint16_t* own;
    uint32_t rows;
    bool left, right;
};

template <bool SHIFT>
__global__ __launch_bounds__(64 * kConvWaves) void conv3d_kernel(ConvArgs a)
{
    __shared__ u32x4 wl[kTaps * 32 * 4];   // 55,296 B
    const int tid = threadIdx.x;
    for (int i = tid; i < kTaps * 32 * 4; i += 64 * kConvWaves) {
        const int row = i >> 2;
        wl[wSlot(row >> 5, row & 31, i & 3)] = reinterpret_cast<const u32x4*>(a.w)[i];
    }
    __syncthreads();

    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63, lr = lane & 15, lq = lane >> 4;
    float bias[2][4];
    for (int m = 0; m < 2; ++m)
        for (int i = 0; i < 4; ++i)
            bias[m][i] = a.bias ? a.bias[16 * m + 4 * lq + i] : 0.0f;

    // XCD-aware order: consecutive work-group ids alternate between the 8
    // XCDs, so map them to contiguous slabs of rows per XCD.
    const int g = gridDim.x;
    const int logical = (blockIdx.x & 7) * (g >> 3) + (blockIdx.x >> 3);
    const int wg_tiles = (a.wave_tiles + kConvWaves - 1) / kConvWaves;
    const int begin = logical * a.chunk;
    const int end = min(begin + a.chunk, wg_tiles);
    const int nxs = a.nx >> 6;
    auto tileAt = [&](int wt) {
        Tile tc { 0, 0, 0, 0, false };
        const int k = wt * kConvWaves + wave;
        if (wt < end && k < a.wave_tiles) {
            const int t = a.tiles ? a.tiles[k] : k;
            int r = t / nxs;
            tc.x0 = (t - r * nxs) * 64;
            tc.y = r % a.ny;
            r /= a.ny;
            tc.z = r % a.nz;
            tc.n = r / a.nz;
            tc.valid = true;
        }
        return tc;
    };

    // Input rows (SHIFT): the nine (dz, dy) rows of a tile in turn, each
    // loaded once.  Tile j, lane n holds voxel x0 + 4n + j (the four tiles
    // interleave), so the x-neighbour operands of tiles 1-3 / 0-2 are simply
    // the neighbouring tile's registers; only tile 0 at dx = -1 and tile 3 at
    // dx = +1 take their voxel from the next lane of the same 16-lane row (one
    // DPP move per register), with the halo voxel x0 - 1 / x0 + 64 in the lane
    // that falls off the row.
    // Rows stream through a ring of three buffers two rows ahead of the
    // MFMAs, across tiles: the next tile's first two rows load during this
    // tile's last two, so a tile does not start by waiting for its first row.
    // Nine rows per tile keep row q of every tile in buffer q % 3.  Every row
    // issues the same six loads (four operand chunks and the two halo voxels)
    // whatever the position: rows outside the volume, halo voxels beyond its x
    // ends and rows of a missing next tile read a clamped in-range address and
    // are dropped afterwards (wave-uniform tests).  The fixed count lets the
    // wait before a row's MFMAs cover that row only.
    constexpr int kRing = 3;
    const u32x4 zero = { 0, 0, 0, 0 };
    u32x4 buf[kRing][4], lbuf[kRing], rbuf[kRing];
    bool have[kRing] = { false, false, false };
    auto loadRow = [&](const Tile& tc, int q, int b) {
        const int zz = tc.z + q / 3 - 1, yy = tc.y + q % 3 - 1;
        have[b] = tc.valid && static_cast<unsigned>(zz) < static_cast<unsigned>(a.nz) && static_cast<unsigned>(yy) < static_cast<unsigned>(a.ny);
        const int zc = min(max(zz, 0), a.nz - 1), yc = min(max(yy, 0), a.ny - 1);
        const size_t rowb = (static_cast<size_t>(tc.n * a.nz + zc) * a.ny + yc) * a.nx;
        const uint16_t* base = a.x + (rowb + tc.x0 + 4 * lr) * kC + 8 * lq;
#pragma unroll
        for (int j = 0; j < 4; ++j)
            buf[b][j] = load16(base + j * kC);
        lbuf[b] = load16(a.x + (rowb + max(tc.x0 - 1, 0)) * kC + 8 * lq);
        rbuf[b] = load16(a.x + (rowb + min(tc.x0 + 64, a.nx - 1)) * kC + 8 * lq);
    };

    Tile cur = tileAt(begin);
    if (SHIFT && cur.valid) {
        loadRow(cur, 0, 0);
        loadRow(cur, 1, 1);
    }
    for (int wt = begin; cur.valid; ++wt) {
        const Tile nxt = tileAt(wt + 1);
        const int x0 = cur.x0, y = cur.y, z = cur.z, n = cur.n;

        f32x4 acc[2][4];
        for (int m = 0; m < 2; ++m)
            for (int j = 0; j < 4; ++j)
                acc[m][j] = f32x4 { 0, 0, 0, 0 };

        // The epilogue's residual (or gate) values, loaded before the MFMAs
        // so that their latency hides behind them; loaded after them, each
        // store's possible alias with the next load serialised the waits.
        // The residual may alias the output: each tile reads and writes only
        // its own voxels.
        const size_t orow = (static_cast<size_t>(n * a.nz + z) * a.ny + y) * a.nx;
        u32x2 side[4][2];
        const uint16_t* side_src = a.res ? a.res : a.gate;
        if (side_src) {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const size_t v = orow + x0 + (SHIFT ? 4 * lr + j : 16 * j + lr);
#pragma unroll
                for (int m = 0; m < 2; ++m)
                    side[j][m] = *reinterpret_cast<const u32x2*>(side_src + v * kC + 16 * m + 4 * lq);
            }
        }

        if (SHIFT) {
#pragma unroll
            for (int q = 0; q < 9; ++q) {
                const int b = q % kRing;
                if (q + 2 < 9)
                    loadRow(cur, q + 2, (q + 2) % kRing);
                else
                    loadRow(nxt, q + 2 - 9, (q + 2) % kRing);
                if (!have[b])
                    continue;
                const u32x4(&c)[4] = buf[b];
#pragma unroll
                for (int dx = -1; dx <= 1; ++dx) {
                    bf16x8 bop[4];
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        if (dx == 0)
                            bop[j] = asOperand(c[j]);
                        else if (dx < 0)
                            bop[j] = asOperand(j > 0 ? c[j - 1] : dpp4<kRowShr1>(x0 > 0 ? lbuf[b] : zero, c[3]));   // lane n <- lane n-1
                        else
                            bop[j] = asOperand(j < 3 ? c[j + 1] : dpp4<kRowShl1>(x0 + 64 < a.nx ? rbuf[b] : zero, c[0]));   // lane n <- lane n+1
                    }
                    const int tap = q * 3 + dx + 1;
                    bf16x8 aop[2];
#pragma unroll
                    for (int m = 0; m < 2; ++m)
                        aop[m] = asOperand(wl[wSlot(tap, 16 * m + lr, lq)]);
#pragma unroll
                    for (int m = 0; m < 2; ++m)
#pragma unroll
                        for (int j = 0; j < 4; ++j)
                            acc[m][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(aop[m], bop[j], acc[m][j], 0, 0, 0);
                }
            }
        }
        for (int dz = -1; dz <= 1 && !SHIFT; ++dz) {
            const int zz = z + dz;
            if (static_cast<unsigned>(zz) >= static_cast<unsigned>(a.nz))
                continue;
            for (int dy = -1; dy <= 1; ++dy) {
                const int yy = y + dy;
                if (static_cast<unsigned>(yy) >= static_cast<unsigned>(a.ny))
                    continue;
                const size_t rowb = (static_cast<size_t>(n * a.nz + zz) * a.ny + yy) * a.nx;
                const int tapb = (dz + 1) * 9 + (dy + 1) * 3;
#pragma unroll
                for (int dx = -1; dx <= 1; ++dx) {
                    bf16x8 bop[4];
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const int xx = x0 + 16 * j + lr + dx;
                        u32x4 v = { 0, 0, 0, 0 };
                        if (static_cast<unsigned>(xx) < static_cast<unsigned>(a.nx))
                            v = load16(a.x + (rowb + xx) * kC + 8 * lq);
                        bop[j] = asOperand(v);
                    }
                    const int tap = tapb + dx + 1;
                    bf16x8 aop[2];
#pragma unroll
                    for (int m = 0; m < 2; ++m)
                        aop[m] = asOperand(wl[wSlot(tap, 16 * m + lr, lq)]);
#pragma unroll
                    for (int m = 0; m < 2; ++m)
#pragma unroll
                        for (int j = 0; j < 4; ++j)
                            acc[m][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(aop[m], bop[j], acc[m][j], 0, 0, 0);
                }
            }
        }

#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const size_t v = orow + x0 + (SHIFT ? 4 * lr + j : 16 * j + lr);
#pragma unroll
            for (int m = 0; m < 2; ++m) {
                const int co0 = 16 * m + 4 * lq;
                float o[4];
                for (int i = 0; i < 4; ++i)
                    o[i] = acc[m][j][i] + bias[m][i];
                if (a.res) {
                    const u32x2 rv = side[j][m];
                    o[0] += fromBf16(rv[0] & 0xffffu);
                    o[1] += fromBf16(rv[0] >> 16);
                    o[2] += fromBf16(rv[1] & 0xffffu);
                    o[3] += fromBf16(rv[1] >> 16);
                }
                if (a.relu)
                    for (int i = 0; i < 4; ++i)
                        o[i] = fmaxf(o[i], 0.0f);
                if (a.gate) {
                    const u32x2 gv = a.res ? *reinterpret_cast<const u32x2*>(a.gate + v * kC + co0) : side[j][m];
                    const uint32_t h[4] = { gv[0] & 0xffffu, gv[0] >> 16, gv[1] & 0xffffu, gv[1] >> 16 };
                    for (int i = 0; i < 4; ++i)
                        if (h[i] == 0 || (h[i] & 0x8000u))
                            o[i] = 0.0f;
                }
                if (a.y)
                    *reinterpret_cast<u32x2*>(a.y + v * kC + co0) = u32x2 { pack2(o[0], o[1]), pack2(o[2], o[3]) };
                if (a.y32) {
                    if (a.y32_channels == kC)
                        *reinterpret_cast<f32x4*>(a.y32 + v * kC + co0) = f32x4 { o[0], o[1], o[2], o[3] };
                    else if (co0 == 0)
                        a.y32[v] = o[0];
                }
            }
        }
        cur = nxt;
    }
}


// Row-pair convolution: a wavefront computes output rows y and y + 1 (64
// voxels each, same x range) together.  Their 3x3 input-row neighbourhoods
// overlap in six of nine rows per pair of rows: the pair streams the twelve
// rows (dz, y - 1 .. y + 2) once and applies each to both outputs (rows y and
// y + 1 feed six taps), so it loads 12 rows where two single-row tiles load
// 18; the input loads, not the MFMAs or the weight reads, bound the
// single-row kernel.  Same operand layout and ring as conv3d_kernel<true>;
// twelve rows per pair keep row q of every pair in buffer q % 3.
// Entries of a.tiles (when given) are pair * 4 + mask: bit 0 stores row y,
// bit 1 row y + 1 (a pair index counts (n, z, y / 2, x / 64) like a tile
// index with ceil(ny / 2) rows); without a list every pair is computed and
// row y + 1 is stored when it exists.
// The epilogue's configuration as template bits (EPI >= 0), so that a layer's
// code holds only its own epilogue: the runtime tests of the residual, gate,
// activation and output kind cost instructions (and branches) on every store,
// and the wavefront's instruction fetch is what the convolution waits on.
// EPI < 0 reads the configuration from the arguments.
constexpr int kEpiRelu = 1, kEpiRes = 2, kEpiGate = 4, kEpiY = 8, kEpiY32 = 16, kEpiY32All = 32;

template <int EPI>
__device__ inline bool epiHas(int flag, bool runtime)
{
    return EPI < 0 ? runtime : (EPI & flag) != 0;
}

template <int R, int EPI>
__global__ __launch_bounds__(64 * kConvWaves) __attribute__((amdgpu_waves_per_eu(kConvWaves / 2))) void conv3d_rows_kernel(ConvArgs a)
{
    const bool e_relu = epiHas<EPI>(kEpiRelu, a.relu != 0);
    const bool e_res = epiHas<EPI>(kEpiRes, a.res != nullptr);
    const bool e_gate = epiHas<EPI>(kEpiGate, a.gate != nullptr);
    const bool e_y = epiHas<EPI>(kEpiY, a.y != nullptr);
    const bool e_y32 = epiHas<EPI>(kEpiY32 | kEpiY32All, a.y32 != nullptr);
    const bool e_y32_all = epiHas<EPI>(kEpiY32All, a.y32_channels == kC);
    // the single-channel FP32 output (the last layer) needs output channels
    // 0-15 only: the second channel half's MFMAs and weight reads are skipped
    constexpr int kHalves = EPI == kEpiY32 ? 1 : 2;
    __shared__ u32x4 wl[kTaps * 32 * 4];   // 55,296 B
    const int tid = threadIdx.x;
    for (int i = tid; i < kTaps * 32 * 4; i += 64 * kConvWaves) {
        const int row = i >> 2;
        wl[wSlot(row >> 5, row & 31, i & 3)] = reinterpret_cast<const u32x4*>(a.w)[i];
    }
    __syncthreads();

    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63, lr = lane & 15, lq = lane >> 4;
    float bias[2][4];
    for (int m = 0; m < 2; ++m)
        for (int i = 0; i < 4; ++i)
            bias[m][i] = a.bias ? a.bias[16 * m + 4 * lq + i] : 0.0f;

    // XCD-aware order: consecutive work-group ids alternate between the 8
    // XCDs, so map them to contiguous slabs of pairs per XCD.
    const int g = gridDim.x;
    const int logical = (blockIdx.x & 7) * (g >> 3) + (blockIdx.x >> 3);
    const int wg_tiles = (a.wave_tiles + kConvWaves - 1) / kConvWaves;
    const int begin = logical * a.chunk;
    const int end = min(begin + a.chunk, wg_tiles);
    const int nxs = a.nx >> 6, nyp = (a.ny + R - 1) / R;
    const ptrdiff_t row_stride = static_cast<ptrdiff_t>(a.nx) * kC, plane_stride = row_stride * a.ny;
    const uint16_t* zrow = a.zrow;
    auto tileAt = [&](int wt, int& mask) {
        RowsTile tc {};
        tc.own = zrow;
        mask = 0;
        const int k = wt * kConvWaves + wave;
        if (wt < end && k < a.wave_tiles) {
            int t = k;
            if (a.tiles) {
                t = a.tiles[k] >> R;
                mask = a.tiles[k] & ((1 << R) - 1);
            }
            int r = t / nxs;
            tc.x0 = (t - r * nxs) * 64;
            tc.y = R * (r % nyp);
            r /= nyp;
            tc.z = r % a.nz;
            tc.n = r / a.nz;
            tc.valid = true;
            if (!a.tiles)
                mask = (1 << min(R, a.ny - tc.y)) - 1;
            tc.own = a.x + ((static_cast<size_t>(tc.n * a.nz + tc.z) * a.ny + tc.y) * a.nx + tc.x0) * kC;
            // bit q: input row y + q % (R + 2) - 1 of plane z + q / (R + 2) - 1 lies inside
            uint32_t ym = (1u << (R + 2)) - 1;
            if (tc.y == 0)
                ym &= ~1u;
            const int below_top = a.ny - tc.y + 1;
            if (below_top < R + 2)
                ym &= (1u << below_top) - 1;
            tc.rows = ym << (R + 2);
            if (tc.z > 0)
                tc.rows |= ym;
            if (tc.z + 1 < a.nz)
                tc.rows |= ym << (2 * (R + 2));
            tc.left = tc.x0 > 0;
            tc.right = tc.x0 + 64 < a.nx;
        }
        return tc;
    };

    constexpr int kRing = kConvRing, kAhead = kRing - 1;
    constexpr int kRows = 3 * (R + 2);
    static_assert(kRows % kRing == 0, "a tile's rows must fill whole rings");
    u32x4 buf[kRing][4], lbuf[kRing], rbuf[kRing];
    bool have[kRing] = {};
    // Input rows are addressed from the tile's own row (voxel x0 of row y,
    // plane z) by whole-row and whole-plane strides.  Rows outside the volume
    // (skipped) and the halo voxels left of x = 0 and right of x = nx - 1 are
    // read from the zero row, so no address is clamped and the shifted
    // operands need no selects; a skipped row's halo voxels are the zero
    // row's padding.
    const uint32_t lane_off = 4 * lr * kC + 8 * lq;
    auto loadRow = [&](const RowsTile& tc, int q, int b) {
        const int dz = q / (R + 2) - 1, dy = q % (R + 2) - 1;
        have[b] = (tc.rows >> q) & 1u;
        const uint16_t* row = have[b] ? tc.own + dz * plane_stride + dy * row_stride : zrow;
        const uint16_t* left = tc.left ? row - kC : zrow;
        const uint16_t* right = tc.right ? row + 64 * kC : zrow;
#pragma unroll
        for (int j = 0; j < 4; ++j)
            buf[b][j] = load16(row + lane_off + j * kC);
        lbuf[b] = load16(left + 8 * lq);
        rbuf[b] = load16(right + 8 * lq);
    };
    // The weight fragments of a step (one tap, both channel halves) are read
    // from LDS one step ahead of their MFMAs.  The steps of a tile run over
    // input rows q, then dx, then the output rows o the tap row dy = q % (R + 2) - o
    // of which lies in 0..2; the step after the last is the next tile's first
    // (q 0, dx -1, o 0: tap 0).
    auto tapOf = [](int q, int dx, int o) { return (q / (R + 2)) * 9 + (q % (R + 2) - o) * 3 + dx + 1; };
    auto nextTap = [&](int q, int dx, int o) {
        const int rr = q % (R + 2), o_hi = min(R - 1, rr);
        if (o < o_hi)
            return tapOf(q, dx, o + 1);
        if (dx < 1)
            return tapOf(q, dx + 1, max(0, rr - 2));
        if (q + 1 < kRows)
            return tapOf(q + 1, -1, max(0, (q + 1) % (R + 2) - 2));
        return 0;
    };
    bf16x8 wcur[2];
#pragma unroll
    for (int m = 0; m < kHalves; ++m)
        wcur[m] = asOperand(wl[wSlot(0, 16 * m + lr, lq)]);

    int cmask = 0;
    RowsTile cur = tileAt(begin, cmask);
    if (cur.valid)
        for (int q = 0; q < kAhead; ++q)
            loadRow(cur, q, q);
    for (int wt = begin; cur.valid; ++wt) {
        int nmask = 0;
        const RowsTile nxt = tileAt(wt + 1, nmask);
        const int x0 = cur.x0, y = cur.y, z = cur.z, n = cur.n;

        f32x4 acc[R][2][4];   // [output row][channel half][voxel group]
        for (int o = 0; o < R; ++o)
            for (int m = 0; m < 2; ++m)
                for (int j = 0; j < 4; ++j)
                    acc[o][m][j] = f32x4 { 0, 0, 0, 0 };

        // the epilogue's residual (or gate) values of both rows, loaded before the MFMAs
        const size_t orow0 = (static_cast<size_t>(n * a.nz + z) * a.ny + y) * a.nx;
        // (up to two rows: more would not leave registers for the ring)
        constexpr bool kPrefetchSide = R <= 2 && kConvWaves <= 4;
        u32x2 side[kPrefetchSide ? R : 1][4][2];
        const uint16_t* side_src = e_res ? a.res : (e_gate ? a.gate : nullptr);
        if (kPrefetchSide && (e_res || e_gate)) {
#pragma unroll
            for (int o = 0; o < R; ++o) {
                if (!((cmask >> o) & 1))
                    continue;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const size_t v = orow0 + static_cast<size_t>(o) * a.nx + x0 + 4 * lr + j;
#pragma unroll
                    for (int m = 0; m < 2; ++m)
                        side[o][j][m] = *reinterpret_cast<const u32x2*>(side_src + v * kC + 16 * m + 4 * lq);
                }
            }
        }

#pragma unroll
        for (int q = 0; q < kRows; ++q) {
            const int b = q % kRing;
            if (q + kAhead < kRows)
                loadRow(cur, q + kAhead, (q + kAhead) % kRing);
            else
                loadRow(nxt, q + kAhead - kRows, (q + kAhead) % kRing);
            if (!have[b]) {
                // a row outside the volume: load the weights of the next row's first step
                const int tap_next = nextTap(q, 1, R - 1);
#pragma unroll
                for (int m = 0; m < kHalves; ++m)
                    wcur[m] = asOperand(wl[wSlot(tap_next, 16 * m + lr, lq)]);
                continue;
            }
            const int rr = q % (R + 2);   // input row y + rr - 1 of plane z + q / (R + 2) - 1
            const u32x4(&c)[4] = buf[b];
#pragma unroll
            for (int dx = -1; dx <= 1; ++dx) {
                bf16x8 bop[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    if (dx == 0)
                        bop[j] = asOperand(c[j]);
                    else if (dx < 0)
                        bop[j] = asOperand(j > 0 ? c[j - 1] : dpp4<kRowShr1>(lbuf[b], c[3]));   // lane n <- lane n-1
                    else
                        bop[j] = asOperand(j < 3 ? c[j + 1] : dpp4<kRowShl1>(rbuf[b], c[0]));   // lane n <- lane n+1
                }
#pragma unroll
                for (int o = 0; o < R; ++o) {
                    const int dy = rr - o;   // tap row of this input row for output row y + o
                    if (dy < 0 || dy > 2)
                        continue;
                    const int tap_next = nextTap(q, dx, o);
                    bf16x8 wnext[2];
#pragma unroll
                    for (int m = 0; m < kHalves; ++m)
                        wnext[m] = asOperand(wl[wSlot(tap_next, 16 * m + lr, lq)]);
#pragma unroll
                    for (int m = 0; m < kHalves; ++m)
#pragma unroll
                        for (int j = 0; j < 4; ++j)
                            acc[o][m][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wcur[m], bop[j], acc[o][m][j], 0, 0, 0);
#pragma unroll
                    for (int m = 0; m < kHalves; ++m)
                        wcur[m] = wnext[m];
                    // one step per scheduling region: without the fence the
                    // scheduler hoists a whole row's weight reads and spills
                    __builtin_amdgcn_sched_barrier(0);
                }
            }
        }

#pragma unroll
        for (int o = 0; o < R; ++o) {
            if (!((cmask >> o) & 1))
                continue;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const size_t v = orow0 + static_cast<size_t>(o) * a.nx + x0 + 4 * lr + j;
#pragma unroll
                for (int m = 0; m < kHalves; ++m) {
                    const int co0 = 16 * m + 4 * lq;
                    float out[4];
                    for (int i = 0; i < 4; ++i)
                        out[i] = acc[o][m][j][i] + bias[m][i];
                    const u32x2 sv = kPrefetchSide ? side[kPrefetchSide ? o : 0][j][m]
                                                   : ((e_res || e_gate) ? *reinterpret_cast<const u32x2*>(side_src + v * kC + co0) : u32x2 { 0, 0 });
                    if (e_res) {
                        const u32x2 rv = sv;
                        out[0] += fromBf16(rv[0] & 0xffffu);
                        out[1] += fromBf16(rv[0] >> 16);
                        out[2] += fromBf16(rv[1] & 0xffffu);
                        out[3] += fromBf16(rv[1] >> 16);
                    }
                    if (e_relu)
                        for (int i = 0; i < 4; ++i)
                            out[i] = fmaxf(out[i], 0.0f);
                    if (e_gate) {
                        const u32x2 gv = e_res ? *reinterpret_cast<const u32x2*>(a.gate + v * kC + co0) : sv;
                        const uint32_t h[4] = { gv[0] & 0xffffu, gv[0] >> 16, gv[1] & 0xffffu, gv[1] >> 16 };
                        for (int i = 0; i < 4; ++i)
                            if (h[i] == 0 || (h[i] & 0x8000u))
                                out[i] = 0.0f;
                    }
                    if (e_y)
                        *reinterpret_cast<u32x2*>(a.y + v * kC + co0) = u32x2 { pack2(out[0], out[1]), pack2(out[2], out[3]) };
                    if (e_y32) {
                        if (e_y32_all)
                            *reinterpret_cast<f32x4*>(a.y32 + v * kC + co0) = f32x4 { out[0], out[1], out[2], out[3] };
                        else if (co0 == 0)
                            a.y32[v] = out[0];
                    }
                }
            }
        }
        cur = nxt;
        cmask = nmask;
    }
}

// First layer: the input has four feature channels (in the 32-channel
// tensor the other 28 are zero), so a 3x3x3 tap uses 4 of the 32 K-values of
// an MFMA and the general kernel spends 7/8 of its MFMAs on zeros.  Here K
// runs over (tap, channel): 27 taps x 4 channels = 108 values, padded to 128
// = four MFMA K-chunks, so an output row costs 4 x 2 x 4 MFMAs instead of
// 27 x 2 x 4.  The input is the features' compact copy (4 BF16 channels, 8
// bytes per voxel).  Lane (lr, lq) holds K-values 8 lq .. 8 lq + 7 of a chunk
// c: taps 8c + 2lq and 8c + 2lq + 1, read as two 8-byte loads from the
// neighbouring voxels of voxel x0 + 16 j + lr (taps past 26 and voxels outside
// the volume read as zero), so the 16 lanes of a quarter read 128 contiguous
// bytes.  The weight fragments (the layer's 27 x 32 x 4 used weights) stay in
// registers.  Same tiles (row pairs) and epilogue (bias, ReLU, BF16 output)
// as conv3d_rows_kernel; results equal up to FP32 summation order.
__global__ __launch_bounds__(64 * kConvWaves) void conv3d_first_kernel(ConvArgs a)
{
    constexpr int R = kConvRows;
    const int tid = threadIdx.x;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63, lr = lane & 15, lq = lane >> 4;

    // weight fragments: aw[m][c] = W[tap][co = 16 m + lr][ci] for the eight
    // (tap, ci) of chunk c and quarter lq
    bf16x8 aw[2][4];
    // the lane's two taps per chunk: voxel offset from the tile row's corner
    // (z - 1, y - 1, x0 - 1), so that every offset is non-negative; taps past
    // 26 take the centre voxel and are dropped
    const int corner = (a.ny + 1) * a.nx + 1;
    unsigned toff[4][2];
    int tdx[4][2], tdy[4][2], tdz[4][2];
    bool tval[4][2];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const int t = 8 * c + 2 * lq + k;
            tval[c][k] = t < kTaps;
            const int tt = tval[c][k] ? t : 13;
            tdz[c][k] = tt / 9 - 1;
            tdy[c][k] = (tt / 3) % 3 - 1;
            tdx[c][k] = tt % 3 - 1;
            toff[c][k] = static_cast<unsigned>(corner + (tdz[c][k] * a.ny + tdy[c][k]) * a.nx + tdx[c][k]);
        }
#pragma unroll
        for (int m = 0; m < 2; ++m) {
            u32x4 v;
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                u32x2 w2 = { 0u, 0u };
                if (tval[c][k])
                    w2 = *reinterpret_cast<const u32x2*>(a.w + (static_cast<size_t>(8 * c + 2 * lq + k) * 32 + 16 * m + lr) * kC);
                v[2 * k] = w2[0];
                v[2 * k + 1] = w2[1];
            }
            aw[m][c] = asOperand(v);
        }
    }
    float bias[2][4];
    for (int m = 0; m < 2; ++m)
        for (int i = 0; i < 4; ++i)
            bias[m][i] = a.bias ? a.bias[16 * m + 4 * lq + i] : 0.0f;

    const int g = gridDim.x;
    const int logical = (blockIdx.x & 7) * (g >> 3) + (blockIdx.x >> 3);
    const int wg_tiles = (a.wave_tiles + kConvWaves - 1) / kConvWaves;
    const int begin = logical * a.chunk;
    const int end = min(begin + a.chunk, wg_tiles);
    const int nxs = a.nx >> 6, nyp = (a.ny + R - 1) / R;
    for (int wt = begin; wt < end; ++wt) {
        const int kt = wt * kConvWaves + wave;
        if (kt >= a.wave_tiles)
            break;
        int t = kt, mask;
        if (a.tiles) {
            t = a.tiles[kt] >> R;
            mask = a.tiles[kt] & ((1 << R) - 1);
        }
        int r = t / nxs;
        const int x0 = (t - r * nxs) * 64;
        const int y = R * (r % nyp);
        r /= nyp;
        const int z = r % a.nz, n = r / a.nz;
        if (!a.tiles)
            mask = (1 << min(R, a.ny - y)) - 1;
        const size_t plane = static_cast<size_t>(n) * a.nz * a.ny * a.nx;
        const bool interior = x0 > 0 && x0 + 64 < a.nx;   // no tap of the row leaves the volume along x
#pragma unroll
        for (int o = 0; o < R; ++o) {
            if (!((mask >> o) & 1))
                continue;
            const int yo = y + o;
            const size_t orow = plane + (static_cast<size_t>(z) * a.ny + yo) * a.nx + x0;
            // the row's corner voxel (may lie before the tensor: only offsets from it are read)
            const uint16_t* base = a.x + (static_cast<long long>(orow) - corner) * 4;
            bool yz[4][2];
#pragma unroll
            for (int c = 0; c < 4; ++c)
#pragma unroll
                for (int k = 0; k < 2; ++k)
                    yz[c][k] = tval[c][k] && static_cast<unsigned>(yo + tdy[c][k]) < static_cast<unsigned>(a.ny)
                               && static_cast<unsigned>(z + tdz[c][k]) < static_cast<unsigned>(a.nz);
            f32x4 acc[2][4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int xl = 16 * j + lr;   // the output voxel within the row
                bf16x8 bop[4];
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    u32x4 v;
#pragma unroll
                    for (int k = 0; k < 2; ++k) {
                        const bool in = yz[c][k] && (interior || static_cast<unsigned>(x0 + xl + tdx[c][k]) < static_cast<unsigned>(a.nx));
                        const unsigned vo = static_cast<unsigned>(xl) + (in ? toff[c][k] : static_cast<unsigned>(corner));
                        const u32x2 x2 = *reinterpret_cast<const u32x2*>(base + vo * 4u);
                        v[2 * k] = in ? x2[0] : 0u;
                        v[2 * k + 1] = in ? x2[1] : 0u;
                    }
                    bop[c] = asOperand(v);
                }
#pragma unroll
                for (int m = 0; m < 2; ++m) {
                    acc[m][j] = f32x4 { 0, 0, 0, 0 };
#pragma unroll
                    for (int c = 0; c < 4; ++c)
                        acc[m][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(aw[m][c], bop[c], acc[m][j], 0, 0, 0);
                }
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const size_t v = orow + 16 * j + lr;
#pragma unroll
                for (int m = 0; m < 2; ++m) {
                    float out[4];
                    for (int i = 0; i < 4; ++i) {
                        out[i] = acc[m][j][i] + bias[m][i];
                        if (a.relu)
                            out[i] = fmaxf(out[i], 0.0f);
                    }
                    *reinterpret_cast<u32x2*>(a.y + v * kC + 16 * m + 4 * lq) = u32x2 { pack2(out[0], out[1]), pack2(out[2], out[3]) };
                }
            }
        }
    }
}

// The row kernel for a launch's epilogue: the configurations the network and
// the probe use are compiled on their own; any other reads its flags at run time.
const void* rowsKernel(const ConvArgs& a)
{
    const int epi = (a.relu ? kEpiRelu : 0) | (a.res ? kEpiRes : 0) | (a.gate ? kEpiGate : 0) | (a.y ? kEpiY : 0)
                    | (a.y32 ? (a.y32_channels == kC ? kEpiY32All : kEpiY32) : 0);
    switch (epi) {
    case kEpiRelu | kEpiY:
        return reinterpret_cast<const void*>(&conv3d_rows_kernel<kConvRows, kEpiRelu | kEpiY>);
    case kEpiRes | kEpiRelu | kEpiY:
        return reinterpret_cast<const void*>(&conv3d_rows_kernel<kConvRows, kEpiRes | kEpiRelu | kEpiY>);
    case kEpiY32:
        return reinterpret_cast<const void*>(&conv3d_rows_kernel<kConvRows, kEpiY32>);
    case kEpiGate | kEpiY:
        return reinterpret_cast<const void*>(&conv3d_rows_kernel<kConvRows, kEpiGate | kEpiY>);
    case kEpiRes | kEpiGate | kEpiY:
        return reinterpret_cast<const void*>(&conv3d_rows_kernel<kConvRows, kEpiRes | kEpiGate | kEpiY>);
    case kEpiY32All:
        return reinterpret_cast<const void*>(&conv3d_rows_kernel<kConvRows, kEpiY32All>);
    default:
        return reinterpret_cast<const void*>(&conv3d_rows_kernel<kConvRows, -1>);
    }
}

struct WgradArgs {
    const uint16_t* x;       // layer input
    const uint16_t* g;       // gradient of the layer's output (before its activation)
    float* part;             // [grid][27][32 co][32 ci] partial sums
    float* part_b;           // [grid][32]
    int nb, nz, ny, nx;
    int units;               // nb*nz*(ny/4)*(nx/32)
    int chunk;
};

__global__ __launch_bounds__(256) void wgrad_kernel(WgradArgs a)
{
    __shared__ __align__(16) uint16_t xT[18 * 32 * kXtLen];   // halo rows (dz, y) x channel x voxel: 64,512 B
    __shared__ __align__(16) uint16_t gT[4 * 32 * kGtLen];    // 4 rows x channel x voxel: 10,240 B
    __shared__ float bsh[32];
    const int tid = threadIdx.x;
    if (tid < 32)
        bsh[tid] = 0.0f;
    const int wave = tid >> 6, lane = tid & 63, lr = lane & 15, lq = lane >> 4;
    const int tap0 = wave * 7;
    const int ntap = min(7, kTaps - tap0);

    f32x4 acc[7][2][2];
    for (int k = 0; k < 7; ++k)
        for (int t = 0; t < 2; ++t)
            for (int s = 0; s < 2; ++s)
                acc[k][t][s] = f32x4 { 0, 0, 0, 0 };
    float bsum = 0.0f;

    const int g = gridDim.x;
    const int logical = (blockIdx.x & 7) * (g >> 3) + (blockIdx.x >> 3);
    const int begin = logical * a.chunk;
    const int end = min(begin + a.chunk, a.units);
    const int nxc = a.nx >> 5, nyq = a.ny >> 2;

    for (int u = begin; u < end; ++u) {
        int r = u / nxc;
        const int x0 = (u - r * nxc) * 32;
        const int y0 = (r % nyq) * 4;
        r /= nyq;
        const int z = r % a.nz;
        const int n = r / a.nz;
        __syncthreads();   // the previous tile is consumed
        for (int i = tid; i < 4 * 32 * 4; i += 256) {
            const int q = i & 3, vox = (i >> 2) & 31, rr = i >> 7;
            const size_t v = (static_cast<size_t>(n * a.nz + z) * a.ny + y0 + rr) * a.nx + x0 + vox;
            const u32x4 d = load16(a.g + v * kC + 8 * q);
            uint16_t* dst = gT + (rr * 32 + 8 * q) * kGtLen + vox;
            for (int e = 0; e < 4; ++e) {
                dst[(2 * e) * kGtLen] = static_cast<uint16_t>(d[e] & 0xffffu);
                dst[(2 * e + 1) * kGtLen] = static_cast<uint16_t>(d[e] >> 16);
            }
        }
        for (int i = tid; i < 18 * 34 * 4; i += 256) {
            const int q = i & 3, t = i >> 2;
            const int row = t / 34, vox = t - row * 34;
            const int zz = z + row / 6 - 1, yy = y0 + row % 6 - 1, xx = x0 + vox - 1;
            u32x4 d = { 0, 0, 0, 0 };
            if (static_cast<unsigned>(zz) < static_cast<unsigned>(a.nz) && static_cast<unsigned>(yy) < static_cast<unsigned>(a.ny)
                && static_cast<unsigned>(xx) < static_cast<unsigned>(a.nx))
                d = load16(a.x + ((static_cast<size_t>(n * a.nz + zz) * a.ny + yy) * a.nx + xx) * kC + 8 * q);
            uint16_t* dst = xT + (row * 32 + 8 * q) * kXtLen + vox;
            for (int e = 0; e < 4; ++e) {
                dst[(2 * e) * kXtLen] = static_cast<uint16_t>(d[e] & 0xffffu);
                dst[(2 * e + 1) * kXtLen] = static_cast<uint16_t>(d[e] >> 16);
            }
        }
        __syncthreads();
        if (tid < 128) {
            const uint16_t* s = gT + ((tid >> 5) * 32 + (tid & 31)) * kGtLen;
            for (int v = 0; v < 32; ++v)
                bsum += fromBf16(s[v]);
        }
        for (int rr = 0; rr < 4; ++rr) {
            bf16x8 ga[2];
#pragma unroll
            for (int t = 0; t < 2; ++t)
                ga[t] = asOperand(load16(gT + (rr * 32 + 16 * t + lr) * kGtLen + 8 * lq));
#pragma unroll
            for (int k = 0; k < 7; ++k) {
                if (k >= ntap)
                    break;
                const int tap = tap0 + k;
                const int dzi = tap / 9, dyi = (tap / 3) % 3, dxi = tap % 3;
                const int row = dzi * 6 + rr + dyi;
#pragma unroll
                for (int s = 0; s < 2; ++s) {
                    const uint16_t* p = xT + (row * 32 + 16 * s + lr) * kXtLen + 8 * lq;
                    const u32x4 lo = load16(p), hi = load16(p + 8);
                    u32x4 b;
                    if (dxi == 0) {
                        b = lo;
                    } else if (dxi == 1) {
                        b = u32x4 { __builtin_amdgcn_alignbyte(lo[1], lo[0], 2), __builtin_amdgcn_alignbyte(lo[2], lo[1], 2),
                                    __builtin_amdgcn_alignbyte(lo[3], lo[2], 2), __builtin_amdgcn_alignbyte(hi[0], lo[3], 2) };
                    } else {
                        b = u32x4 { lo[1], lo[2], lo[3], hi[0] };
                    }
                    const bf16x8 bop = asOperand(b);
#pragma unroll
                    for (int t = 0; t < 2; ++t)
                        acc[k][t][s] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ga[t], bop, acc[k][t][s], 0, 0, 0);
                }
            }
        }
    }

    float* part = a.part + static_cast<size_t>(blockIdx.x) * kLayerW;
#pragma unroll
    for (int k = 0; k < 7; ++k) {
        if (k >= ntap)
            break;
        const int tap = tap0 + k;
        for (int t = 0; t < 2; ++t)
            for (int s = 0; s < 2; ++s)
                for (int i = 0; i < 4; ++i)
                    part[(tap * 32 + 16 * t + 4 * lq + i) * 32 + 16 * s + lr] = acc[k][t][s][i];
    }
    __syncthreads();
    if (tid < 128)
        atomicAdd(&bsh[tid & 31], bsum);
    __syncthreads();
    if (tid < 32)
        a.part_b[static_cast<size_t>(blockIdx.x) * 32 + tid] = bsh[tid];
}

// Sum the per-work-group partials into one layer's gradient, in two fixed-
// order stages: slice y of the partials into sums[y], then the slices into the
// parameter layout ([co][tap][ci] weights, then 32 biases).
constexpr int kReduceSlices = 8;

__global__ __launch_bounds__(256) void wgrad_reduce_kernel(const float* part, const float* part_b, int parts, float* sums)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int per = (parts + kReduceSlices - 1) / kReduceSlices;
    const int p0 = blockIdx.y * per, p1 = min(parts, p0 + per);
    if (i < kLayerW) {
        float s0 = 0, s1 = 0, s2 = 0, s3 = 0;
        int p = p0;
        for (; p + 4 <= p1; p += 4) {
            s0 += part[static_cast<size_t>(p) * kLayerW + i];
            s1 += part[static_cast<size_t>(p + 1) * kLayerW + i];
            s2 += part[static_cast<size_t>(p + 2) * kLayerW + i];
            s3 += part[static_cast<size_t>(p + 3) * kLayerW + i];
        }
        for (; p < p1; ++p)
            s0 += part[static_cast<size_t>(p) * kLayerW + i];
        sums[blockIdx.y * kLayerP + i] = (s0 + s1) + (s2 + s3);
    } else if (i < kLayerP) {
        const int co = i - kLayerW;
        float s = 0;
        for (int p = p0; p < p1; ++p)
            s += part_b[p * 32 + co];
        sums[blockIdx.y * kLayerP + i] = s;
    }
}

__global__ __launch_bounds__(256) void wgrad_finish_kernel(const float* sums, float* grad)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= kLayerP)
        return;
    float s = 0;
    for (int k = 0; k < kReduceSlices; ++k)
        s += sums[k * kLayerP + i];
    if (i < kLayerW) {
        const int tap = i >> 10, co = (i >> 5) & 31, ci = i & 31;
        grad[(co * kTaps + tap) * kC + ci] = s;
    } else {
        grad[i] = s;
    }
}

// FP32 master parameters -> BF16 forward weights [tap][co][ci] and input-
// gradient weights wd[tap][co][ci] = W[ci][26 - tap][co] (flipped, transposed).
__global__ __launch_bounds__(256) void prep_weights_kernel(const float* params, uint16_t* wf, uint16_t* wd, int layers)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= layers * kLayerW)
        return;
    const int l = i / kLayerW, r = i - l * kLayerW;
    const int tap = r >> 10, co = (r >> 5) & 31, ci = r & 31;
    const float* p = params + static_cast<size_t>(l) * kLayerP;
    wf[i] = toBf16(p[(co * kTaps + tap) * kC + ci]);
    wd[i] = toBf16(p[(ci * kTaps + (kTaps - 1 - tap)) * kC + co]);
}

__global__ __launch_bounds__(256) void adam_kernel(float* p, float* m, float* v, const float* g, int n, float lr, float b1,
                                                   float b2, float eps, float bc1, float bc2)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n)
        return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi * bc1) / (sqrtf(vi * bc2) + eps);
}

// Features (4 FP32 per voxel) -> 32-channel BF16 input tensor and an FP32 copy
// of feature 0 (the relative dose the network corrects).
__global__ __launch_bounds__(256) void pack_features_kernel(const float* f, uint16_t* xin, uint16_t* xin4, float* feat0, size_t nv)
{
    const size_t v = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
    if (v >= nv)
        return;
    const f32x4 q = *reinterpret_cast<const f32x4*>(f + 4 * v);
    u32x4* dst = reinterpret_cast<u32x4*>(xin + v * kC);
    const u32x2 f4 = { pack2(q[0], q[1]), pack2(q[2], q[3]) };
    dst[0] = u32x4 { f4[0], f4[1], 0u, 0u };
    *reinterpret_cast<u32x2*>(xin4 + 4 * v) = f4;
    dst[1] = u32x4 { 0u, 0u, 0u, 0u };
    dst[2] = u32x4 { 0u, 0u, 0u, 0u };
    dst[3] = u32x4 { 0u, 0u, 0u, 0u };
    feat0[v] = q[0];
}

// Masked squared error of pred = feat0 + correction against the target, and
// its gradient into channel 0 of a 32-channel BF16 tensor.
__global__ __launch_bounds__(256) void loss_kernel(const float* y32, const float* feat0, const float* target,
                                                   const float* mask, float inv_count, uint16_t* gout, float* lpart,
                                                   size_t nv)
{
    __shared__ float red[4];
    const size_t v = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
    float l = 0.0f;
    if (v < nv) {
        const float diff = feat0[v] + y32[v] - target[v];
        const float mk = mask[v];
        l = mk * diff * diff;
        u32x4* dst = reinterpret_cast<u32x4*>(gout + v * kC);
        dst[0] = u32x4 { static_cast<uint32_t>(toBf16(2.0f * mk * diff * inv_count)), 0u, 0u, 0u };
        dst[1] = u32x4 { 0u, 0u, 0u, 0u };
        dst[2] = u32x4 { 0u, 0u, 0u, 0u };
        dst[3] = u32x4 { 0u, 0u, 0u, 0u };
    }
    for (int off = 32; off > 0; off >>= 1)
        l += __shfl_down(l, off, 64);
    if ((threadIdx.x & 63) == 0)
        red[threadIdx.x >> 6] = l;
    __syncthreads();
    if (threadIdx.x == 0)
        lpart[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

__global__ __launch_bounds__(256) void finalize_kernel(const float* y32, const float* feat0, float* pred, size_t nv)
{
    const size_t v = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
    if (v < nv)
        pred[v] = feat0[v] + y32[v];
}

// ---- features and dose straight from a world's device tally

__device__ inline float wordDensity(uint32_t word) { return __uint_as_float(word & 0xFFFFFF00u); }

__global__ __launch_bounds__(64) void zero2_kernel(double* p)
{
    if (threadIdx.x < 2)
        p[threadIdx.x] = 0.0;
}

// acc[0] += sum over body voxels (material != 0) of sum_wdE / density, acc[1] += their count.
__global__ __launch_bounds__(256) void dose_mean_kernel(const uint32_t* voxels, const double* tally, size_t nv, double* acc)
{
    __shared__ double red[2][4];
    double s = 0.0, c = 0.0;
    for (size_t v = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x; v < nv; v += static_cast<size_t>(gridDim.x) * 256) {
        const uint32_t word = voxels[v];
        if ((word & 0xFFu) != 0) {
            const float rho = wordDensity(word);
            if (rho > 0.0f)
                s += tally[v * dxg::kTallyRecord] / static_cast<double>(rho);
            c += 1.0;
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        s += __shfl_down(s, off, 64);
        c += __shfl_down(c, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = s;
        red[1][threadIdx.x >> 6] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        atomicAdd(&acc[0], (red[0][0] + red[0][1]) + (red[0][2] + red[0][3]));
        atomicAdd(&acc[1], (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]));
    }
}

__device__ inline double doseRef(const double* acc)
{
    const double r = acc[1] > 0.0 ? acc[0] / acc[1] : 0.0;
    return r > 0.0 ? r : 1.0;
}

struct WorldFeatArgs {
    const uint32_t* voxels;
    const double* tally;    // records of dxg::kTallyRecord doubles: sum, sum of squares, ...
    const float* emb;
    int n_emb;
    int nx, nxp;            // grid x extent and its padded extent
    size_t nvp;             // padded voxels
    double histories;
    const double* acc;      // dose_mean_kernel sums
    uint16_t* xin;          // 32-channel BF16 network input (null: only the compact copy is needed)
    uint16_t* xin4;         // its four feature channels alone (8 bytes per voxel)
    float* feat0;
    float* features;        // optional FP32 copy (nvp x 4)
};

// The four input features of every (padded) voxel, as the host computes them
// from DoseScore: D / D_ref, density, material embedding, relative standard
// error.  Padding voxels get zero features.
__global__ __launch_bounds__(256) void world_features_kernel(WorldFeatArgs a)
{
    const size_t v = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
    if (v >= a.nvp)
        return;
    const size_t row = v / a.nxp;
    const int x = static_cast<int>(v - row * a.nxp);
    float f[4] = { 0.0f, 0.0f, 0.0f, 0.0f };
    if (x < a.nx) {
        const size_t i = row * a.nx + x;
        const uint32_t word = a.voxels[i];
        const uint32_t m = word & 0xFFu;
        const float rho = wordDensity(word);
        const double e = a.tally[i * dxg::kTallyRecord], e2 = a.tally[i * dxg::kTallyRecord + 1];
        const double d = rho > 0.0f ? e / static_cast<double>(rho) : 0.0;
        f[0] = static_cast<float>(d / doseRef(a.acc));
        f[1] = rho;
        f[2] = static_cast<int>(m) < a.n_emb ? a.emb[m] : 0.0f;
        float rel = 1.0f;
        if (d > 0.0 && e > 0.0) {
            const double h = a.histories;
            const double var = h > 1.0 ? fmax(0.0, h / (h - 1.0) * (e2 - e * e / h)) : 0.0;
            rel = fminf(static_cast<float>(sqrt(var) / e), 1.0f);
        }
        f[3] = rel;
    }
    const u32x2 f4 = { pack2(f[0], f[1]), pack2(f[2], f[3]) };
    *reinterpret_cast<u32x2*>(a.xin4 + 4 * v) = f4;
    if (a.xin) {
        u32x4* dst = reinterpret_cast<u32x4*>(a.xin + v * kC);
        dst[0] = u32x4 { f4[0], f4[1], 0u, 0u };
        dst[1] = u32x4 { 0u, 0u, 0u, 0u };
        dst[2] = u32x4 { 0u, 0u, 0u, 0u };
        dst[3] = u32x4 { 0u, 0u, 0u, 0u };
    }
    a.feat0[v] = f[0];
    if (a.features)
        *reinterpret_cast<f32x4*>(a.features + 4 * v) = f32x4 { f[0], f[1], f[2], f[3] };
}

// Denoised dose (mGy) from the predicted relative dose.
__global__ __launch_bounds__(256) void world_dose_kernel(const float* feat0, const float* corr, int nx, int nxp, size_t nv,
                                                         const double* acc, double scale, double* out)
{
    const size_t v = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
    if (v >= nv)
        return;
    const size_t row = v / nx;
    const size_t i = row * nxp + (v - row * nx);
    out[v] = fmax(static_cast<double>(feat0[i] + corr[i]), 0.0) * doseRef(acc) * scale;
}

// ---------------------------------------------------------------- host side

int computeUnits()
{
    static int cus = 0;
    if (cus == 0) {
        int dev = 0, n = 0;
        if (hipGetDevice(&dev) == hipSuccess && hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess
            && n > 0)
            cus = n;
        else
            cus = 256;
    }
    return cus;
}

int roundUp8(int v) { return (v + 7) & ~7; }

// Activation operands for the x-neighbour taps: moved across lanes (default)
// or loaded again per tap (DXG_CONV_SHIFT=0).
bool convShift()
{
    static const bool on = [] {
        const char* e = std::getenv("DXG_CONV_SHIFT");
        return !(e && e[0] == '0');
    }();
    return on;
}

// Row-pair convolution (conv3d_pair_kernel) unless DXG_CONV_PAIRS=0.
bool convPairs()
{
    static const bool on = [] {
        const char* e = std::getenv("DXG_CONV_PAIRS");
        return !(e && e[0] == '0');
    }();
    return on;
}

// The first layer by conv3d_first_kernel (its four input channels packed
// over the taps) unless DXG_CONV_FIRST=0.
int convFirst()
{
    static const int on = [] {
        const char* e = std::getenv("DXG_CONV_FIRST");
        return !(e && e[0] == '0') ? 1 : 0;
    }();
    return on;
}

struct Dims {
    int nb, nz, ny, nx;
    size_t voxels() const { return static_cast<size_t>(nb) * nz * ny * nx; }
};

// Whether the first layer runs conv3d_first_kernel on the compact input: row
// pairs on, and a plane small enough for its 32-bit offsets.
bool firstLayerCompact(const Dims& d)
{
    return convFirst() && convPairs() && static_cast<long long>(d.ny + 1) * d.nx < (1ll << 27);
}

int checkDims(const char* fn, uint64_t batch, const uint64_t* dims, Dims& d)
{
    if (!dims)
        return fail(DXG_ERR_ARGUMENT, std::string(fn) + ": null dims");
    if (batch == 0 || dims[0] == 0 || dims[1] == 0 || dims[2] == 0)
        return fail(DXG_ERR_ARGUMENT, std::string(fn) + ": empty volume");
    if (dims[2] % 64 != 0)
        return fail(DXG_ERR_ARGUMENT, std::string(fn) + ": the x extent must be a multiple of 64 (pad the volume)");
    const double nv = static_cast<double>(batch) * dims[0] * dims[1] * dims[2];
    if (nv > 4.0e9 || dims[0] > (1u << 20) || dims[1] > (1u << 20) || dims[2] > (1u << 20))
        return fail(DXG_ERR_ARGUMENT, std::string(fn) + ": volume too large");
    d = { static_cast<int>(batch), static_cast<int>(dims[0]), static_cast<int>(dims[1]), static_cast<int>(dims[2]) };
    return DXG_OK;
}

// kZeroRow's address on the current device (each device holds its own copy).
const uint16_t* zeroRow()
{
    static std::mutex mu;
    static std::map<int, const uint16_t*> rows;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess)
        return nullptr;
    std::lock_guard<std::mutex> lock(mu);
    auto it = rows.find(dev);
    if (it != rows.end())
        return it->second;
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(kZeroRow)) != hipSuccess)
        return nullptr;
    return rows[dev] = static_cast<const uint16_t*>(p) + kC;
}

hipError_t launchConv(ConvArgs a, const Dims& d, hipStream_t s)
{
    a.zrow = zeroRow();
    if (!a.zrow)
        return hipErrorInvalidSymbol;
    a.nb = d.nb;
    a.nz = d.nz;
    a.ny = d.ny;
    a.nx = d.nx;
    const bool pairs = convPairs();
    if (!a.tiles)
        a.wave_tiles = d.nb * d.nz * (pairs ? (d.ny + kConvRows - 1) / kConvRows : d.ny) * (d.nx / 64);
    const int wg_tiles = (a.wave_tiles + kConvWaves - 1) / kConvWaves;
    const int grid = roundUp8(std::min(wg_tiles, 2 * computeUnits()));
    a.chunk = (wg_tiles + grid - 1) / grid;
    // the first layer reads the compact 4-channel input: only its own kernel
    // can (its 32-bit element offsets from a row's corner stay below 2^31
    // while (ny + 1) nx < 2^27)
    if (a.in4) {
        if (!pairs || !a.y || a.y32 || a.res || a.gate || static_cast<long long>(d.ny + 1) * d.nx >= (1ll << 27))
            return hipErrorInvalidValue;
        hipLaunchKernelGGL(conv3d_first_kernel, dim3(grid), dim3(64 * kConvWaves), 0, s, a);
    } else if (pairs) {
        void* args[] = { &a };
        (void)hipLaunchKernel(rowsKernel(a), dim3(grid), dim3(64 * kConvWaves), args, 0, s);
    } else if (convShift())
        hipLaunchKernelGGL(conv3d_kernel<true>, dim3(grid), dim3(64 * kConvWaves), 0, s, a);
    else
        hipLaunchKernelGGL(conv3d_kernel<false>, dim3(grid), dim3(64 * kConvWaves), 0, s, a);
    return hipGetLastError();
}

int wgradGrid(const Dims& d)
{
    const int units = d.nb * d.nz * (d.ny / 4) * (d.nx / 32);
    return roundUp8(std::min(units, 2 * computeUnits()));
}

hipError_t launchWgrad(const uint16_t* x, const uint16_t* g, float* part, float* part_b, float* sums, float* grad, const Dims& d,
                       hipStream_t s)
{
    WgradArgs a {};
    a.x = x;
    a.g = g;
    a.part = part;
    a.part_b = part_b;
    a.nb = d.nb;
    a.nz = d.nz;
    a.ny = d.ny;
    a.nx = d.nx;
    a.units = d.nb * d.nz * (d.ny / 4) * (d.nx / 32);
    const int grid = wgradGrid(d);
    a.chunk = (a.units + grid - 1) / grid;
    hipLaunchKernelGGL(wgrad_kernel, dim3(grid), dim3(256), 0, s, a);
    hipLaunchKernelGGL(wgrad_reduce_kernel, dim3((kLayerP + 255) / 256, kReduceSlices), dim3(256), 0, s, part, part_b, grid, sums);
    hipLaunchKernelGGL(wgrad_finish_kernel, dim3((kLayerP + 255) / 256), dim3(256), 0, s, sums, grad);
    return hipGetLastError();
}

// Receptive-field radius of the 18 convolutions, plus one voxel: a voxel of
// the body depends on nothing farther than this.
constexpr int kReach = kLayers + 1;

// Body-tile lists.  Layer l's outputs feed 17 - l more convolutions, so they
// matter within layerReach(l) = kReach - l voxels (Chebyshev) of a voxel with
// material != 0; computing layer l on the wave tiles (64 voxels along x of
// one row) that hold such a voxel keeps the body's predictions equal to the
// full computation: values outside a layer's tiles are wrong, and wrong
// values move inward by one voxel per layer, never reaching that radius.
constexpr int layerReach(int l) { return kReach - l; }

// A tile's span: the first and last x offset (0..63) of its body voxels,
// packed as first | last << 8; first = 255 (last = 0) when it has none.
constexpr uint16_t kNoSpan = 255;
__host__ __device__ inline uint16_t spanJoin(uint16_t a, uint16_t b)
{
    const unsigned lo = min(a & 0xFFu, b & 0xFFu), hi = max(a >> 8, b >> 8);
    return static_cast<uint16_t>(lo | hi << 8);
}
// Whether tile t of a row holds a voxel within r voxels along x of a body
// voxel, given the spans of the row's tiles already joined over the rows and
// planes within r (r < 64: only the neighbouring tiles can reach).
__host__ __device__ inline bool spanReaches(const uint16_t* row, int t, int nxt, int r)
{
    return (row[t] & 0xFFu) != kNoSpan || (t > 0 && static_cast<int>(row[t - 1] >> 8) >= 64 - r)
           || (t + 1 < nxt && static_cast<int>(row[t + 1] & 0xFFu) < r);
}

// The wave tiles of every layer, ascending, from the host copy of the
// material array: spans per tile, joined over +-1 row and plane r times for
// reach r (a box of half-width r is r boxes of half-width 1), then the x rule.
void hostLayerTiles(const uint8_t* mat, int nx, int ny, int nz, int nxp, std::vector<int> (&out)[kLayers])
{
    const int nxt = nxp / 64;
    const size_t rows = static_cast<size_t>(nz) * ny, n = rows * nxt;
    std::vector<uint16_t> span(n, kNoSpan), tmp(n);
    for (size_t r = 0; r < rows; ++r)
        for (int t = 0; t * 64 < nx; ++t) {
            const uint8_t* v = mat + r * nx + t * 64;
            const int m = std::min(64, nx - t * 64);
            for (int x = 0; x < m; ++x)
                if (v[x])
                    span[r * nxt + t] = spanJoin(span[r * nxt + t], static_cast<uint16_t>(x | x << 8));
        }
    for (int reach = 1; reach <= kReach; ++reach) {
        for (size_t i = 0; i < n; ++i) {   // +-1 row
            const size_t rr = i / nxt;
            const int y = static_cast<int>(rr % ny);
            uint16_t v = span[i];
            if (y > 0)
                v = spanJoin(v, span[i - nxt]);
            if (y + 1 < ny)
                v = spanJoin(v, span[i + nxt]);
            tmp[i] = v;
        }
        const size_t plane = static_cast<size_t>(ny) * nxt;
        for (size_t i = 0; i < n; ++i) {   // +-1 plane
            const int z = static_cast<int>(i / plane);
            uint16_t v = tmp[i];
            if (z > 0)
                v = spanJoin(v, tmp[i - plane]);
            if (z + 1 < nz)
                v = spanJoin(v, tmp[i + plane]);
            span[i] = v;
        }
        const int l = kReach - reach;
        if (l < 0 || l >= kLayers)
            continue;
        out[l].clear();
        for (size_t r = 0; r < rows; ++r)
            for (int t = 0; t < nxt; ++t)
                if (spanReaches(&span[r * nxt], t, nxt, reach))
                    out[l].push_back(static_cast<int>(r * nxt + t));
    }
}

// The row pairs of conv3d_pair_kernel that hold the given wave tiles
// (ascending tile indices of a volume with ny rows and nxt tiles per row), as
// pair * 4 + rows to store (bit 0 row y, bit 1 row y + 1), ascending.
std::vector<int> pairTiles(const std::vector<int>& tiles, int ny, int nxt)
{
    if (tiles.empty())
        return {};
    const int nyp = (ny + kConvRows - 1) / kConvRows;
    const int planes = tiles.back() / nxt / ny + 1;
    std::vector<uint8_t> mask(static_cast<size_t>(planes) * nyp * nxt, 0);   // rows to store per group
    for (const int t : tiles) {
        const int r = t / nxt, xt = t - r * nxt;
        const int y = r % ny, plane = r / ny;
        mask[(static_cast<size_t>(plane) * nyp + y / kConvRows) * nxt + xt] |= static_cast<uint8_t>(1 << (y % kConvRows));
    }
    std::vector<int> out;
    for (size_t g = 0; g < mask.size(); ++g)
        if (mask[g])
            out.push_back((static_cast<int>(g) << kConvRows) | mask[g]);
    return out;
}

// ---- the row-pair tile lists on the device (the same lists as pairTiles of hostLayerTiles)

// span[r * nxt + t]: the span of tile t of row r (spanJoin), one wavefront per tile.
__global__ __launch_bounds__(256) void tile_span_kernel(const uint32_t* voxels, int nx, int nxt, size_t tiles, uint16_t* span)
{
    const size_t w = (static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x) >> 6;
    const int lane = static_cast<int>(threadIdx.x & 63);
    if (w >= tiles)
        return;   // the whole wavefront
    const size_t r = w / static_cast<size_t>(nxt);
    const int x = static_cast<int>(w - r * nxt) * 64 + lane;
    const bool body = x < nx && (voxels[r * static_cast<size_t>(nx) + x] & 0xFFu) != 0;
    const unsigned long long m = __ballot(body);
    if (lane == 0)
        span[w] = m ? static_cast<uint16_t>((__ffsll(static_cast<long long>(m)) - 1) | (63 - __clzll(static_cast<long long>(m))) << 8)
                    : kNoSpan;
}

// out = src joined over +-1 along y (axis 0) or z (axis 1).
__global__ __launch_bounds__(256) void tile_span_grow_kernel(const uint16_t* src, uint16_t* out, int nz, int ny, int nxt, int axis)
{
    const size_t i = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
    const size_t plane = static_cast<size_t>(ny) * nxt;
    if (i >= plane * nz)
        return;
    uint16_t v = src[i];
    if (axis == 0) {
        const int y = static_cast<int>((i / nxt) % ny);
        if (y > 0)
            v = spanJoin(v, src[i - nxt]);
        if (y + 1 < ny)
            v = spanJoin(v, src[i + nxt]);
    } else {
        const int z = static_cast<int>(i / plane);
        if (z > 0)
            v = spanJoin(v, src[i - plane]);
        if (z + 1 < nz)
            v = spanJoin(v, src[i + plane]);
    }
    out[i] = v;
}

// flags = spanReaches at reach r
__global__ __launch_bounds__(256) void tile_reach_kernel(const uint16_t* span, uint8_t* flags, size_t tiles, int nxt, int r)
{
    const size_t i = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
    if (i >= tiles)
        return;
    const int t = static_cast<int>(i % nxt);
    flags[i] = spanReaches(span + (i - t), t, nxt, r);
}

// The store mask of row pair g = (z * nyp + yp) * nxt + t: bit o for row 2 yp + o.
__device__ inline int pairMask(const uint8_t* f, size_t g, int ny, int nyp, int nxt)
{
    const int t = static_cast<int>(g % nxt);
    const size_t q = g / nxt;
    const int yp = static_cast<int>(q % nyp);
    const size_t z = q / nyp;
    const size_t row0 = z * ny + 2 * yp;
    int m = f[row0 * nxt + t];
    if (2 * yp + 1 < ny)
        m |= f[(row0 + 1) * nxt + t] << 1;
    return m;
}

// Row pairs with a row to store, per 256-pair block: counts, then the blocks'
// exclusive offsets (one work-group, in order), then the entries (g << 2 | mask)
// written in ascending order.
__global__ __launch_bounds__(256) void pair_count_kernel(const uint8_t* f, size_t groups, int ny, int nyp, int nxt, int* counts)
{
    const size_t g = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
    const int keep = g < groups && pairMask(f, g, ny, nyp, nxt) != 0;
    const int n = __syncthreads_count(keep);
    if (threadIdx.x == 0)
        counts[blockIdx.x] = n;
}

__global__ __launch_bounds__(256) void pair_scan_kernel(const int* counts, int blocks, int* offsets, int* total)
{
    __shared__ int s_wave[4];
    __shared__ int s_base;
    if (threadIdx.x == 0)
        s_base = 0;
    __syncthreads();
    const int lane = static_cast<int>(threadIdx.x & 63), wave = static_cast<int>(threadIdx.x >> 6);
    for (int b0 = 0; b0 < blocks; b0 += 256) {
        const int b = b0 + static_cast<int>(threadIdx.x);
        const int c = b < blocks ? counts[b] : 0;
        // inclusive scan within the wavefront
        int x = c;
        for (int d = 1; d < 64; d <<= 1) {
            const int y = __shfl_up(x, d);
            if (lane >= d)
                x += y;
        }
        if (lane == 63)
            s_wave[wave] = x;
        __syncthreads();
        int before = s_base;
        for (int k = 0; k < wave; ++k)
            before += s_wave[k];
        if (b < blocks)
            offsets[b] = before + x - c;
        __syncthreads();
        if (threadIdx.x == 255)
            s_base = before + x;
        __syncthreads();
    }
    if (threadIdx.x == 0)
        *total = s_base;
}

__global__ __launch_bounds__(256) void pair_write_kernel(const uint8_t* f, size_t groups, int ny, int nyp, int nxt, const int* offsets,
                                                         int* list)
{
    __shared__ int s_wave[4];
    const size_t g = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
    const int m = g < groups ? pairMask(f, g, ny, nyp, nxt) : 0;
    const unsigned long long bits = __ballot(m != 0);
    const int lane = static_cast<int>(threadIdx.x & 63), wave = static_cast<int>(threadIdx.x >> 6);
    if (lane == 0)
        s_wave[wave] = __popcll(bits);
    __syncthreads();
    int base = offsets[blockIdx.x];
    for (int k = 0; k < wave; ++k)
        base += s_wave[k];
    if (m != 0)
        list[base + __popcll(bits & ((1ull << lane) - 1ull))] = static_cast<int>(g << kConvRows) | m;
}

template <typename T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() { release(); }
    void release()
    {
        if (p)
            (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    hipError_t reserve(size_t count)
    {
        if (count <= n && p)
            return hipSuccess;
        release();
        const hipError_t e = hipMalloc(reinterpret_cast<void**>(&p), std::max<size_t>(count, 1) * sizeof(T));
        if (e == hipSuccess)
            n = count;
        return e;
    }
};

}  // namespace

struct dxg_denoiser {
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    DBuf<float> params, adam_m, adam_v, grads;
    DBuf<uint16_t> wf, wd;
    DBuf<uint16_t> xin, xin4, acts, gbuf;
    DBuf<float> feat0, y32, target, mask, features, lpart, part, part_b, sums, wemb;
    DBuf<double> wacc, wdose;
    DBuf<int> tiles;                    // kLayers lists of up to tiles_stride entries, layer l's at l * tiles_stride
    size_t tiles_stride = 0;
    int layer_tiles[kLayers] = {};      // entries of each layer's list
    DBuf<uint16_t> tspan_a, tspan_b;   // device tile lists: tile spans (before / after a join along y)
    DBuf<uint8_t> tflag;               // tiles a layer computes
    DBuf<int> tcounts, toffsets, ttotal;
    uint64_t tiles_generation = 0;   // world generation the tile list belongs to
    bool tiles_zeroed = false;       // acts and y32 are zero outside the tile list ...
    const void* zeroed_acts = nullptr;   // ... in these allocations
    const void* zeroed_y32 = nullptr;
    uint64_t step = 0;
    bool weights_ready = false;
    double last_ms = 0.0;

    ~dxg_denoiser()
    {
        if (ev0)
            (void)hipEventDestroy(ev0);
        if (ev1)
            (void)hipEventDestroy(ev1);
        if (stream)
            (void)hipStreamDestroy(stream);
    }

    int prepWeights()
    {
        const int n = kLayers * kLayerW;
        hipLaunchKernelGGL(prep_weights_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, params.p, wf.p, wd.p, kLayers);
        HIP_TRY(hipGetLastError());
        weights_ready = true;
        return DXG_OK;
    }

    const uint16_t* layerW(int l) const { return wf.p + static_cast<size_t>(l) * kLayerW; }
    const uint16_t* layerWd(int l) const { return wd.p + static_cast<size_t>(l) * kLayerW; }
    const float* layerBias(int l) const { return params.p + static_cast<size_t>(l) * kLayerP + kLayerW; }
    float* layerGrad(int l) const { return grads.p + static_cast<size_t>(l) * kLayerP; }

    // Stages the features on the device and packs them.
    int upload(const float* features_host, const Dims& d)
    {
        const size_t nv = d.voxels();
        HIP_TRY(features.reserve(4 * nv));
        HIP_TRY(xin.reserve(nv * kC));
        HIP_TRY(xin4.reserve(nv * 4));
        HIP_TRY(feat0.reserve(nv));
        HIP_TRY(y32.reserve(nv));
        HIP_TRY(hipMemcpyAsync(features.p, features_host, 4 * nv * sizeof(float), hipMemcpyHostToDevice, stream));
        hipLaunchKernelGGL(pack_features_kernel, dim3((nv + 255) / 256), dim3(256), 0, stream, features.p, xin.p, xin4.p, feat0.p, nv);
        HIP_TRY(hipGetLastError());
        return DXG_OK;
    }

    // Forward pass.  With keep, activations a_0..a_8 and h_0..h_7 are kept
    // for the backward pass (17 tensors); otherwise two tensors are reused.
    // Tiles the convolutions compute (off: all); set by apply for one pass.
    bool use_tiles = false;
    void setTiles(ConvArgs& c, int layer) const
    {
        if (use_tiles) {
            c.tiles = tiles.p + static_cast<size_t>(layer) * tiles_stride;
            c.wave_tiles = layer_tiles[layer];
        }
    }

    int forward(const Dims& d, bool keep)
    {
        if (!use_tiles)
            tiles_zeroed = false;   // a full pass writes every tile
        const size_t nv = d.voxels(), ts = nv * kC;
        HIP_TRY(acts.reserve((keep ? 17 : 2) * ts));
        auto aT = [&](int b) { return acts.p + (keep ? static_cast<size_t>(b) : 0) * ts; };
        auto hT = [&](int b) { return acts.p + (keep ? static_cast<size_t>(9 + b) : 1) * ts; };
        ConvArgs c {};
        c.in4 = firstLayerCompact(d);   // the four features, packed over the taps
        c.x = c.in4 ? xin4.p : xin.p;
        c.w = layerW(0);
        c.bias = layerBias(0);
        c.y = aT(0);
        c.relu = 1;
        setTiles(c, 0);
        HIP_TRY(launchConv(c, d, stream));
        for (int b = 0; b < kBlocks; ++b) {
            c = ConvArgs {};
            c.x = aT(b);
            c.w = layerW(2 * b + 1);
            c.bias = layerBias(2 * b + 1);
            c.y = hT(b);
            c.relu = 1;
            setTiles(c, 2 * b + 1);
            HIP_TRY(launchConv(c, d, stream));
            c = ConvArgs {};
            c.x = hT(b);
            c.w = layerW(2 * b + 2);
            c.bias = layerBias(2 * b + 2);
            c.res = aT(b);
            c.y = aT(b + 1);
            c.relu = 1;
            setTiles(c, 2 * b + 2);
            HIP_TRY(launchConv(c, d, stream));
        }
        c = ConvArgs {};
        c.x = aT(kBlocks);
        c.w = layerW(kLayers - 1);
        c.bias = layerBias(kLayers - 1);
        c.y32 = y32.p;
        c.y32_channels = 1;
        setTiles(c, kLayers - 1);
        HIP_TRY(launchConv(c, d, stream));
        return DXG_OK;
    }

    // The row-pair lists of the world's body tiles for every layer, built on
    // the device: the same lists as pairTiles of hostLayerTiles.  Tile spans
    // from the material bytes of the voxel words, joined over +-1 row and
    // plane once per unit of reach, and at each layer's reach the x rule and
    // an ordered compaction of the pairs with a row to store.
    int devicePairTiles(const dxg::WorldView& v, const Dims& dm)
    {
        static_assert(kConvRows == 2, "pair masks of two rows");
        const int nxt = dm.nx / 64, nyp = (v.ny + 1) / 2;
        const size_t ntiles = static_cast<size_t>(v.nz) * v.ny * nxt;
        const size_t groups = static_cast<size_t>(v.nz) * nyp * nxt;
        const int blocks = static_cast<int>((groups + 255) / 256);
        tiles_stride = std::max<size_t>(groups, 1);
        HIP_TRY(tspan_a.reserve(ntiles));
        HIP_TRY(tspan_b.reserve(ntiles));
        HIP_TRY(tflag.reserve(ntiles));
        HIP_TRY(tcounts.reserve(static_cast<size_t>(blocks)));
        HIP_TRY(toffsets.reserve(static_cast<size_t>(blocks)));
        HIP_TRY(ttotal.reserve(kLayers));
        HIP_TRY(tiles.reserve(kLayers * tiles_stride));
        hipLaunchKernelGGL(tile_span_kernel, dim3(static_cast<unsigned>((ntiles * 64 + 255) / 256)), dim3(256), 0, stream, v.voxels,
                           v.nx, nxt, ntiles, tspan_a.p);
        const dim3 tgrid(static_cast<unsigned>((ntiles + 255) / 256));
        for (int reach = 1; reach <= kReach; ++reach) {
            hipLaunchKernelGGL(tile_span_grow_kernel, tgrid, dim3(256), 0, stream, tspan_a.p, tspan_b.p, v.nz, v.ny, nxt, 0);
            hipLaunchKernelGGL(tile_span_grow_kernel, tgrid, dim3(256), 0, stream, tspan_b.p, tspan_a.p, v.nz, v.ny, nxt, 1);
            const int l = kReach - reach;
            if (l < 0 || l >= kLayers)
                continue;
            hipLaunchKernelGGL(tile_reach_kernel, tgrid, dim3(256), 0, stream, tspan_a.p, tflag.p, ntiles, nxt, reach);
            hipLaunchKernelGGL(pair_count_kernel, dim3(blocks), dim3(256), 0, stream, tflag.p, groups, v.ny, nyp, nxt, tcounts.p);
            hipLaunchKernelGGL(pair_scan_kernel, dim3(1), dim3(256), 0, stream, tcounts.p, blocks, toffsets.p, ttotal.p + l);
            hipLaunchKernelGGL(pair_write_kernel, dim3(blocks), dim3(256), 0, stream, tflag.p, groups, v.ny, nyp, nxt, toffsets.p,
                               tiles.p + static_cast<size_t>(l) * tiles_stride);
        }
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(layer_tiles, ttotal.p, kLayers * sizeof(int), hipMemcpyDeviceToHost, stream));
        HIP_TRY(hipStreamSynchronize(stream));
        return DXG_OK;
    }

    // The body-tile lists the convolutions take: on the device for row pairs
    // unless DXG_DENOISE_TILES_HOST=1 (or always on the host for single rows).
    int buildTiles(const dxg::WorldView& v, const Dims& dm, bool host)
    {
        if (!host && convPairs())
            return devicePairTiles(v, dm);
        std::vector<int> t[kLayers];
        hostLayerTiles(v.material_host, v.nx, v.ny, v.nz, dm.nx, t);
        size_t most = 1;
        for (int l = 0; l < kLayers; ++l) {
            if (convPairs())
                t[l] = pairTiles(t[l], dm.ny, dm.nx / 64);
            most = std::max(most, t[l].size());
        }
        tiles_stride = most;
        HIP_TRY(tiles.reserve(kLayers * tiles_stride));
        for (int l = 0; l < kLayers; ++l) {
            if (!t[l].empty())
                HIP_TRY(hipMemcpyAsync(tiles.p + static_cast<size_t>(l) * tiles_stride, t[l].data(), t[l].size() * sizeof(int),
                                       hipMemcpyHostToDevice, stream));
            layer_tiles[l] = static_cast<int>(t[l].size());
        }
        HIP_TRY(hipStreamSynchronize(stream));   // the host vectors go out of scope
        return DXG_OK;
    }

    // Features of a built world's current scores into xin / feat0 (and the
    // FP32 features buffer when asked); dm receives the padded volume.
    int worldFeatures(const dxg_world* w, dxg::WorldView& v, Dims& dm, bool keep_fp32)
    {
        if (int rc = dxg::worldView(w, v); rc != DXG_OK)
            return rc;
        int dev = 0;
        HIP_TRY(hipGetDevice(&dev));
        if (dev != v.device)
            return fail(DXG_ERR_ARGUMENT, "denoiser and world live on different devices");
        if (v.embedding.empty())
            return fail(DXG_ERR_STATE, "world has no materials");
        const int nxp = (v.nx + 63) / 64 * 64;
        dm = { 1, v.nz, v.ny, nxp };
        const size_t nvp = dm.voxels();
        const size_t nv = static_cast<size_t>(v.nx) * v.ny * v.nz;
        // the first layer reads the compact copy alone unless it runs the general kernel
        const bool full_input = !firstLayerCompact(dm);
        if (full_input)
            HIP_TRY(xin.reserve(nvp * kC));
        HIP_TRY(xin4.reserve(nvp * 4));
        HIP_TRY(feat0.reserve(nvp));
        HIP_TRY(y32.reserve(nvp));
        HIP_TRY(wacc.reserve(2));
        HIP_TRY(wemb.reserve(v.embedding.size()));
        if (keep_fp32)
            HIP_TRY(features.reserve(4 * nvp));
        HIP_TRY(hipMemcpyAsync(wemb.p, v.embedding.data(), v.embedding.size() * sizeof(float), hipMemcpyHostToDevice, stream));
        hipLaunchKernelGGL(zero2_kernel, dim3(1), dim3(64), 0, stream, wacc.p);
        const int mblocks = static_cast<int>(std::min<size_t>((nv + 255) / 256, 4096));
        hipLaunchKernelGGL(dose_mean_kernel, dim3(mblocks), dim3(256), 0, stream, v.voxels, v.tally, nv, wacc.p);
        WorldFeatArgs a {};
        a.voxels = v.voxels;
        a.tally = v.tally;
        a.emb = wemb.p;
        a.n_emb = static_cast<int>(v.embedding.size());
        a.nx = v.nx;
        a.nxp = nxp;
        a.nvp = nvp;
        a.histories = static_cast<double>(v.histories);
        a.acc = wacc.p;
        a.xin = full_input ? xin.p : nullptr;
        a.xin4 = xin4.p;
        a.feat0 = feat0.p;
        a.features = keep_fp32 ? features.p : nullptr;
        hipLaunchKernelGGL(world_features_kernel, dim3((nvp + 255) / 256), dim3(256), 0, stream, a);
        HIP_TRY(hipGetLastError());
        return DXG_OK;
    }

    // Backward pass from the loss gradient in gbuf tensor 0; writes every
    // layer's gradient into grads.
    int backward(const Dims& d)
    {
        const size_t ts = d.voxels() * kC;
        auto aT = [&](int b) { return acts.p + static_cast<size_t>(b) * ts; };
        auto hT = [&](int b) { return acts.p + static_cast<size_t>(9 + b) * ts; };
        uint16_t* gout = gbuf.p;
        uint16_t* gs = gbuf.p + ts;
        uint16_t* gh = gbuf.p + 2 * ts;
        const int grid = wgradGrid(d);
        HIP_TRY(part.reserve(static_cast<size_t>(grid) * kLayerW));
        HIP_TRY(part_b.reserve(static_cast<size_t>(grid) * 32));
        HIP_TRY(sums.reserve(static_cast<size_t>(kReduceSlices) * kLayerP));

        HIP_TRY(launchWgrad(aT(kBlocks), gout, part.p, part_b.p, sums.p, layerGrad(kLayers - 1), d, stream));
        ConvArgs c {};
        c.x = gout;
        c.w = layerWd(kLayers - 1);
        c.gate = aT(kBlocks);
        c.y = gs;
        HIP_TRY(launchConv(c, d, stream));
        for (int b = kBlocks - 1; b >= 0; --b) {
            HIP_TRY(launchWgrad(hT(b), gs, part.p, part_b.p, sums.p, layerGrad(2 * b + 2), d, stream));
            c = ConvArgs {};
            c.x = gs;
            c.w = layerWd(2 * b + 2);
            c.gate = hT(b);
            c.y = gh;
            HIP_TRY(launchConv(c, d, stream));
            HIP_TRY(launchWgrad(aT(b), gh, part.p, part_b.p, sums.p, layerGrad(2 * b + 1), d, stream));
            c = ConvArgs {};
            c.x = gh;
            c.w = layerWd(2 * b + 1);
            c.res = gs;
            c.gate = aT(b);
            c.y = gs;
            HIP_TRY(launchConv(c, d, stream));
        }
        HIP_TRY(launchWgrad(xin.p, gs, part.p, part_b.p, sums.p, layerGrad(0), d, stream));
        return DXG_OK;
    }
};

namespace {

void initParams(std::vector<float>& p, uint64_t seed)
{
    p.assign(static_cast<size_t>(kLayers) * kLayerP, 0.0f);
    std::mt19937_64 rng(seed);
    for (int l = 0; l < kLayers - 1; ++l) {
        const int ci_active = l == 0 ? DXG_DENOISE_FEATURES : kC;
        // He initialisation; the second convolution of each residual block
        // starts small so every block begins close to the identity.
        const bool second = l > 0 && (l % 2) == 0;
        std::normal_distribution<float> nd(0.0f, (second ? 0.1f : 1.0f) * std::sqrt(2.0f / (kTaps * ci_active)));
        float* w = p.data() + static_cast<size_t>(l) * kLayerP;
        for (int co = 0; co < kC; ++co)
            for (int tap = 0; tap < kTaps; ++tap)
                for (int ci = 0; ci < ci_active; ++ci)
                    w[(co * kTaps + tap) * kC + ci] = nd(rng);
    }
    // the correction layer starts at zero: the untrained network is the identity on the log dose
}

bool hasDevice()
{
    int n = 0;
    return hipGetDeviceCount(&n) == hipSuccess && n > 0;
}

}  // namespace

extern "C" {

uint64_t dxg_denoiser_num_params(void) { return static_cast<uint64_t>(kLayers) * kLayerP; }

int dxg_denoiser_create(dxg_denoiser** out, uint64_t seed)
{
    if (!out)
        return fail(DXG_ERR_ARGUMENT, "dxg_denoiser_create: null output pointer");
    *out = nullptr;
    if (!hasDevice())
        return fail(DXG_ERR_DEVICE, "dxg_denoiser_create: no HIP device");
    auto* d = new (std::nothrow) dxg_denoiser();
    if (!d)
        return fail(DXG_ERR_DEVICE, "dxg_denoiser_create: out of host memory");
    auto cleanup = [&](int code) {
        delete d;
        return code;
    };
    const size_t np = dxg_denoiser_num_params();
    std::vector<float> host;
    initParams(host, seed);
#define TRY_CLEAN(expr)                                                                                   \
    do {                                                                                                  \
        const hipError_t e_ = (expr);                                                                     \
        if (e_ != hipSuccess)                                                                             \
            return cleanup(fail(DXG_ERR_DEVICE, std::string(#expr ": ") + hipGetErrorString(e_)));        \
    } while (0)
    TRY_CLEAN(hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking));
    TRY_CLEAN(hipEventCreate(&d->ev0));
    TRY_CLEAN(hipEventCreate(&d->ev1));
    TRY_CLEAN(d->params.reserve(np));
    TRY_CLEAN(d->adam_m.reserve(np));
    TRY_CLEAN(d->adam_v.reserve(np));
    TRY_CLEAN(d->grads.reserve(np));
    TRY_CLEAN(d->wf.reserve(static_cast<size_t>(kLayers) * kLayerW));
    TRY_CLEAN(d->wd.reserve(static_cast<size_t>(kLayers) * kLayerW));
    TRY_CLEAN(hipMemcpyAsync(d->params.p, host.data(), np * sizeof(float), hipMemcpyHostToDevice, d->stream));
    TRY_CLEAN(hipMemsetAsync(d->adam_m.p, 0, np * sizeof(float), d->stream));
    TRY_CLEAN(hipMemsetAsync(d->adam_v.p, 0, np * sizeof(float), d->stream));
    TRY_CLEAN(hipMemsetAsync(d->grads.p, 0, np * sizeof(float), d->stream));
    TRY_CLEAN(hipStreamSynchronize(d->stream));
#undef TRY_CLEAN
    *out = d;
    return DXG_OK;
}

void dxg_denoiser_destroy(dxg_denoiser* d)
{
    if (d && d->stream)
        (void)hipStreamSynchronize(d->stream);
    delete d;
}

int dxg_denoiser_get_params(const dxg_denoiser* d, float* params, uint64_t n)
{
    if (!d || !params)
        return fail(DXG_ERR_ARGUMENT, "dxg_denoiser_get_params: null argument");
    if (n != dxg_denoiser_num_params())
        return fail(DXG_ERR_ARGUMENT, "dxg_denoiser_get_params: wrong parameter count");
    HIP_TRY(hipMemcpyAsync(params, d->params.p, n * sizeof(float), hipMemcpyDeviceToHost, d->stream));
    HIP_TRY(hipStreamSynchronize(d->stream));
    return DXG_OK;
}

int dxg_denoiser_get_grads(const dxg_denoiser* d, float* grads, uint64_t n)
{
    if (!d || !grads)
        return fail(DXG_ERR_ARGUMENT, "dxg_denoiser_get_grads: null argument");
    if (n != dxg_denoiser_num_params())
        return fail(DXG_ERR_ARGUMENT, "dxg_denoiser_get_grads: wrong parameter count");
    HIP_TRY(hipMemcpyAsync(grads, d->grads.p, n * sizeof(float), hipMemcpyDeviceToHost, d->stream));
    HIP_TRY(hipStreamSynchronize(d->stream));
    return DXG_OK;
}

int dxg_denoiser_set_params(dxg_denoiser* d, const float* params, uint64_t n)
{
    if (!d || !params)
        return fail(DXG_ERR_ARGUMENT, "dxg_denoiser_set_params: null argument");
    if (n != dxg_denoiser_num_params())
        return fail(DXG_ERR_ARGUMENT, "dxg_denoiser_set_params: wrong parameter count");
    HIP_TRY(hipMemcpyAsync(d->params.p, params, n * sizeof(float), hipMemcpyHostToDevice, d->stream));
    HIP_TRY(hipMemsetAsync(d->adam_m.p, 0, n * sizeof(float), d->stream));
    HIP_TRY(hipMemsetAsync(d->adam_v.p, 0, n * sizeof(float), d->stream));
    HIP_TRY(hipStreamSynchronize(d->stream));
    d->step = 0;
    d->weights_ready = false;
    return DXG_OK;
}

int dxg_denoiser_train_step(dxg_denoiser* d, const float* features, const float* target, const float* mask,
                            uint64_t batch, const uint64_t dims[3], float lr, float* loss)
{
    if (!d || !features || !target || !mask || !loss)
        return fail(DXG_ERR_ARGUMENT, "dxg_denoiser_train_step: null argument");
    Dims dm {};
    if (int rc = checkDims("dxg_denoiser_train_step", batch, dims, dm); rc != DXG_OK)
        return rc;
    if (dm.ny % 4 != 0)
        return fail(DXG_ERR_ARGUMENT, "dxg_denoiser_train_step: the y extent must be a multiple of 4");
    if (!(lr > 0.0f) || !std::isfinite(lr))
        return fail(DXG_ERR_ARGUMENT, "dxg_denoiser_train_step: learning rate must be positive");
    const size_t nv = dm.voxels();
    double count = 0.0;
    for (size_t i = 0; i < nv; ++i)
        count += mask[i];
    if (!(count > 0.0))
        return fail(DXG_ERR_ARGUMENT, "dxg_denoiser_train_step: empty mask");

    HIP_TRY(hipEventRecord(d->ev0, d->stream));
    if (int rc = d->upload(features, dm); rc != DXG_OK)
        return rc;
    HIP_TRY(d->target.reserve(nv));
    HIP_TRY(d->mask.reserve(nv));
    HIP_TRY(hipMemcpyAsync(d->target.p, target, nv * sizeof(float), hipMemcpyHostToDevice, d->stream));
    HIP_TRY(hipMemcpyAsync(d->mask.p, mask, nv * sizeof(float), hipMemcpyHostToDevice, d->stream));
    if (int rc = d->prepWeights(); rc != DXG_OK)
        return rc;
    if (int rc = d->forward(dm, true); rc != DXG_OK)
        return rc;
    HIP_TRY(d->gbuf.reserve(3 * nv * kC));
    const int lblocks = static_cast<int>((nv + 255) / 256);
    HIP_TRY(d->lpart.reserve(lblocks));
    hipLaunchKernelGGL(loss_kernel, dim3(lblocks), dim3(256), 0, d->stream, d->y32.p, d->feat0.p, d->target.p, d->mask.p,
                       static_cast<float>(1.0 / count), d->gbuf.p, d->lpart.p, nv);
    HIP_TRY(hipGetLastError());
    if (int rc = d->backward(dm); rc != DXG_OK)
        return rc;
    d->step += 1;
    const double b1 = 0.9, b2 = 0.999;
    const float bc1 = static_cast<float>(1.0 / (1.0 - std::pow(b1, static_cast<double>(d->step))));
    const float bc2 = static_cast<float>(1.0 / (1.0 - std::pow(b2, static_cast<double>(d->step))));
    const int np = static_cast<int>(dxg_denoiser_num_params());
    hipLaunchKernelGGL(adam_kernel, dim3((np + 255) / 256), dim3(256), 0, d->stream, d->params.p, d->adam_m.p, d->adam_v.p,
                       d->grads.p, np, lr, static_cast<float>(b1), static_cast<float>(b2), 1e-8f, bc1, bc2);
    HIP_TRY(hipGetLastError());
    d->weights_ready = false;
    HIP_TRY(hipEventRecord(d->ev1, d->stream));
    std::vector<float> lp(lblocks);
    HIP_TRY(hipMemcpyAsync(lp.data(), d->lpart.p, lblocks * sizeof(float), hipMemcpyDeviceToHost, d->stream));
    HIP_TRY(hipStreamSynchronize(d->stream));
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, d->ev0, d->ev1));
    d->last_ms = ms;
    double s = 0.0;
    for (float v : lp)
        s += v;
    *loss = static_cast<float>(s / count);
    return DXG_OK;
}

int dxg_denoiser_infer(dxg_denoiser* d, const float* features, const uint64_t dims[3], float* prediction)
{
    if (!d || !features || !prediction)
        return fail(DXG_ERR_ARGUMENT, "dxg_denoiser_infer: null argument");
    Dims dm {};
    if (int rc = checkDims("dxg_denoiser_infer", 1, dims, dm); rc != DXG_OK)
        return rc;
    const size_t nv = dm.voxels();
    if (int rc = d->upload(features, dm); rc != DXG_OK)
        return rc;
    HIP_TRY(hipEventRecord(d->ev0, d->stream));
    if (!d->weights_ready)
        if (int rc = d->prepWeights(); rc != DXG_OK)
            return rc;
    if (int rc = d->forward(dm, false); rc != DXG_OK)
        return rc;
    hipLaunchKernelGGL(finalize_kernel, dim3((nv + 255) / 256), dim3(256), 0, d->stream, d->y32.p, d->feat0.p, d->y32.p, nv);
    d->tiles_zeroed = false;
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(d->ev1, d->stream));
    HIP_TRY(hipMemcpyAsync(prediction, d->y32.p, nv * sizeof(float), hipMemcpyDeviceToHost, d->stream));
    HIP_TRY(hipStreamSynchronize(d->stream));
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, d->ev0, d->ev1));
    d->last_ms = ms;
    return DXG_OK;
}

double dxg_denoiser_last_ms(const dxg_denoiser* d) { return d ? d->last_ms : 0.0; }

int dxg_denoiser_world_features(dxg_denoiser* d, const dxg_world* w, float* features, uint64_t n)
{
    if (!d || !w || !features)
        return fail(DXG_ERR_ARGUMENT, "dxg_denoiser_world_features: null argument");
    dxg::WorldView v;
    Dims dm {};
    if (int rc = d->worldFeatures(w, v, dm, true); rc != DXG_OK)
        return rc;
    if (n != 4 * dm.voxels())
        return fail(DXG_ERR_ARGUMENT, "dxg_denoiser_world_features: n must be 4 x nz x ny x (nx padded to 64)");
    HIP_TRY(hipMemcpyAsync(features, d->features.p, n * sizeof(float), hipMemcpyDeviceToHost, d->stream));
    HIP_TRY(hipStreamSynchronize(d->stream));
    return DXG_OK;
}

int dxg_denoiser_apply(dxg_denoiser* d, const dxg_world* w, double* dose_mgy)
{
    if (!d || !w)
        return fail(DXG_ERR_ARGUMENT, "dxg_denoiser_apply: null argument");
    HIP_TRY(hipEventRecord(d->ev0, d->stream));
    dxg::WorldView v;
    Dims dm {};
    if (int rc = d->worldFeatures(w, v, dm, false); rc != DXG_OK)
        return rc;
    if (!d->weights_ready)
        if (int rc = d->prepWeights(); rc != DXG_OK)
            return rc;
    const size_t nvp = dm.voxels();
    const size_t nv = static_cast<size_t>(v.nx) * v.ny * v.nz;
    // only the tiles within the network's reach of the body (DXG_DENOISE_SKIP=0: all)
    const char* env = std::getenv("DXG_DENOISE_SKIP");
    const bool skip = !(env && env[0] == '0');
    if (skip) {
        if (d->tiles_generation != v.generation || !d->tiles.p) {
            const char* th = std::getenv("DXG_DENOISE_TILES_HOST");
            if (int rc = d->buildTiles(v, dm, th && th[0] == '1'); rc != DXG_OK)
                return rc;
            d->tiles_generation = v.generation;
            d->tiles_zeroed = false;
        }
        const size_t ts = nvp * kC;
        HIP_TRY(d->acts.reserve(2 * ts));
        if (!d->tiles_zeroed || d->zeroed_acts != d->acts.p || d->zeroed_y32 != d->y32.p) {
            HIP_TRY(hipMemsetAsync(d->acts.p, 0, 2 * ts * sizeof(uint16_t), d->stream));
            HIP_TRY(hipMemsetAsync(d->y32.p, 0, nvp * sizeof(float), d->stream));
            d->tiles_zeroed = true;
            d->zeroed_acts = d->acts.p;
            d->zeroed_y32 = d->y32.p;
        }
        d->use_tiles = true;
    }
    const int frc = d->layer_tiles[0] > 0 || !skip ? d->forward(dm, false) : DXG_OK;
    d->use_tiles = false;
    if (frc != DXG_OK)
        return frc;
    HIP_TRY(d->wdose.reserve(nv));
    const double voxel_cm3 = v.spacing[0] * v.spacing[1] * v.spacing[2];
    const double scale = dxg::kKeVToJ / (voxel_cm3 * 1e-3) * 1e3;   // keV per g/cm3 -> mGy
    hipLaunchKernelGGL(world_dose_kernel, dim3((nv + 255) / 256), dim3(256), 0, d->stream, d->feat0.p, d->y32.p, v.nx, dm.nx,
                       nv, d->wacc.p, scale, d->wdose.p);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(d->ev1, d->stream));
    if (dose_mgy)
        HIP_TRY(hipMemcpyAsync(dose_mgy, d->wdose.p, nv * sizeof(double), hipMemcpyDeviceToHost, d->stream));
    HIP_TRY(hipStreamSynchronize(d->stream));
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, d->ev0, d->ev1));
    d->last_ms = ms;
    return DXG_OK;
}

int dxg_denoiser_debug_tiles(dxg_denoiser* d, const dxg_world* w, int host, int layer, int32_t* out, uint64_t cap,
                             uint64_t* count)
{
    if (!d || !w || !count || (!out && cap))
        return fail(DXG_ERR_ARGUMENT, "dxg_denoiser_debug_tiles: null argument");
    if (layer < 0 || layer >= kLayers)
        return fail(DXG_ERR_ARGUMENT, "dxg_denoiser_debug_tiles: layer must be 0..17");
    dxg::WorldView v;
    if (int rc = dxg::worldView(w, v); rc != DXG_OK)
        return rc;
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev != v.device)
        return fail(DXG_ERR_ARGUMENT, "denoiser and world live on different devices");
    const Dims dm { 1, v.nz, v.ny, (v.nx + 63) / 64 * 64 };
    if (int rc = d->buildTiles(v, dm, host != 0); rc != DXG_OK)
        return rc;
    d->tiles_generation = ~0ull;   // the next apply builds its own list
    *count = static_cast<uint64_t>(d->layer_tiles[layer]);
    const size_t k = std::min<size_t>(cap, static_cast<size_t>(d->layer_tiles[layer]));
    if (k)
        HIP_TRY(hipMemcpyAsync(out, d->tiles.p + static_cast<size_t>(layer) * d->tiles_stride, k * sizeof(int),
                               hipMemcpyDeviceToHost, d->stream));
    HIP_TRY(hipStreamSynchronize(d->stream));
    return DXG_OK;
}

int dxg_conv3d_probe(int mode, const float* x, const float* w, const float* bias_or_dy, uint64_t n, const uint64_t dims[3],
                     float* out)
{
    if (!x || !w || !out || (mode != 1 && !bias_or_dy))
        return fail(DXG_ERR_ARGUMENT, "dxg_conv3d_probe: null argument");
    if (mode < 0 || mode > 3)
        return fail(DXG_ERR_ARGUMENT, "dxg_conv3d_probe: mode must be 0, 1, 2 or 3");
    Dims dm {};
    if (int rc = checkDims("dxg_conv3d_probe", n, dims, dm); rc != DXG_OK)
        return rc;
    if (mode == 2 && dm.ny % 4 != 0)
        return fail(DXG_ERR_ARGUMENT, "dxg_conv3d_probe: the y extent must be a multiple of 4");
    if (!hasDevice())
        return fail(DXG_ERR_DEVICE, "dxg_conv3d_probe: no HIP device");
    const size_t nv = dm.voxels(), ts = nv * kC;
    std::vector<uint16_t> xh(ts);
    for (size_t i = 0; i < ts; ++i)
        xh[i] = toBf16(x[i]);
    std::vector<float> layer(kLayerP, 0.0f);
    std::copy(w, w + kLayerW, layer.begin());
    if (mode == 0 || mode == 3)
        std::copy(bias_or_dy, bias_or_dy + kC, layer.begin() + kLayerW);
    DBuf<uint16_t> dx, dy, wf, wd;
    DBuf<float> dp, dout, part, part_b;
    HIP_TRY(dx.reserve(ts));
    HIP_TRY(dp.reserve(kLayerP));
    HIP_TRY(wf.reserve(kLayerW));
    HIP_TRY(wd.reserve(kLayerW));
    HIP_TRY(hipMemcpy(dx.p, xh.data(), ts * 2, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dp.p, layer.data(), kLayerP * sizeof(float), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(prep_weights_kernel, dim3((kLayerW + 255) / 256), dim3(256), 0, nullptr, dp.p, wf.p, wd.p, 1);
    HIP_TRY(hipGetLastError());
    if (mode == 2) {
        std::vector<uint16_t> gh(ts);
        for (size_t i = 0; i < ts; ++i)
            gh[i] = toBf16(bias_or_dy[i]);
        HIP_TRY(dy.reserve(ts));
        HIP_TRY(hipMemcpy(dy.p, gh.data(), ts * 2, hipMemcpyHostToDevice));
        const int grid = wgradGrid(dm);
        HIP_TRY(part.reserve(static_cast<size_t>(grid) * kLayerW));
        HIP_TRY(part_b.reserve(static_cast<size_t>(grid) * 32));
        HIP_TRY(dout.reserve(kLayerP));
        DBuf<float> sums;
        HIP_TRY(sums.reserve(static_cast<size_t>(kReduceSlices) * kLayerP));
        HIP_TRY(launchWgrad(dx.p, dy.p, part.p, part_b.p, sums.p, dout.p, dm, nullptr));
        HIP_TRY(hipMemcpy(out, dout.p, kLayerP * sizeof(float), hipMemcpyDeviceToHost));
        return DXG_OK;
    }
    if (mode == 3) {
        // the network's first layer: channels 4..31 of x must be zero
        for (size_t i = 0; i < ts; ++i)
            if (i % kC >= 4 && x[i] != 0.0f)
                return fail(DXG_ERR_ARGUMENT, "dxg_conv3d_probe: mode 3 needs channels 4..31 of x to be zero");
        if (!firstLayerCompact(dm))
            return fail(DXG_ERR_ARGUMENT, "dxg_conv3d_probe: mode 3 needs row pairs and a plane below 2^27 voxels");
        std::vector<uint16_t> x4(nv * 4);
        for (size_t v = 0; v < nv; ++v)
            for (int k = 0; k < 4; ++k)
                x4[4 * v + k] = xh[v * kC + k];
        DBuf<uint16_t> yb, dx4;
        HIP_TRY(yb.reserve(ts));
        HIP_TRY(dx4.reserve(nv * 4));
        HIP_TRY(hipMemcpy(dx4.p, x4.data(), nv * 4 * sizeof(uint16_t), hipMemcpyHostToDevice));
        ConvArgs c {};
        c.x = dx4.p;
        c.w = wf.p;
        c.bias = dp.p + kLayerW;
        c.y = yb.p;
        c.relu = 1;
        c.in4 = 1;
        HIP_TRY(launchConv(c, dm, nullptr));
        std::vector<uint16_t> yh(ts);
        HIP_TRY(hipMemcpy(yh.data(), yb.p, ts * sizeof(uint16_t), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < ts; ++i)
            out[i] = fromBf16(yh[i]);
        return DXG_OK;
    }
    HIP_TRY(dout.reserve(ts));
    ConvArgs c {};
    c.x = dx.p;
    c.w = mode == 0 ? wf.p : wd.p;
    c.bias = mode == 0 ? dp.p + kLayerW : nullptr;
    c.y32 = dout.p;
    c.y32_channels = kC;
    HIP_TRY(launchConv(c, dm, nullptr));
    HIP_TRY(hipMemcpy(out, dout.p, ts * sizeof(float), hipMemcpyDeviceToHost));
    return DXG_OK;
}

int dxg_mfma_probe(const float* a, const float* b, float* d)
{
    if (!a || !b || !d)
        return fail(DXG_ERR_ARGUMENT, "dxg_mfma_probe: null argument");
    if (!hasDevice())
        return fail(DXG_ERR_DEVICE, "dxg_mfma_probe: no HIP device");
    DBuf<float> da, db, dd;
    HIP_TRY(da.reserve(16 * 32));
    HIP_TRY(db.reserve(32 * 16));
    HIP_TRY(dd.reserve(16 * 16));
    HIP_TRY(hipMemcpy(da.p, a, 16 * 32 * sizeof(float), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(db.p, b, 32 * 16 * sizeof(float), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(mfma_probe_kernel, dim3(1), dim3(64), 0, nullptr, da.p, db.p, dd.p);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpy(d, dd.p, 16 * 16 * sizeof(float), hipMemcpyDeviceToHost));
    return DXG_OK;
}

}  // extern "C"
