// C-ABI of the transport engine: World / AAVoxelGrid / Transport /
// TransportProgress / DoseScore on one MI355X (include/dxmc_hip.h).
#include "dxmc_hip.h"

#include "device.hpp"
#include "physics.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <atomic>
#include <cmath>
#include <cstring>
#include <memory>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

namespace {

thread_local std::string g_error;

int fail(int code, const std::string& msg) { return dxg::setError(code, msg); }

const dxg::PhysicsList* physics(std::string& err) { return dxg::sharedPhysics(err); }

#define HIP_TRY(expr)                                                                     \
    do {                                                                                  \
        const hipError_t e_ = (expr);                                                     \
        if (e_ != hipSuccess)                                                             \
            return fail(DXG_ERR_DEVICE, std::string(#expr ": ") + hipGetErrorString(e_)); \
    } while (0)

template <typename T>
struct DevBuf {
    T* ptr = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release()
    {
        if (ptr)
            (void)hipFree(ptr);
        ptr = nullptr;
        n = 0;
    }
    hipError_t alloc(size_t count)
    {
        if (count == n && ptr)
            return hipSuccess;
        release();
        n = count;
        return hipMalloc(reinterpret_cast<void**>(&ptr), std::max<size_t>(count, 1) * sizeof(T));
    }
    hipError_t upload(const std::vector<T>& v, hipStream_t s)
    {
        hipError_t e = alloc(v.size());
        if (e != hipSuccess || v.empty())
            return e;
        return hipMemcpyAsync(ptr, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s);
    }
};

}  // namespace

struct dxg_progress {
    std::atomic<uint64_t> n { 0 };
    std::atomic<uint64_t> total { 0 };
    std::atomic<bool> stop { false };
    mutable std::mutex msg_mutex;
    std::string msg;
    void setMessage(const std::string& m)
    {
        std::lock_guard<std::mutex> lk(msg_mutex);
        msg = m;
    }
};

struct dxg_world {
    int level = 1;
    bool has_data = false, has_spacing = false, built = false;
    uint64_t dims[3] = { 0, 0, 0 };
    double spacing[3] = { 0, 0, 0 };
    std::vector<double> density;
    std::vector<uint8_t> material;
    std::vector<dxg::MixMaterial> mats;
    std::vector<double> rho_max;
    dxg::DeviceTables tables;

    int device = 0;
    hipStream_t stream = nullptr;
    DevBuf<uint32_t> d_voxels;
    // box of the voxels whose word differs from the grid's corner voxel (the
    // outer medium), voxel index bounds inclusive; outer_on false when it is
    // the whole grid
    bool outer_on = false;
    uint32_t outer_word = 0;
    uint64_t outer_box[6] = { 0, 0, 0, 0, 0, 0 };
    DevBuf<float> d_mu_tot;
    DevBuf<float4> d_mu_parts;
    DevBuf<float> d_majorant, d_rfwd, d_rinv;
    DevBuf<float2> d_kn, d_kns;
    DevBuf<float4> d_ia_shell;
    DevBuf<float> d_ia_inv, d_fl_cum, d_fl_line;
    DevBuf<double> d_tally;   // one record of dxg::kTallyRecord doubles per voxel, `replicas` copies
    int replicas = 1;
    size_t tallyDoubles() const { return dxg::kTallyRecord * static_cast<size_t>(dims[0] * dims[1] * dims[2]); }
    DevBuf<float> d_cells;   // macro-cell majorant ratios
    DevBuf<uint8_t> d_cells_q;   // the same, as bytes of cell_qscale (padded to 4)
    float cell_qscale = 0.0f;
    int cell = 0, cn[3] = { 0, 0, 0 };
    DevBuf<unsigned long long> d_pool;
    DevBuf<dxg::DevCounters> d_counters;
    // per-run beam tables
    DevBuf<dxg::DevExposure> d_ex;
    DevBuf<unsigned long long> d_ex_start;
    DevBuf<uint32_t> d_blob;

    uint64_t histories = 0;
    dxg_counters counters {};
    uint64_t generation = 0;   // world builds so far

    ~dxg_world()
    {
        if (stream) {
            (void)hipStreamSynchronize(stream);
            (void)hipStreamDestroy(stream);
        }
    }
    uint64_t size() const { return dims[0] * dims[1] * dims[2]; }
};

namespace dxg {
int setError(int code, const std::string& msg)
{
    g_error = msg;
    return code;
}

int worldView(const dxg_world* w, WorldView& out)
{
    if (!w || !w->built)
        return setError(DXG_ERR_STATE, "world not built");
    std::string err;
    const PhysicsList* pl = sharedPhysics(err);
    if (!pl)
        return setError(DXG_ERR_PHYSICS, err);
    const int32_t wz[2] = { 1, 8 };
    const double ww[2] = { 0.111894, 0.888106 };
    MixMaterial water;
    if (!makeMaterial(*pl, wz, ww, 2, water, err))
        return setError(DXG_ERR_PHYSICS, err);
    double pe, inc, coh;
    water.attenuation(*pl, 60.0, pe, inc, coh);
    const double mu_w = pe + inc + coh;
    out.embedding.clear();
    for (const auto& m : w->mats) {
        m.attenuation(*pl, 60.0, pe, inc, coh);
        out.embedding.push_back(static_cast<float>((pe + inc + coh) / mu_w));
    }
    out.voxels = w->d_voxels.ptr;
    out.tally = w->d_tally.ptr;
    out.nx = static_cast<int>(w->dims[0]);
    out.ny = static_cast<int>(w->dims[1]);
    out.nz = static_cast<int>(w->dims[2]);
    for (int a = 0; a < 3; ++a)
        out.spacing[a] = w->spacing[a];
    out.histories = w->histories;
    out.device = w->device;
    out.material_host = w->material.data();
    out.generation = w->generation;
    return DXG_OK;
}
}  // namespace dxg

namespace {

void fillParams(const dxg_world* w, dxg::DevParams& p)
{
    std::memset(&p, 0, sizeof p);
    p.voxels = w->d_voxels.ptr;
    p.nx = static_cast<int>(w->dims[0]);
    p.ny = static_cast<int>(w->dims[1]);
    p.nz = static_cast<int>(w->dims[2]);
    p.n_mat = static_cast<int>(w->mats.size());
    p.lut_n = dxg::kLutN;
    for (int a = 0; a < 3; ++a) {
        const float half = 0.5f * (static_cast<float>(w->dims[a]) * static_cast<float>(w->spacing[a]));
        p.bmin[a] = -half;
        p.bmax[a] = half;
        p.inv_spacing[a] = 1.0f / static_cast<float>(w->spacing[a]);
        p.vox_off[a] = -p.bmin[a] * p.inv_spacing[a];
    }
    p.level = w->level;
    p.outer_on = w->outer_on && w->cell == 0 ? 1 : 0;
    if (const char* env = std::getenv("DXG_OUTER"))   // 0: Woodcock steps through the outer medium too
        p.outer_on = std::atoi(env) != 0 ? p.outer_on : 0;
    p.outer_word = w->outer_word;
    for (int a = 0; a < 3; ++a) {
        // the box in world coordinates, widened by 1e-3 voxel so that rounding
        // in the voxel lookup (a few 1e-5 voxel on grids of up to ~1000 voxels
        // a side) never puts a point outside it into a voxel inside
        const double sp = w->spacing[a], b0 = -0.5 * static_cast<double>(w->dims[a]) * sp;
        if (w->outer_box[a] > w->outer_box[3 + a]) {   // empty box
            p.outer_lo[a] = 1.0f;
            p.outer_hi[a] = -1.0f;
        } else {
            p.outer_lo[a] = static_cast<float>(b0 + (static_cast<double>(w->outer_box[a]) - 1e-3) * sp);
            p.outer_hi[a] = static_cast<float>(b0 + (static_cast<double>(w->outer_box[3 + a]) + 1.0 + 1e-3) * sp);
        }
    }
    p.index24 = w->dims[0] < (1u << 24) && w->dims[1] * w->dims[2] < (1u << 24);
    p.mu_tot = w->d_mu_tot.ptr;
    p.mu_parts = w->d_mu_parts.ptr;
    p.majorant = w->d_majorant.ptr;
    p.ray_fwd = w->d_rfwd.ptr;
    p.ray_inv = w->d_rinv.ptr;
    p.kn_inv = w->d_kn.ptr;
    p.kns_inv = w->d_kns.ptr;
    p.ia_shell = w->d_ia_shell.ptr;
    p.ia_inv = w->d_ia_inv.ptr;
    p.fl_cum = w->d_fl_cum.ptr;
    p.fl_line = w->d_fl_line.ptr;
    p.tally = w->d_tally.ptr;
    p.tally_stride = w->tallyDoubles();
    p.tally_replicas = w->replicas;
    // same-record deposits of a wavefront merge in LDS before the atomics where
    // the grid is small (the tally copies are for the same contention): a
    // pencil beam into 64^3 water 8.58 -> 7.82 ms per 1e8 histories; large
    // grids gain nothing (C2 10.51 against 10.56 ms)
    p.tally_merge = w->replicas > 1 ? 1 : 0;
    if (const char* env = std::getenv("DXG_TALLY_MERGE"))
        p.tally_merge = std::atoi(env) != 0 ? 1 : 0;
    p.pool = w->d_pool.ptr;
    p.counters = w->d_counters.ptr;
    if (w->cell > 0) {
        p.cell_ratio = w->d_cells.ptr;
        p.cell_q = w->d_cells_q.ptr;
        p.cell_qscale = w->cell_qscale;
        p.cell_lds = 0;
        p.cnx = w->cn[0];
        p.cny = w->cn[1];
        p.cnz = w->cn[2];
        for (int a = 0; a < 3; ++a) {
            p.cell_len[a] = static_cast<float>(w->cell) * static_cast<float>(w->spacing[a]);
            p.cell_inv[a] = 1.0f / p.cell_len[a];
        }
    }
}

// Walker alias table (Vose's construction).
void aliasTable(const std::vector<double>& weight, std::vector<float>& prob, std::vector<int>& alias)
{
    const size_t n = weight.size();
    double tot = 0;
    for (double v : weight)
        tot += v;
    std::vector<double> q(n);
    for (size_t i = 0; i < n; ++i)
        q[i] = weight[i] / tot * static_cast<double>(n);
    std::vector<size_t> small, large;
    for (size_t i = 0; i < n; ++i)
        (q[i] < 1.0 ? small : large).push_back(i);
    prob.assign(n, 1.0f);
    alias.assign(n, 0);
    for (size_t i = 0; i < n; ++i)
        alias[i] = static_cast<int>(i);
    while (!small.empty() && !large.empty()) {
        const size_t s = small.back();
        small.pop_back();
        const size_t l = large.back();
        prob[s] = static_cast<float>(q[s]);
        alias[s] = static_cast<int>(l);
        q[l] = (q[l] + q[s]) - 1.0;
        if (q[l] < 1.0) {
            large.pop_back();
            small.push_back(l);
        }
    }
    for (size_t i : large)
        prob[i] = 1.0f;
    for (size_t i : small)
        prob[i] = 1.0f;
}

}  // namespace

extern "C" {

const char* dxg_last_error(void) { return g_error.c_str(); }

int dxg_world_create(dxg_world** out, int correction_level)
{
    if (!out)
        return fail(DXG_ERR_ARGUMENT, "dxg_world_create: null output pointer");
    *out = nullptr;
    if (correction_level < 0 || correction_level > 2)
        return fail(DXG_ERR_ARGUMENT, "correction level must be 0, 1 or 2");
    auto w = std::make_unique<dxg_world>();
    w->level = correction_level;
    *out = w.release();
    return DXG_OK;
}

void dxg_world_destroy(dxg_world* w) { delete w; }

int dxg_grid_set_data(dxg_world* w, const uint64_t dims[3], const double* density, const uint8_t* material,
                      const dxg_material* mats, uint64_t n_mats)
{
    if (!w || !dims || !density || !material || (!mats && n_mats))
        return fail(DXG_ERR_ARGUMENT, "dxg_grid_set_data: null argument");
    if (dims[0] == 0 || dims[1] == 0 || dims[2] == 0)
        return fail(DXG_ERR_ARGUMENT, "dxg_grid_set_data: zero dimension");
    const uint64_t n = dims[0] * dims[1] * dims[2];
    if (n >= (1ull << 31))
        return fail(DXG_ERR_ARGUMENT, "dxg_grid_set_data: grid has 2^31 or more voxels");
    if (n_mats == 0 || n_mats > 255)
        return fail(DXG_ERR_ARGUMENT, "dxg_grid_set_data: need 1..255 materials");
    std::string err;
    const dxg::PhysicsList* pl = physics(err);
    if (!pl)
        return fail(DXG_ERR_PHYSICS, err);
    std::vector<dxg::MixMaterial> built(n_mats);
    for (uint64_t i = 0; i < n_mats; ++i)
        if (!dxg::makeMaterial(*pl, mats[i].Z, mats[i].mass_fraction, mats[i].n, built[i], err))
            return fail(DXG_ERR_MATERIAL, "material " + std::to_string(i) + ": " + err);
    std::vector<double> rho_max(n_mats, 0.0);
    for (uint64_t i = 0; i < n; ++i) {
        const uint8_t m = material[i];
        if (m >= n_mats)
            return fail(DXG_ERR_DATA, "voxel " + std::to_string(i) + " has material index " + std::to_string(m) + " >= " + std::to_string(n_mats));
        const double d = density[i];
        if (!(d >= 0) || !std::isfinite(d))
            return fail(DXG_ERR_DATA, "voxel " + std::to_string(i) + " has a negative or non-finite density");
        rho_max[m] = std::max(rho_max[m], d);
    }
    for (int a = 0; a < 3; ++a)
        w->dims[a] = dims[a];
    w->density.assign(density, density + n);
    w->material.assign(material, material + n);
    w->mats = std::move(built);
    w->rho_max = std::move(rho_max);
    w->has_data = true;
    w->built = false;
    return DXG_OK;
}

int dxg_grid_set_spacing(dxg_world* w, const double spacing_cm[3])
{
    if (!w || !spacing_cm)
        return fail(DXG_ERR_ARGUMENT, "dxg_grid_set_spacing: null argument");
    for (int a = 0; a < 3; ++a)
        if (!(spacing_cm[a] > 0) || !std::isfinite(spacing_cm[a]))
            return fail(DXG_ERR_ARGUMENT, "dxg_grid_set_spacing: spacing must be positive");
    for (int a = 0; a < 3; ++a)
        w->spacing[a] = spacing_cm[a];
    w->has_spacing = true;
    w->built = false;
    return DXG_OK;
}

int dxg_world_build(dxg_world* w)
{
    if (!w)
        return fail(DXG_ERR_ARGUMENT, "dxg_world_build: null world");
    if (!w->has_data || !w->has_spacing)
        return fail(DXG_ERR_STATE, "dxg_world_build: setData and setSpacing must come first");
    std::string err;
    const dxg::PhysicsList* pl = physics(err);
    if (!pl)
        return fail(DXG_ERR_PHYSICS, err);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(DXG_ERR_DEVICE, "no HIP device is visible");
    HIP_TRY(hipGetDevice(&w->device));
    if (!w->stream)
        HIP_TRY(hipStreamCreateWithFlags(&w->stream, hipStreamNonBlocking));
    dxg::buildTables(*pl, w->mats, w->rho_max, w->level, w->tables);
    const uint64_t n = w->size();
    std::vector<uint32_t> packed(n);
    for (uint64_t i = 0; i < n; ++i)
        packed[i] = dxg::packVoxel(static_cast<float>(w->density[i]), w->material[i]);
    HIP_TRY(w->d_voxels.upload(packed, w->stream));
    {
        // The uniform outer medium around the object (usually air): photons
        // outside the box of the other voxels sample their flights exactly in
        // it instead of taking Woodcock steps through it.
        const uint64_t nx = w->dims[0], ny = w->dims[1], nz = w->dims[2];
        const uint32_t ow = packed[0];
        uint64_t lo[3] = { nx, ny, nz }, hi[3] = { 0, 0, 0 };
        bool any = false;
        for (uint64_t z = 0; z < nz; ++z)
            for (uint64_t y = 0; y < ny; ++y) {
                // a row: its first and last voxel that is not the outer medium
                const uint32_t* row = packed.data() + (z * ny + y) * nx;
                uint64_t first = 0;
                while (first < nx && row[first] == ow)
                    ++first;
                if (first == nx)
                    continue;
                uint64_t last = nx - 1;
                while (row[last] == ow)
                    --last;
                any = true;
                lo[0] = std::min(lo[0], first), hi[0] = std::max(hi[0], last);
                lo[1] = std::min(lo[1], y), hi[1] = std::max(hi[1], y);
                lo[2] = std::min(lo[2], z), hi[2] = std::max(hi[2], z);
            }
        if (!any)   // a uniform grid: an empty box (first index above the last)
            for (int a = 0; a < 3; ++a)
                lo[a] = 1, hi[a] = 0;
        w->outer_word = ow;
        w->outer_on = !(any && lo[0] == 0 && lo[1] == 0 && lo[2] == 0 && hi[0] == nx - 1 && hi[1] == ny - 1 && hi[2] == nz - 1);
        for (int a = 0; a < 3; ++a) {
            w->outer_box[a] = lo[a];
            w->outer_box[3 + a] = hi[a];
        }
    }
    HIP_TRY(w->d_mu_tot.upload(w->tables.mu_tot, w->stream));
    {
        std::vector<float4> parts(w->tables.mu_parts.size() / 4);
        std::memcpy(parts.data(), w->tables.mu_parts.data(), w->tables.mu_parts.size() * sizeof(float));
        HIP_TRY(w->d_mu_parts.upload(parts, w->stream));
    }
    HIP_TRY(w->d_majorant.upload(w->tables.majorant, w->stream));
    HIP_TRY(w->d_rfwd.upload(w->tables.ray_fwd, w->stream));
    HIP_TRY(w->d_rinv.upload(w->tables.ray_inv, w->stream));
    {
        static_assert(dxg::kKnRowsDev == dxg::kKnRows && dxg::kKnNDev == dxg::kKnN && dxg::kKnInvDEDev * dxg::kKnDE == 1.0
                          && dxg::kKnsRowsDev == dxg::kKnsRows && dxg::kKnsPerOctaveDev == dxg::kKnsPerOctave,
                      "Klein-Nishina table shapes");
        std::vector<float2> kn(w->tables.kn_inv.size() / 2);
        std::memcpy(kn.data(), w->tables.kn_inv.data(), w->tables.kn_inv.size() * sizeof(float));
        HIP_TRY(w->d_kn.upload(kn, w->stream));
        if (!w->tables.kns_inv.empty()) {
            std::vector<float2> kns(w->tables.kns_inv.size() / 2);
            std::memcpy(kns.data(), w->tables.kns_inv.data(), w->tables.kns_inv.size() * sizeof(float));
            HIP_TRY(w->d_kns.upload(kns, w->stream));
        }
    }
    if (w->level == 2) {
        std::vector<float4> shells(w->tables.ia_shell.size() / 4);
        std::memcpy(shells.data(), w->tables.ia_shell.data(), w->tables.ia_shell.size() * sizeof(float));
        HIP_TRY(w->d_ia_shell.upload(shells, w->stream));
        HIP_TRY(w->d_ia_inv.upload(w->tables.ia_inv, w->stream));
        HIP_TRY(w->d_fl_cum.upload(w->tables.fl_cum, w->stream));
        HIP_TRY(w->d_fl_line.upload(w->tables.fl_line, w->stream));
    }
    // Global majorant with exact flights through the outer medium, or
    // macro-cell majorants.  Cells pay off for grids whose voxel words no
    // longer sit in the 256 MB last-level cache (>= 32 M voxels, 128 MB) and
    // that the outer medium does not surround: on a 512x512x400 CT body
    // (the box of the non-air voxels holds 42 % of the grid) the global
    // majorant with exact air flights ran 5.6e9 histories/s against 4.7e9 for
    // 64^3 cells (16^3 3.7e9, 32^3 4.0e9, 128^3 4.3e9).  So grids of >= 32 M
    // voxels use 64^3 cells when the box holds more than 3/4 of the grid.
    // DXG_CELL = cell edge in voxels overrides (0 = global majorant).
    double box_frac = 1.0;
    if (w->outer_on) {
        box_frac = 0.0;
        if (w->outer_box[0] <= w->outer_box[3])
            box_frac = static_cast<double>(w->outer_box[3] - w->outer_box[0] + 1) * static_cast<double>(w->outer_box[4] - w->outer_box[1] + 1)
                       * static_cast<double>(w->outer_box[5] - w->outer_box[2] + 1) / static_cast<double>(n);
    }
    w->cell = n >= (32ull << 20) && box_frac > 0.75 ? 64 : 0;
    if (const char* env = std::getenv("DXG_CELL"))
        w->cell = std::clamp(std::atoi(env), 0, 1024);
    if (w->cell > 0) {
        const std::vector<float> ratio = dxg::cellRatios(w->tables, w->dims, w->density, w->material, w->cell, w->cn);
        if (w->cn[0] > 1023 || w->cn[1] > 1023 || w->cn[2] > 1023)
            w->cell = 0;   // cell coordinates are packed in 10 bits each
        else {
            HIP_TRY(w->d_cells.upload(ratio, w->stream));
            // bytes q with ratio <= q * qscale in FP32, qscale = max ratio / 255 (inflated)
            float top = 0.0f;
            for (float r : ratio)
                top = std::max(top, r);
            const float qs = top > 0.0f ? top / 255.0f * (1.0f + 1.0f / 1048576.0f) : 1.0f;
            std::vector<uint8_t> q((ratio.size() + 3) / 4 * 4, 0);
            for (size_t c = 0; c < ratio.size(); ++c) {
                int v = static_cast<int>(std::ceil(ratio[c] / qs));
                while (static_cast<float>(v) * qs < ratio[c])
                    ++v;
                q[c] = static_cast<uint8_t>(std::min(v, 255));
            }
            w->cell_qscale = qs;
            HIP_TRY(w->d_cells_q.upload(q, w->stream));
        }
    }
    // Copies of the tally for small grids: the FP64 atomics of one voxel record
    // are applied one after another where the memory holds it, so a narrow
    // beam on a small grid (a pencil beam's axis voxels take most first
    // interactions) waits on a few records; work-groups spread over the copies
    // and a fold after each run sums them.  Up to 16 copies within 128 MiB.
    {
        const size_t bytes = w->tallyDoubles() * sizeof(double);
        w->replicas = static_cast<int>(std::clamp<size_t>((size_t(128) << 20) / std::max<size_t>(bytes, 1), 1, 16));
        if (const char* env = std::getenv("DXG_TALLY_REPLICAS"))
            w->replicas = std::clamp(std::atoi(env), 1, 64);
    }
    HIP_TRY(w->d_tally.alloc(w->tallyDoubles() * static_cast<size_t>(w->replicas)));
    HIP_TRY(hipMemsetAsync(w->d_tally.ptr, 0, w->tallyDoubles() * static_cast<size_t>(w->replicas) * sizeof(double), w->stream));
    HIP_TRY(w->d_pool.alloc(1));
    HIP_TRY(w->d_counters.alloc(1));
    HIP_TRY(hipStreamSynchronize(w->stream));
    w->histories = 0;
    w->counters = dxg_counters {};
    w->built = true;
    static std::atomic<uint64_t> builds { 0 };
    w->generation = ++builds;
    return DXG_OK;
}

uint64_t dxg_grid_size(const dxg_world* w) { return w ? w->size() : 0; }

uint64_t dxg_world_histories(const dxg_world* w) { return w ? w->histories : 0; }

int dxg_world_clear_scores(dxg_world* w)
{
    if (!w || !w->built)
        return fail(DXG_ERR_STATE, "dxg_world_clear_scores: world not built");
    HIP_TRY(hipMemsetAsync(w->d_tally.ptr, 0, w->tallyDoubles() * static_cast<size_t>(w->replicas) * sizeof(double), w->stream));
    HIP_TRY(hipStreamSynchronize(w->stream));
    w->histories = 0;
    return DXG_OK;
}

}  // extern "C"

namespace dxg {

// The beam's device tables built on the host: exposure records, their first
// beam history, the spectrum/bowtie blob staged into LDS, the history count.
struct HostBeam {
    std::vector<dxg::DevExposure> ex;
    std::vector<unsigned long long> ex_start;
    std::vector<uint32_t> blob;
    dxg::BeamBlob bb {};
    uint64_t total = 0;
    double e_max = 0.0;   // highest spectrum energy (keV): no photon of the beam ever has more
};

// Validates a beam and builds its device tables on the host.
int buildHostBeam(const dxg_beam* beam, HostBeam& hb)
{
    if (beam->n_exposures == 0 || !beam->exposures)
        return fail(DXG_ERR_ARGUMENT, "dxg_transport_run: beam has no exposures");
    if (beam->n_spectra == 0 || !beam->spectra)
        return fail(DXG_ERR_ARGUMENT, "dxg_transport_run: beam has no spectrum");
    // ---- spectra -> alias tables
    std::vector<float> se, sp;
    std::vector<int> sa, soff, sn;
    for (uint64_t s = 0; s < beam->n_spectra; ++s) {
        const dxg_spectrum& spec = beam->spectra[s];
        if (spec.n == 0 || !spec.energy_kev || !spec.weight)
            return fail(DXG_ERR_ARGUMENT, "spectrum " + std::to_string(s) + " is empty");
        double tot = 0;
        std::vector<double> wv(spec.n);
        for (uint64_t k = 0; k < spec.n; ++k) {
            const double e = spec.energy_kev[k];
            if (!(e >= dxg::kCutoffKeV && e <= dxg::kLutEMax))
                return fail(DXG_ERR_ARGUMENT, "spectrum energy " + std::to_string(e) + " keV outside [1, 200] keV");
            if (!(spec.weight[k] >= 0) || !std::isfinite(spec.weight[k]))
                return fail(DXG_ERR_ARGUMENT, "spectrum weights must be finite and non-negative");
            wv[k] = spec.weight[k];
            tot += wv[k];
            hb.e_max = std::max(hb.e_max, e);
        }
        if (!(tot > 0))
            return fail(DXG_ERR_ARGUMENT, "spectrum weights sum to zero");
        // a spectrum equal to an earlier one (both tubes of a dual-source beam
        // at one voltage) shares its table: the tables sit in LDS, where every
        // kilobyte counts against work-groups per CU
        uint64_t same = s;
        for (uint64_t j = 0; j < s && same == s; ++j) {
            const dxg_spectrum& o = beam->spectra[j];
            if (o.n == spec.n && std::equal(spec.energy_kev, spec.energy_kev + spec.n, o.energy_kev)
                && std::equal(spec.weight, spec.weight + spec.n, o.weight))
                same = j;
        }
        if (same != s) {
            soff.push_back(soff[same]);
            sn.push_back(sn[same]);
            continue;
        }
        std::vector<float> prob;
        std::vector<int> alias;
        aliasTable(wv, prob, alias);
        soff.push_back(static_cast<int>(se.size()));
        sn.push_back(static_cast<int>(spec.n));
        for (uint64_t k = 0; k < spec.n; ++k) {
            se.push_back(static_cast<float>(spec.energy_kev[k]));
            sp.push_back(prob[k]);
            sa.push_back(alias[k]);
        }
    }
    // ---- bowties, sorted by |angle| and normalised to the central ray
    std::vector<float> ba, bw;
    std::vector<int> boff, bn;
    for (uint64_t b = 0; b < beam->n_bowties; ++b) {
        const dxg_bowtie& bt = beam->bowties[b];
        if (bt.n == 0 || !bt.angle_rad || !bt.weight)
            return fail(DXG_ERR_ARGUMENT, "bowtie " + std::to_string(b) + " is empty");
        std::vector<std::pair<double, double>> nodes;
        for (uint64_t k = 0; k < bt.n; ++k)
            nodes.emplace_back(std::fabs(bt.angle_rad[k]), bt.weight[k]);
        std::sort(nodes.begin(), nodes.end());
        double w0 = nodes.front().second;
        if (nodes.front().first > 0 && nodes.size() > 1) {
            // value at angle 0 is the first node (clamped)
            w0 = nodes.front().second;
        }
        if (!(w0 > 0))
            return fail(DXG_ERR_ARGUMENT, "bowtie central weight must be positive");
        boff.push_back(static_cast<int>(ba.size()));
        bn.push_back(static_cast<int>(nodes.size()));
        for (auto& nd : nodes) {
            ba.push_back(static_cast<float>(nd.first));
            bw.push_back(static_cast<float>(nd.second / w0));
        }
    }
    if (ba.empty()) {
        ba.push_back(0.f);
        bw.push_back(1.f);
    }
    // ---- exposures
    std::vector<dxg::DevExposure>& ex = hb.ex;
    std::vector<unsigned long long>& ex_start = hb.ex_start;
    ex.assign(beam->n_exposures, dxg::DevExposure {});
    ex_start.assign(beam->n_exposures, 0);
    uint64_t& total = hb.total;
    total = 0;
    for (uint64_t i = 0; i < beam->n_exposures; ++i) {
        const dxg_exposure& e = beam->exposures[i];
        if (e.spectrum_index >= beam->n_spectra)
            return fail(DXG_ERR_ARGUMENT, "exposure " + std::to_string(i) + " references a missing spectrum");
        if (e.bowtie_index >= 0 && static_cast<uint64_t>(e.bowtie_index) >= beam->n_bowties)
            return fail(DXG_ERR_ARGUMENT, "exposure " + std::to_string(i) + " references a missing bowtie");
        if (!(std::fabs(e.collimation_half_angles[0]) < 1.5707963) || !(std::fabs(e.collimation_half_angles[1]) < 1.5707963))
            return fail(DXG_ERR_ARGUMENT, "collimation half angles must be below pi/2");
        dxg::DevExposure& d = ex[i];
        std::memset(&d, 0, sizeof d);
        const double* c0 = e.cosines;
        const double* c1 = e.cosines + 3;
        const double ax[3] = { c0[1] * c1[2] - c0[2] * c1[1], c0[2] * c1[0] - c0[0] * c1[2], c0[0] * c1[1] - c0[1] * c1[0] };
        const double axn = std::sqrt(ax[0] * ax[0] + ax[1] * ax[1] + ax[2] * ax[2]);
        if (!(axn > 1e-12))
            return fail(DXG_ERR_ARGUMENT, "exposure direction cosines are parallel");
        for (int a = 0; a < 3; ++a) {
            d.pos[a] = static_cast<float>(e.position[a]);
            d.c0[a] = static_cast<float>(c0[a]);
            d.c1[a] = static_cast<float>(c1[a]);
            d.axis[a] = static_cast<float>(ax[a] / axn);
        }
        d.half_angle[0] = static_cast<float>(e.collimation_half_angles[0]);
        d.half_angle[1] = static_cast<float>(e.collimation_half_angles[1]);
        d.weight = static_cast<float>(e.weight);
        d.spectrum = static_cast<int32_t>(e.spectrum_index);
        d.bowtie = e.bowtie_index;
        ex_start[i] = total;
        total += e.n_particles;
    }
    if (total == 0)
        return fail(DXG_ERR_ARGUMENT, "beam has zero particles");
    // ---- the small per-beam tables travel as one blob staged into LDS
    dxg::BeamBlob& bb = hb.bb;
    bb.n_spec = static_cast<int>(soff.size());
    bb.n_bow = static_cast<int>(boff.size());
    bb.n_bins = static_cast<int>(se.size());
    bb.n_nodes = static_cast<int>(ba.size());
    bb.o_spec_off = 0;
    bb.o_spec_n = bb.n_spec;
    bb.o_bow_off = 2 * bb.n_spec;
    bb.o_bow_n = 2 * bb.n_spec + bb.n_bow;
    bb.o_energy = 2 * (bb.n_spec + bb.n_bow);
    bb.o_prob = bb.o_energy + bb.n_bins;
    bb.o_alias = bb.o_prob + bb.n_bins;
    bb.o_angle = bb.o_alias + bb.n_bins;
    bb.o_weight = bb.o_angle + bb.n_nodes;
    bb.words = bb.o_weight + bb.n_nodes;
    std::vector<uint32_t>& blob = hb.blob;
    blob.assign(static_cast<size_t>(bb.words), 0u);
    auto fbits = [](float f) {
        uint32_t u;
        std::memcpy(&u, &f, 4);
        return u;
    };
    for (int i = 0; i < bb.n_spec; ++i) {
        blob[static_cast<size_t>(bb.o_spec_off + i)] = static_cast<uint32_t>(soff[static_cast<size_t>(i)]);
        blob[static_cast<size_t>(bb.o_spec_n + i)] = static_cast<uint32_t>(sn[static_cast<size_t>(i)]);
    }
    for (int i = 0; i < bb.n_bow; ++i) {
        blob[static_cast<size_t>(bb.o_bow_off + i)] = static_cast<uint32_t>(boff[static_cast<size_t>(i)]);
        blob[static_cast<size_t>(bb.o_bow_n + i)] = static_cast<uint32_t>(bn[static_cast<size_t>(i)]);
    }
    for (int i = 0; i < bb.n_bins; ++i) {
        blob[static_cast<size_t>(bb.o_energy + i)] = fbits(se[static_cast<size_t>(i)]);
        blob[static_cast<size_t>(bb.o_prob + i)] = fbits(sp[static_cast<size_t>(i)]);
        blob[static_cast<size_t>(bb.o_alias + i)] = static_cast<uint32_t>(sa[static_cast<size_t>(i)]);
    }
    for (int i = 0; i < bb.n_nodes; ++i) {
        blob[static_cast<size_t>(bb.o_angle + i)] = fbits(ba[static_cast<size_t>(i)]);
        blob[static_cast<size_t>(bb.o_weight + i)] = fbits(bw[static_cast<size_t>(i)]);
    }
    return DXG_OK;
}

// Fills the launch parameters of a run of `beam` on world w (everything but
// the history range and the device addresses of the beam tables, the pool and
// the counters) and plans the LDS use.
int runParams(dxg_world* w, const dxg_beam* beam, const HostBeam& hb, uint64_t seed, DevParams& p)
{
    const dxg::BeamBlob& bb = hb.bb;
    fillParams(w, p);
    p.n_ex = static_cast<int>(beam->n_exposures);
    p.ppe = beam->exposures[0].n_particles;
    for (uint64_t i = 1; i < beam->n_exposures; ++i)
        if (beam->exposures[i].n_particles != p.ppe)
            p.ppe = 0;
    p.inv_ppe = p.ppe ? 1.0 / static_cast<double>(p.ppe) : 0.0;
    p.bb = bb;
    // The LDS tables stop at the beam's highest energy: energies only fall
    // from there (Compton, fluorescence) or stay (Rayleigh), and a photon at
    // E reads bin j = (int)((E - E0) * (1 / dE)) (FP32, as the kernel computes
    // it) and the next, as with the full table; one bin of margin.
    {
        const float f = (static_cast<float>(hb.e_max) - static_cast<float>(dxg::kLutE0)) * static_cast<float>(1.0 / dxg::kLutDE);
        p.lut_n = std::min(dxg::kLutN, static_cast<int>(f) + 3);
    }
    p.seed = seed;
    p.rng_offset = w->histories;
    p.dose_scale = 1.0f;
    p.refill_threshold = 16;
    if (const char* env = std::getenv("DXG_REFILL"))
        p.refill_threshold = std::clamp(std::atoi(env), 1, 64);
    p.interact_threshold = 24;
    if (const char* env = std::getenv("DXG_INTERACT"))
        p.interact_threshold = std::clamp(std::atoi(env), 1, 64);
    p.outer_threshold = 16;
    if (const char* env = std::getenv("DXG_OUTERT"))
        p.outer_threshold = std::clamp(std::atoi(env), 1, 64);
    p.batch = 256;
    if (const char* env = std::getenv("DXG_BATCH"))
        p.batch = std::max<unsigned long long>(64, std::strtoull(env, nullptr, 10));

    // ---- LDS plan: the attenuation table joins the LDS tables while a block
    // still fits 3 to a CU (<= 48 KiB); larger material sets read it from L1/L2
    constexpr size_t kLdsLutLimit = 48 * 1024;
    p.lds_lut = 1;
    if (dxg::transportLdsBytes(p) > kLdsLutLimit)
        p.lds_lut = 0;
    if (dxg::transportLdsBytes(p) > 64 * 1024)
        return fail(DXG_ERR_ARGUMENT, "beam tables (spectra + bowties) exceed the LDS budget");
    // macro-cell bytes join LDS when that costs no work-group per CU
    if (p.cell_ratio) {
        HIP_TRY(hipSetDevice(w->device));
        const size_t without = dxg::transportLdsBytes(p);
        p.cell_lds = 1;
        const size_t with = dxg::transportLdsBytes(p);
        if (with > 64 * 1024
            || dxg::transportBlocksPerCU(p.lds_lut != 0, p.level, true, with) < dxg::transportBlocksPerCU(p.lds_lut != 0, p.level, true, without))
            p.cell_lds = 0;
        if (const char* env = std::getenv("DXG_CELL_LDS"))
            p.cell_lds = std::atoi(env) != 0 && with <= 64 * 1024 ? 1 : 0;
    }
    return DXG_OK;
}

// Validates a beam, uploads its exposure and spectrum/bowtie tables to the
// world's device buffers and fills the launch parameters (everything but the
// history range).  total receives the beam's history count.
int prepareRun(dxg_world* w, const dxg_beam* beam, uint64_t seed, DevParams& p, uint64_t& total)
{
    if (!w || !beam)
        return fail(DXG_ERR_ARGUMENT, "dxg_transport_run: null argument");
    if (!w->built)
        return fail(DXG_ERR_STATE, "dxg_transport_run: World::build() has not run");
    HostBeam hb;
    if (int rc = buildHostBeam(beam, hb); rc != DXG_OK)
        return rc;
    HIP_TRY(hipSetDevice(w->device));
    HIP_TRY(w->d_ex.upload(hb.ex, w->stream));
    HIP_TRY(w->d_ex_start.upload(hb.ex_start, w->stream));
    HIP_TRY(w->d_blob.upload(hb.blob, w->stream));
    if (int rc = runParams(w, beam, hb, seed, p); rc != DXG_OK)
        return rc;
    p.exposures = w->d_ex.ptr;
    p.ex_start = w->d_ex_start.ptr;
    p.blob = w->d_blob.ptr;
    total = hb.total;
    return DXG_OK;
}

void addHistories(dxg_world* w, uint64_t n, const dxg_counters& c)
{
    w->histories += n;
    w->counters = c;
}

int worldDevice(const dxg_world* w) { return w->device; }

int foldTally(dxg_world* w)
{
    HIP_TRY(hipSetDevice(w->device));
    dxg::launchTallyFold(w->d_tally.ptr, w->tallyDoubles(), w->replicas, w->stream);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(w->stream));
    return DXG_OK;
}

// Launch settings chosen by the run's own timing, kept per world build and
// beam layout: dead lanes that start a refill, parked lanes that start an
// interaction phase, lanes outside the box that start an outer-flight phase,
// and histories a wavefront claims per pool atomic.  Events, tallies and
// scores do not depend on them (per-history random streams), only time does.
struct Tuned {
    int refill, interact, outer;
    unsigned long long batch;
};
constexpr Tuned kTuneCandidates[] = {
    { 16, 24, 16, 256 }, { 12, 20, 12, 256 }, { 16, 32, 24, 256 }, { 20, 24, 16, 512 }, { 12, 24, 8, 128 },
};
constexpr int kTuneRounds = 2;   // every candidate timed twice, interleaved; the faster launch counts

struct TuneKey {
    const dxg_world* w;
    uint64_t generation;
    int level, n_ex, lds_lut;
    bool cells, uniform_ppe;
    bool operator<(const TuneKey& o) const
    {
        return std::tie(w, generation, level, n_ex, lds_lut, cells, uniform_ppe)
               < std::tie(o.w, o.generation, o.level, o.n_ex, o.lds_lut, o.cells, o.uniform_ppe);
    }
};
std::mutex g_tune_mutex;
std::map<TuneKey, Tuned> g_tuned;

bool autotuneEnabled()
{
    if (const char* env = std::getenv("DXG_AUTOTUNE"); env && std::atoi(env) == 0)
        return false;
    for (const char* name : { "DXG_REFILL", "DXG_INTERACT", "DXG_OUTERT", "DXG_BATCH" })
        if (std::getenv(name))
            return false;
    return true;
}

void applyTuned(DevParams& p, const Tuned& t)
{
    p.refill_threshold = t.refill;
    p.interact_threshold = t.interact;
    p.outer_threshold = t.outer;
    p.batch = t.batch;
}

// The launches of one run: the persistent transport kernel over the history
// range in equal chunks (progress and cancel act between them), event
// counters summed into the world.  With `events` the kernel time comes from
// HIP events around each launch, else from the host clock.  With `events` and
// autotuning on, the first run of a beam layout on a world build with enough
// histories starts with short launches of each candidate setting
// (kTuneCandidates, twice each, interleaved); the fastest per history is kept
// for the rest of the run and for later runs of that layout.
int runChunks(dxg_world* w, DevParams& p, uint64_t total, dxg_progress* progress, bool events)
{
    const size_t lds_bytes = dxg::transportLdsBytes(p);

    // ---- launch geometry: fill every CU to the occupancy limit
    int cus = 0;
    HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, w->device));
    const int block = dxg::transportBlockSize();
    const int per_cu = dxg::transportBlocksPerCU(p.lds_lut != 0, p.level, p.cell_ratio != nullptr, lds_bytes);
    int blocks = cus * per_cu;
    if (const char* env = std::getenv("DXG_BLOCKS"))
        blocks = std::max(1, std::atoi(env));
    const uint64_t lanes = static_cast<uint64_t>(blocks) * block;

    if (progress) {
        progress->stop.store(false);
        progress->n.store(0);
        progress->total.store(total);
        progress->setMessage("Simulating " + std::to_string(total) + " histories");
    }
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    if (events) {
        HIP_TRY(hipEventCreate(&ev0));
        HIP_TRY(hipEventCreate(&ev1));
    }
    dxg_counters acc {};
    int status = DXG_OK;
    // one launch over beam histories [h_begin, h_end); ms receives its time
    auto launchRange = [&](uint64_t h_begin, uint64_t h_end, float& ms) -> int {
        p.h_begin = h_begin;
        p.h_end = h_end;
        dxg::launchReset(p.pool, p.counters, w->stream);
        const auto t0 = std::chrono::steady_clock::now();
        if (events)
            HIP_TRY(hipEventRecord(ev0, w->stream));
        dxg::launchTransport(p, blocks, block, w->stream);
        HIP_TRY(hipGetLastError());
        if (events)
            HIP_TRY(hipEventRecord(ev1, w->stream));
        dxg::DevCounters c {};
        HIP_TRY(hipMemcpyAsync(&c, p.counters, sizeof c, hipMemcpyDeviceToHost, w->stream));
        HIP_TRY(hipStreamSynchronize(w->stream));
        ms = 0;
        if (events)
            HIP_TRY(hipEventElapsedTime(&ms, ev0, ev1));
        else   // no events without allocation: the host's clock around launch and wait
            ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
        acc.kernel_ms += ms;
        acc.kernel_launches += 1;
        acc.histories += c.histories;
        acc.steps += c.steps;
        acc.real_events += c.real_events;
        acc.deposits += c.deposits;
        acc.emitted_energy += c.emitted;
        acc.escaped_energy += c.escaped;
        acc.deposited_energy += c.deposited;
        acc.uncollided += c.uncollided;
        acc.cell_crossings += c.crossings;
        for (int k = 0; k < 4; ++k) {
            acc.phase_cycles[k] += c.phase_cycles[k];
            acc.phase_lanes[k] += c.phase_lanes[k];
            acc.phase_iters[k] += c.phase_iters[k];
        }
        w->histories += h_end - h_begin;
        if (progress)
            progress->n.store(h_end);
        return DXG_OK;
    };

    // ---- autotuned launch settings
    uint64_t h0 = 0;
    int autotuned = 0;
    if (events && autotuneEnabled()) {
        const TuneKey key { w, w->generation, p.level, p.n_ex, p.lds_lut, p.cell_ratio != nullptr, p.ppe != 0 };
        bool found = false;
        {
            std::lock_guard<std::mutex> lock(g_tune_mutex);
            if (auto it = g_tuned.find(key); it != g_tuned.end()) {
                applyTuned(p, it->second);
                found = true;
            }
        }
        autotuned = found ? 1 : 0;
        constexpr int n_cand = static_cast<int>(sizeof kTuneCandidates / sizeof kTuneCandidates[0]);
        const uint64_t sub = std::max<uint64_t>(lanes * 16, 1ull << 20);
        if (!found && total >= sub * (kTuneRounds * n_cand + 2)) {
            float best[n_cand];
            for (int c = 0; c < n_cand; ++c)
                best[c] = 3.0e38f;
            for (int r = 0; r < kTuneRounds && status == DXG_OK; ++r)
                for (int c = 0; c < n_cand; ++c) {
                    if (progress && progress->stop.load()) {
                        status = fail(DXG_ERR_CANCELLED, "transport cancelled by the progress stop flag");
                        break;
                    }
                    applyTuned(p, kTuneCandidates[c]);
                    float ms = 0;
                    if (int rc = launchRange(h0, h0 + sub, ms); rc != DXG_OK)
                        return rc;
                    best[c] = std::min(best[c], ms);
                    h0 += sub;
                }
            int pick = 0;
            for (int c = 1; c < n_cand; ++c)
                if (best[c] < best[pick])
                    pick = c;
            applyTuned(p, kTuneCandidates[pick]);
            if (status == DXG_OK) {
                std::lock_guard<std::mutex> lock(g_tune_mutex);
                g_tuned[key] = kTuneCandidates[pick];
                autotuned = 1;
            }
        }
    }

    // ---- the rest of the run in equal launches of up to 256 histories per
    // lane: the drain at the end of a launch is paid rarely, and progress /
    // cancel still act every few tens of ms
    const uint64_t rest = total - h0;
    uint64_t chunk = std::max<uint64_t>(lanes * 256, 1ull << 20);
    if (const char* env = std::getenv("DXG_CHUNK"))
        chunk = std::max<uint64_t>(1, std::strtoull(env, nullptr, 10));
    if (rest > chunk) {   // equal launches: no short last launch that is mostly drain
        const uint64_t n_launch = (rest + chunk - 1) / chunk;
        chunk = (rest + n_launch - 1) / n_launch;
    }
    for (; status == DXG_OK && h0 < total; h0 += chunk) {
        if (progress && progress->stop.load()) {
            status = fail(DXG_ERR_CANCELLED, "transport cancelled by the progress stop flag");
            break;
        }
        float ms = 0;
        if (int rc = launchRange(h0, std::min(total, h0 + chunk), ms); rc != DXG_OK)
            return rc;
    }
    if (events) {
        (void)hipEventDestroy(ev0);
        (void)hipEventDestroy(ev1);
    }
    dxg::launchTallyFold(p.tally, p.tally_stride, p.tally_replicas, w->stream);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(w->stream));
    acc.refill_threshold = p.refill_threshold;
    acc.interact_threshold = p.interact_threshold;
    acc.outer_threshold = p.outer_threshold;
    acc.batch = p.batch;
    acc.autotuned = autotuned;
    w->counters = acc;
    if (progress && status == DXG_OK)
        progress->setMessage("Simulation finished");
    return status;
}

}  // namespace dxg

extern "C" {

int dxg_transport_run(dxg_world* w, const dxg_beam* beam, dxg_progress* progress, int apply_calibration, uint64_t seed)
{
    dxg::DevParams p;
    uint64_t total = 0;
    if (int rc = dxg::prepareRun(w, beam, seed, p, total); rc != DXG_OK)
        return rc;
    if (apply_calibration && beam->calibration > 0.0)
        p.dose_scale = static_cast<float>(beam->calibration);
    return dxg::runChunks(w, p, total, progress, true);
}

int dxg_transport_run_static(dxg_world* w, const dxg_beam* beam, dxg_progress* progress, int apply_calibration, uint64_t seed)
{
    if (!w || !beam)
        return fail(DXG_ERR_ARGUMENT, "dxg_transport_run_static: null argument");
    if (!w->built)
        return fail(DXG_ERR_STATE, "dxg_transport_run_static: World::build() has not run");
    dxg::HostBeam hb;
    if (int rc = dxg::buildHostBeam(beam, hb); rc != DXG_OK)
        return rc;
    // scratch layout: [0, 64) pool counter, [64, 256) event counters, then the
    // beam blob, the exposures' first histories and the exposure records, each
    // at a 256-byte boundary
    auto up256 = [](size_t b) { return (b + 255) / 256 * 256; };
    static_assert(sizeof(dxg::DevCounters) <= 192, "counters fit their scratch slot");
    const size_t off_blob = 256;
    const size_t off_start = off_blob + up256(hb.blob.size() * sizeof(uint32_t));
    const size_t off_ex = off_start + up256(hb.ex_start.size() * sizeof(unsigned long long));
    const size_t need = off_ex + hb.ex.size() * sizeof(dxg::DevExposure);
    if (need > dxg::staticScratchBytes())
        return fail(DXG_ERR_ARGUMENT, "dxg_transport_run_static: the beam's tables need " + std::to_string(need)
                                          + " bytes of scratch, more than its " + std::to_string(dxg::staticScratchBytes())
                                          + " (about 12,000 exposures with small spectra); use dxg_transport_run");
    HIP_TRY(hipSetDevice(w->device));
    void* base = nullptr;
    HIP_TRY(static_cast<hipError_t>(dxg::staticScratch(w->device, &base)));
    char* sc = static_cast<char*>(base);
    HIP_TRY(hipMemcpyAsync(sc + off_blob, hb.blob.data(), hb.blob.size() * sizeof(uint32_t), hipMemcpyHostToDevice, w->stream));
    HIP_TRY(hipMemcpyAsync(sc + off_start, hb.ex_start.data(), hb.ex_start.size() * sizeof(unsigned long long), hipMemcpyHostToDevice,
                           w->stream));
    HIP_TRY(hipMemcpyAsync(sc + off_ex, hb.ex.data(), hb.ex.size() * sizeof(dxg::DevExposure), hipMemcpyHostToDevice, w->stream));
    dxg::DevParams p;
    if (int rc = dxg::runParams(w, beam, hb, seed, p); rc != DXG_OK)
        return rc;
    p.pool = reinterpret_cast<unsigned long long*>(sc);
    p.counters = reinterpret_cast<dxg::DevCounters*>(sc + 64);
    p.blob = reinterpret_cast<const uint32_t*>(sc + off_blob);
    p.ex_start = reinterpret_cast<const unsigned long long*>(sc + off_start);
    p.exposures = reinterpret_cast<const dxg::DevExposure*>(sc + off_ex);
    if (apply_calibration && beam->calibration > 0.0)
        p.dose_scale = static_cast<float>(beam->calibration);
    return dxg::runChunks(w, p, hb.total, progress, false);
}

int dxg_read_energy_scores(const dxg_world* w, double* sum_e, double* sum_e2, uint64_t* n_events)
{
    if (!w || !w->built)
        return fail(DXG_ERR_STATE, "dxg_read_energy_scores: world not built");
    const uint64_t n = w->size();
    constexpr int R = dxg::kTallyRecord;
    std::vector<double> t(R * n);
    HIP_TRY(hipMemcpyAsync(t.data(), w->d_tally.ptr, t.size() * sizeof(double), hipMemcpyDeviceToHost, w->stream));
    HIP_TRY(hipStreamSynchronize(w->stream));
    for (uint64_t i = 0; i < n; ++i) {
        if (sum_e)
            sum_e[i] = t[R * i];
        if (sum_e2)
            sum_e2[i] = t[R * i + 1];
        if (n_events)   // an FP64 integer (exact below 2^53)
            n_events[i] = static_cast<uint64_t>(t[R * i + 2]);
    }
    return DXG_OK;
}

int dxg_read_dose_scores(const dxg_world* w, double* dose_mgy, double* variance_mgy2, uint64_t* n_events)
{
    if (!w || !w->built)
        return fail(DXG_ERR_STATE, "dxg_read_dose_scores: world not built");
    const uint64_t n = w->size();
    std::vector<double> e(n), e2(n);
    int rc = dxg_read_energy_scores(w, e.data(), e2.data(), n_events);
    if (rc != DXG_OK)
        return rc;
    const double voxel_cm3 = w->spacing[0] * w->spacing[1] * w->spacing[2];
    const double H = static_cast<double>(w->histories);
    for (uint64_t i = 0; i < n; ++i) {
        const double mass_kg = w->density[i] * voxel_cm3 * 1e-3;
        const double f = mass_kg > 0 ? dxg::kKeVToJ / mass_kg * 1e3 : 0.0;   // keV -> mGy
        if (dose_mgy)
            dose_mgy[i] = e[i] * f;
        if (variance_mgy2) {
            double v = 0;
            if (H > 1)
                v = std::max(0.0, H / (H - 1.0) * (e2[i] - e[i] * e[i] / H));
            variance_mgy2[i] = v * f * f;
        }
    }
    return DXG_OK;
}

int dxg_get_counters(const dxg_world* w, dxg_counters* out)
{
    if (!w || !out)
        return fail(DXG_ERR_ARGUMENT, "dxg_get_counters: null argument");
    *out = w->counters;
    return DXG_OK;
}

int dxg_progress_create(dxg_progress** out)
{
    if (!out)
        return fail(DXG_ERR_ARGUMENT, "dxg_progress_create: null output pointer");
    *out = new dxg_progress();
    return DXG_OK;
}

void dxg_progress_destroy(dxg_progress* p) { delete p; }

void dxg_progress_request_stop(dxg_progress* p)
{
    if (p)
        p->stop.store(true);
}

int dxg_progress_continue(const dxg_progress* p) { return p && !p->stop.load() ? 1 : 0; }

void dxg_progress_get(const dxg_progress* p, uint64_t* n, uint64_t* total)
{
    if (n)
        *n = p ? p->n.load() : 0;
    if (total)
        *total = p ? p->total.load() : 0;
}

void dxg_progress_message(const dxg_progress* p, char* buf, uint64_t cap)
{
    if (!buf || cap == 0)
        return;
    std::string m;
    if (p) {
        std::lock_guard<std::mutex> lk(p->msg_mutex);
        m = p->msg;
    }
    const size_t k = std::min<size_t>(m.size(), cap - 1);
    std::memcpy(buf, m.data(), k);
    buf[k] = '\0';
}

int dxg_debug_compton_cos(double energy_kev, uint64_t n, const float* u, float* cos_out)
{
    if ((!u && n) || (!cos_out && n))
        return fail(DXG_ERR_ARGUMENT, "dxg_debug_compton_cos: null argument");
    if (!(energy_kev >= dxg::kKnE0 && energy_kev <= dxg::kLutEMax))
        return fail(DXG_ERR_ARGUMENT, "dxg_debug_compton_cos: energy outside 1-200 keV");
    const auto& t = dxg::comptonTable();
    const float2* tab = reinterpret_cast<const float2*>(t.data());
    const float e = static_cast<float>(energy_kev);
    for (uint64_t i = 0; i < n; ++i) {
        if (!(u[i] >= 0.0f && u[i] < 1.0f))
            return fail(DXG_ERR_ARGUMENT, "dxg_debug_compton_cos: u must lie in [0, 1)");
        cos_out[i] = dxg::knCos(tab, e, u[i]);
    }
    return DXG_OK;
}

int dxg_debug_compton_cos_material(const dxg_material* m, double energy_kev, uint64_t n, const float* u, float* cos_out)
{
    if (!m || (!u && n) || (!cos_out && n))
        return fail(DXG_ERR_ARGUMENT, "dxg_debug_compton_cos_material: null argument");
    if (!(energy_kev >= dxg::kKnE0 && energy_kev <= dxg::kLutEMax))
        return fail(DXG_ERR_ARGUMENT, "dxg_debug_compton_cos_material: energy outside 1-200 keV");
    std::string err;
    const dxg::PhysicsList* pl = physics(err);
    if (!pl)
        return fail(DXG_ERR_PHYSICS, err);
    dxg::MixMaterial mm;
    if (!dxg::makeMaterial(*pl, m->Z, m->mass_fraction, m->n, mm, err))
        return fail(DXG_ERR_MATERIAL, err);
    std::vector<float> t;
    dxg::buildComptonSTable(*pl, mm, t);
    const float2* tab = reinterpret_cast<const float2*>(t.data());
    const float e = static_cast<float>(energy_kev);
    for (uint64_t i = 0; i < n; ++i) {
        if (!(u[i] >= 0.0f && u[i] < 1.0f))
            return fail(DXG_ERR_ARGUMENT, "dxg_debug_compton_cos_material: u must lie in [0, 1)");
        cos_out[i] = dxg::knsLookup(tab, std::log2(e) * dxg::kKnsPerOctaveDev,
                                    (1.0f - std::sqrt(1.0f - u[i])) * static_cast<float>(dxg::kKnNDev - 1));
    }
    return DXG_OK;
}

int dxg_material_attenuation(const dxg_material* m, const double* energy_kev, uint64_t n, double* out)
{
    if (!m || (!energy_kev && n) || (!out && n))
        return fail(DXG_ERR_ARGUMENT, "dxg_material_attenuation: null argument");
    std::string err;
    const dxg::PhysicsList* pl = physics(err);
    if (!pl)
        return fail(DXG_ERR_PHYSICS, err);
    dxg::MixMaterial mm;
    if (!dxg::makeMaterial(*pl, m->Z, m->mass_fraction, m->n, mm, err))
        return fail(DXG_ERR_MATERIAL, err);
    for (uint64_t i = 0; i < n; ++i) {
        if (!(energy_kev[i] > 0))
            return fail(DXG_ERR_ARGUMENT, "energy must be positive");
        double pe, inc, coh;
        mm.attenuation(*pl, energy_kev[i], pe, inc, coh);
        out[4 * i] = pe;
        out[4 * i + 1] = inc;
        out[4 * i + 2] = coh;
        out[4 * i + 3] = pe + inc + coh;
    }
    return DXG_OK;
}

int dxg_debug_traverse(const dxg_world* w, uint64_t n_rays, const float* pos, const float* dir, uint64_t n_steps,
                       const float* steps, int64_t* out_index, int32_t* out_material)
{
    if (!w || !w->built)
        return fail(DXG_ERR_STATE, "dxg_debug_traverse: world not built");
    if (n_rays == 0 || n_steps == 0)
        return DXG_OK;
    if (!pos || !dir || !steps || !out_index || !out_material)
        return fail(DXG_ERR_ARGUMENT, "dxg_debug_traverse: null argument");
    HIP_TRY(hipSetDevice(w->device));
    DevBuf<float> dp, dd, ds;
    DevBuf<int64_t> di;
    DevBuf<int32_t> dm;
    HIP_TRY(dp.upload(std::vector<float>(pos, pos + 3 * n_rays), w->stream));
    HIP_TRY(dd.upload(std::vector<float>(dir, dir + 3 * n_rays), w->stream));
    HIP_TRY(ds.upload(std::vector<float>(steps, steps + n_rays * n_steps), w->stream));
    HIP_TRY(di.alloc(n_rays * n_steps));
    HIP_TRY(dm.alloc(n_rays * n_steps));
    dxg::DevParams p;
    fillParams(w, p);
    dxg::launchTraverse(p, n_rays, dp.ptr, dd.ptr, n_steps, ds.ptr, di.ptr, dm.ptr, w->stream);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(out_index, di.ptr, n_rays * n_steps * sizeof(int64_t), hipMemcpyDeviceToHost, w->stream));
    HIP_TRY(hipMemcpyAsync(out_material, dm.ptr, n_rays * n_steps * sizeof(int32_t), hipMemcpyDeviceToHost, w->stream));
    HIP_TRY(hipStreamSynchronize(w->stream));
    return DXG_OK;
}

int dxg_debug_source(dxg_world* w, const dxg_beam* beam, uint64_t seed, uint64_t n, float* out)
{
    if (!out && n)
        return fail(DXG_ERR_ARGUMENT, "dxg_debug_source: null argument");
    dxg::DevParams p;
    uint64_t total = 0;
    if (int rc = dxg::prepareRun(w, beam, seed, p, total); rc != DXG_OK)
        return rc;
    if (n > total)
        return fail(DXG_ERR_ARGUMENT, "dxg_debug_source: n exceeds the beam's " + std::to_string(total) + " histories");
    if (n == 0)
        return DXG_OK;
    p.rng_offset = 0;
    DevBuf<float> d;
    HIP_TRY(d.alloc(8 * n));
    dxg::launchSource(p, n, d.ptr, w->stream);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(out, d.ptr, 8 * n * sizeof(float), hipMemcpyDeviceToHost, w->stream));
    HIP_TRY(hipStreamSynchronize(w->stream));
    return DXG_OK;
}

int dxg_debug_device_copy(uint64_t bytes, int iters, double* ms)
{
    if (bytes == 0 || bytes % 16 != 0 || iters < 1)
        return fail(DXG_ERR_ARGUMENT, "dxg_debug_device_copy: bytes must be a positive multiple of 16, iters >= 1");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(DXG_ERR_DEVICE, "no HIP device is visible");
    DevBuf<uint8_t> a, b;
    HIP_TRY(a.alloc(bytes));
    HIP_TRY(b.alloc(bytes));
    HIP_TRY(hipMemset(a.ptr, 1, bytes));
    HIP_TRY(hipDeviceSynchronize());
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipEventRecord(e0, nullptr));
    for (int i = 0; i < iters; ++i)
        dxg::launchCopy(a.ptr, b.ptr, bytes, nullptr);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(e1, nullptr));
    HIP_TRY(hipEventSynchronize(e1));
    float t = 0;
    HIP_TRY(hipEventElapsedTime(&t, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (ms)
        *ms = t / iters;
    return DXG_OK;
}

}  // extern "C"
