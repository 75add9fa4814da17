// Batched transport: many (world, beam) instances of uneven size in one
// persistent launch, captured once into a HIP graph and replayed
// (include/dxmc_hip.h, dxg_batch_*).
#include "dxmc_hip.h"

#include "device.hpp"
#include "physics.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <numeric>
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

// A deep copy of one beam (the C-ABI copies its inputs).
struct BeamCopy {
    std::vector<dxg_exposure> exposures;
    std::vector<std::vector<double>> spec_e, spec_w, bow_a, bow_w;
    std::vector<dxg_spectrum> spectra;
    std::vector<dxg_bowtie> bowties;
    dxg_beam view {};

    int copy(const dxg_beam* b)
    {
        if (!b || !b->exposures || b->n_exposures == 0 || !b->spectra || b->n_spectra == 0)
            return fail(DXG_ERR_ARGUMENT, "dxg_batch_add: beam needs exposures and spectra");
        exposures.assign(b->exposures, b->exposures + b->n_exposures);
        for (uint64_t i = 0; i < b->n_spectra; ++i) {
            const dxg_spectrum& s = b->spectra[i];
            if (s.n == 0 || !s.energy_kev || !s.weight)
                return fail(DXG_ERR_ARGUMENT, "dxg_batch_add: empty spectrum");
            spec_e.emplace_back(s.energy_kev, s.energy_kev + s.n);
            spec_w.emplace_back(s.weight, s.weight + s.n);
        }
        for (uint64_t i = 0; i < b->n_bowties; ++i) {
            const dxg_bowtie& t = b->bowties[i];
            if (t.n == 0 || !t.angle_rad || !t.weight)
                return fail(DXG_ERR_ARGUMENT, "dxg_batch_add: empty bowtie");
            bow_a.emplace_back(t.angle_rad, t.angle_rad + t.n);
            bow_w.emplace_back(t.weight, t.weight + t.n);
        }
        for (size_t i = 0; i < spec_e.size(); ++i)
            spectra.push_back({ spec_e[i].data(), spec_w[i].data(), spec_e[i].size() });
        for (size_t i = 0; i < bow_a.size(); ++i)
            bowties.push_back({ bow_a[i].data(), bow_w[i].data(), bow_a[i].size() });
        view = { exposures.data(), exposures.size(), spectra.data(), spectra.size(),
                 bowties.empty() ? nullptr : bowties.data(), bowties.size(), b->calibration };
        return DXG_OK;
    }
};

struct Instance {
    dxg_world* world;
    BeamCopy beam;
    uint64_t seed;
    uint64_t histories = 0;
};

}  // namespace

struct dxg_batch {
    std::vector<Instance> inst;
    int device = -1;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    dxg::DevParams* d_params = nullptr;
    int* d_start = nullptr;
    size_t cap = 0, start_cap = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // launch shape the graph was captured with
    int g_blocks = 0, g_level = -1, g_n = -1;
    bool g_lut = false, g_cells = false;
    size_t g_lds = 0;

    void dropGraph()
    {
        if (exec)
            (void)hipGraphExecDestroy(exec);
        if (graph)
            (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
    ~dxg_batch()
    {
        if (stream)
            (void)hipStreamSynchronize(stream);
        dropGraph();
        if (d_params)
            (void)hipFree(d_params);
        if (d_start)
            (void)hipFree(d_start);
        if (ev0)
            (void)hipEventDestroy(ev0);
        if (ev1)
            (void)hipEventDestroy(ev1);
        if (stream)
            (void)hipStreamDestroy(stream);
    }
};

extern "C" {

int dxg_batch_create(dxg_batch** out)
{
    if (!out)
        return fail(DXG_ERR_ARGUMENT, "dxg_batch_create: null output pointer");
    *out = new dxg_batch();
    return DXG_OK;
}

void dxg_batch_destroy(dxg_batch* b) { delete b; }

int dxg_batch_add(dxg_batch* b, dxg_world* w, const dxg_beam* beam, uint64_t seed)
{
    if (!b || !w || !beam)
        return fail(DXG_ERR_ARGUMENT, "dxg_batch_add: null argument");
    for (const Instance& i : b->inst)
        if (i.world == w)
            return fail(DXG_ERR_ARGUMENT, "dxg_batch_add: a world can hold one instance of a batch");
    const int dev = dxg::worldDevice(w);
    if (b->device >= 0 && dev != b->device)
        return fail(DXG_ERR_ARGUMENT, "dxg_batch_add: every world of a batch must live on one device");
    Instance in { w, {}, seed, 0 };
    if (int rc = in.beam.copy(beam); rc != DXG_OK)
        return rc;
    // validate now (uploads the tables once; every run uploads them again)
    dxg::DevParams p;
    if (int rc = dxg::prepareRun(w, &in.beam.view, seed, p, in.histories); rc != DXG_OK)
        return rc;
    if (!b->inst.empty()) {
        dxg::DevParams q;
        uint64_t t = 0;
        if (int rc = dxg::prepareRun(b->inst.front().world, &b->inst.front().beam.view, 0, q, t); rc != DXG_OK)
            return rc;
        if (q.level != p.level)
            return fail(DXG_ERR_ARGUMENT, "dxg_batch_add: every world of a batch must have the same correction level");
    }
    b->device = dev;
    // BeamCopy holds pointers into its own vectors: build the view after the move
    b->inst.push_back(std::move(in));
    Instance& last = b->inst.back();
    last.beam.view.exposures = last.beam.exposures.data();
    last.beam.spectra.clear();
    for (size_t i = 0; i < last.beam.spec_e.size(); ++i)
        last.beam.spectra.push_back({ last.beam.spec_e[i].data(), last.beam.spec_w[i].data(), last.beam.spec_e[i].size() });
    last.beam.bowties.clear();
    for (size_t i = 0; i < last.beam.bow_a.size(); ++i)
        last.beam.bowties.push_back({ last.beam.bow_a[i].data(), last.beam.bow_w[i].data(), last.beam.bow_a[i].size() });
    last.beam.view.spectra = last.beam.spectra.data();
    last.beam.view.bowties = last.beam.bowties.empty() ? nullptr : last.beam.bowties.data();
    b->dropGraph();
    return DXG_OK;
}

uint64_t dxg_batch_size(const dxg_batch* b) { return b ? b->inst.size() : 0; }

int dxg_batch_run(dxg_batch* b, double* kernel_ms)
{
    if (!b)
        return fail(DXG_ERR_ARGUMENT, "dxg_batch_run: null batch");
    if (b->inst.empty())
        return fail(DXG_ERR_STATE, "dxg_batch_run: the batch is empty");
    HIP_TRY(hipSetDevice(b->device));
    if (!b->stream) {
        HIP_TRY(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
        HIP_TRY(hipEventCreate(&b->ev0));
        HIP_TRY(hipEventCreate(&b->ev1));
    }
    const int n = static_cast<int>(b->inst.size());
    // per-run parameters: tables uploaded again (the world may have run other
    // beams since), RNG offset = the world's histories so far
    std::vector<dxg::DevParams> ps(static_cast<size_t>(n));
    std::vector<double> work(static_cast<size_t>(n));
    bool lut = true;
    int level = -1;
    for (int i = 0; i < n; ++i) {
        Instance& in = b->inst[static_cast<size_t>(i)];
        uint64_t total = 0;
        if (int rc = dxg::prepareRun(in.world, &in.beam.view, in.seed, ps[static_cast<size_t>(i)], total); rc != DXG_OK)
            return rc;
        in.histories = total;
        if (in.beam.view.calibration > 0.0)
            ps[static_cast<size_t>(i)].dose_scale = static_cast<float>(in.beam.view.calibration);
        ps[static_cast<size_t>(i)].h_begin = 0;
        ps[static_cast<size_t>(i)].h_end = total;
        lut = lut && ps[static_cast<size_t>(i)].lds_lut != 0;
        level = ps[static_cast<size_t>(i)].level;
        work[static_cast<size_t>(i)] = static_cast<double>(total);
    }
    // one kernel instance serves the batch: macro cells only when every world has them
    bool cells = true;
    for (const auto& p : ps)
        cells = cells && p.cell_ratio != nullptr;
    size_t lds = 0;
    for (auto& p : ps) {
        p.lds_lut = lut ? 1 : 0;
        if (!cells)
            p.cell_ratio = nullptr;
        lds = std::max(lds, dxg::transportLdsBytes(p));
    }
    HIP_TRY(hipDeviceSynchronize());   // the worlds' table uploads are on their own streams

    int cus = 0;
    HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device));
    const int blocks = cus * dxg::transportBatchBlocksPerCU(lut, level, cells, lds);
    // initial block -> instance assignment proportional to each instance's histories
    std::vector<int> start(static_cast<size_t>(blocks));
    const double tot = std::accumulate(work.begin(), work.end(), 0.0);
    {
        double acc = 0.0;
        int i = 0;
        for (int k = 0; k < blocks; ++k) {
            const double pos = (k + 0.5) / blocks * tot;
            while (i < n - 1 && acc + work[static_cast<size_t>(i)] <= pos) {
                acc += work[static_cast<size_t>(i)];
                ++i;
            }
            start[static_cast<size_t>(k)] = i;
        }
    }
    if (b->cap < static_cast<size_t>(n)) {
        if (b->d_params)
            (void)hipFree(b->d_params);
        b->d_params = nullptr;
        HIP_TRY(hipMalloc(reinterpret_cast<void**>(&b->d_params), sizeof(dxg::DevParams) * static_cast<size_t>(n)));
        b->cap = static_cast<size_t>(n);
        b->dropGraph();
    }
    if (b->start_cap < static_cast<size_t>(blocks)) {
        if (b->d_start)
            (void)hipFree(b->d_start);
        b->d_start = nullptr;
        HIP_TRY(hipMalloc(reinterpret_cast<void**>(&b->d_start), sizeof(int) * static_cast<size_t>(blocks)));
        b->start_cap = static_cast<size_t>(blocks);
        b->dropGraph();
    }
    HIP_TRY(hipMemcpyAsync(b->d_params, ps.data(), sizeof(dxg::DevParams) * static_cast<size_t>(n), hipMemcpyHostToDevice, b->stream));
    HIP_TRY(hipMemcpyAsync(b->d_start, start.data(), sizeof(int) * static_cast<size_t>(blocks), hipMemcpyHostToDevice, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));

    // the reset and transport kernels read everything from d_params, so one
    // captured graph serves every run with the same launch shape
    if (b->exec && (b->g_blocks != blocks || b->g_level != level || b->g_lut != lut || b->g_cells != cells || b->g_lds != lds
                    || b->g_n != n))
        b->dropGraph();
    if (!b->exec) {
        HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        dxg::launchTransportBatch(b->d_params, n, b->d_start, lut, level, cells, lds, blocks, b->stream);
        hipGraph_t g = nullptr;
        const hipError_t launch_err = hipGetLastError();
        const hipError_t cap_err = hipStreamEndCapture(b->stream, &g);
        if (launch_err != hipSuccess || cap_err != hipSuccess) {
            if (g)
                (void)hipGraphDestroy(g);
            return fail(DXG_ERR_DEVICE, std::string("dxg_batch_run: graph capture failed: ")
                                            + hipGetErrorString(launch_err != hipSuccess ? launch_err : cap_err));
        }
        b->graph = g;
        HIP_TRY(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
        b->g_blocks = blocks;
        b->g_level = level;
        b->g_lut = lut;
        b->g_cells = cells;
        b->g_lds = lds;
        b->g_n = n;
    }
    HIP_TRY(hipEventRecord(b->ev0, b->stream));
    HIP_TRY(hipGraphLaunch(b->exec, b->stream));
    HIP_TRY(hipEventRecord(b->ev1, b->stream));
    std::vector<dxg::DevCounters> cs(static_cast<size_t>(n));
    for (int i = 0; i < n; ++i)
        HIP_TRY(hipMemcpyAsync(&cs[static_cast<size_t>(i)], ps[static_cast<size_t>(i)].counters, sizeof(dxg::DevCounters),
                               hipMemcpyDeviceToHost, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    for (int i = 0; i < n; ++i) {
        const dxg::DevCounters& c = cs[static_cast<size_t>(i)];
        dxg_counters acc {};
        acc.histories = c.histories;
        acc.steps = c.steps;
        acc.real_events = c.real_events;
        acc.deposits = c.deposits;
        acc.emitted_energy = c.emitted;
        acc.escaped_energy = c.escaped;
        acc.deposited_energy = c.deposited;
        acc.uncollided = c.uncollided;
        acc.cell_crossings = c.crossings;
        for (int k = 0; k < 4; ++k) {
            acc.phase_cycles[k] = c.phase_cycles[k];
            acc.phase_lanes[k] = c.phase_lanes[k];
            acc.phase_iters[k] = c.phase_iters[k];
        }
        acc.kernel_ms = ms;          // the shared launch
        acc.kernel_launches = 1;
        dxg::addHistories(b->inst[static_cast<size_t>(i)].world, b->inst[static_cast<size_t>(i)].histories, acc);
        if (int rc = dxg::foldTally(b->inst[static_cast<size_t>(i)].world); rc != DXG_OK)
            return rc;
    }
    if (kernel_ms)
        *kernel_ms = ms;
    return DXG_OK;
}

}  // extern "C"
