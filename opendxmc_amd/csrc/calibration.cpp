// CT beam calibration from a simulated CTDI phantom (include/dxmc_beams.h,
// dxg_calibration_ctdi).  The reference keeps CTDIw / CTDIvol and the CTDI
// phantom diameter on its CT beams (beamsettingsmodel.cpp:1054-1071,
// 1276-1295) and lets DXMClib turn them into a dose scale; this file builds
// the PMMA phantom as a voxel world of its own and measures the beam in it.
#include "dxmc_beams.h"
#include "dxmc_hip.h"

#include "physics.hpp"

#include <algorithm>
#include <cmath>
#include <memory>
#include <string>
#include <vector>

namespace {

int fail(int code, const std::string& msg) { return dxg::setError(code, msg); }

struct WorldDeleter {
    void operator()(dxg_world* w) const { dxg_world_destroy(w); }
};

constexpr double kVoxel = 0.25;          // cm
constexpr double kLength = 15.0;         // phantom length, cm
constexpr double kChamberHalf = 5.0;     // 100 mm chambers
constexpr double kChamberRadius = 0.5;   // cm
constexpr double kDepth = 1.0;           // peripheral chambers below the surface, cm
constexpr double kPmmaDensity = 1.19;
constexpr double kAirDensity = 0.001205;

}  // namespace

extern "C" int dxg_calibration_ctdi(const dxg_beam* rotation, double ctdi_w_mgy, double diameter_cm, double collimation_cm,
                                    uint64_t run_particles_per_exposure, uint64_t seed, double* factor, double* ctdiw_sim)
{
    if (!factor)
        return fail(DXG_ERR_ARGUMENT, "dxg_calibration_ctdi: null output");
    if (!rotation || !rotation->exposures || rotation->n_exposures == 0)
        return fail(DXG_ERR_ARGUMENT, "dxg_calibration_ctdi: the rotation has no exposures");
    if (!(ctdi_w_mgy > 0.0) || !(collimation_cm > 0.0) || run_particles_per_exposure == 0)
        return fail(DXG_ERR_ARGUMENT, "dxg_calibration_ctdi: CTDIw, collimation and particle count must be positive");
    if (!(diameter_cm >= 4.0 && diameter_cm <= 60.0))
        return fail(DXG_ERR_ARGUMENT, "dxg_calibration_ctdi: phantom diameter must be within 4..60 cm");
    const uint64_t n_c = rotation->exposures[0].n_particles;
    for (uint64_t i = 1; i < rotation->n_exposures; ++i)
        if (rotation->exposures[i].n_particles != n_c)
            return fail(DXG_ERR_ARGUMENT, "dxg_calibration_ctdi: rotation exposures must share one particle count");

    // ---- the phantom: PMMA cylinder along z in air, chambers of PMMA
    const int nxy = 2 * static_cast<int>(std::ceil((0.5 * diameter_cm + 1.0) / kVoxel));
    const int nz = 2 * static_cast<int>(std::ceil((0.5 * kLength + 0.5) / kVoxel));
    const uint64_t dims[3] = { static_cast<uint64_t>(nxy), static_cast<uint64_t>(nxy), static_cast<uint64_t>(nz) };
    const size_t n = static_cast<size_t>(nxy) * nxy * nz;
    std::vector<double> density(n, kAirDensity);
    std::vector<uint8_t> material(n, 0);
    std::vector<int8_t> chamber(n, -1);   // 0 centre, 1..4 periphery
    const double R = 0.5 * diameter_cm;
    const double cx[5] = { 0.0, R - kDepth, 0.0, -(R - kDepth), 0.0 };
    const double cy[5] = { 0.0, 0.0, R - kDepth, 0.0, -(R - kDepth) };
    for (int k = 0; k < nz; ++k) {
        const double z = (k + 0.5 - 0.5 * nz) * kVoxel;
        for (int j = 0; j < nxy; ++j) {
            const double y = (j + 0.5 - 0.5 * nxy) * kVoxel;
            for (int i = 0; i < nxy; ++i) {
                const double x = (i + 0.5 - 0.5 * nxy) * kVoxel;
                const size_t v = (static_cast<size_t>(k) * nxy + j) * nxy + i;
                if (x * x + y * y <= R * R && std::fabs(z) <= 0.5 * kLength) {
                    material[v] = 1;
                    density[v] = kPmmaDensity;
                    if (std::fabs(z) <= kChamberHalf)
                        for (int c = 0; c < 5; ++c)
                            if ((x - cx[c]) * (x - cx[c]) + (y - cy[c]) * (y - cy[c]) <= kChamberRadius * kChamberRadius)
                                chamber[v] = static_cast<int8_t>(c);
                }
            }
        }
    }
    const int32_t airZ[4] = { 6, 7, 8, 18 };
    const double airW[4] = { 0.000124, 0.755268, 0.231781, 0.012827 };
    const int32_t pmmaZ[3] = { 1, 6, 8 };
    const double pmmaW[3] = { 0.080538, 0.599848, 0.319614 };
    const dxg_material mats[2] = { { airZ, airW, 4 }, { pmmaZ, pmmaW, 3 } };
    const double spacing[3] = { kVoxel, kVoxel, kVoxel };

    dxg_world* raw = nullptr;
    if (int rc = dxg_world_create(&raw, 1); rc != DXG_OK)
        return rc;
    std::unique_ptr<dxg_world, WorldDeleter> w(raw);
    if (int rc = dxg_grid_set_data(w.get(), dims, density.data(), material.data(), mats, 2); rc != DXG_OK)
        return rc;
    if (int rc = dxg_grid_set_spacing(w.get(), spacing); rc != DXG_OK)
        return rc;
    if (int rc = dxg_world_build(w.get()); rc != DXG_OK)
        return rc;
    if (int rc = dxg_transport_run(w.get(), rotation, nullptr, 0, seed); rc != DXG_OK)
        return rc;
    std::vector<double> e(n);
    if (int rc = dxg_read_energy_scores(w.get(), e.data(), nullptr, nullptr); rc != DXG_OK)
        return rc;

    // ---- chamber doses (energy over mass) -> CTDI100 -> CTDIw
    double energy[5] = { 0, 0, 0, 0, 0 }, mass[5] = { 0, 0, 0, 0, 0 };
    const double voxel_kg = kVoxel * kVoxel * kVoxel * kPmmaDensity * 1e-3;
    for (size_t v = 0; v < n; ++v)
        if (chamber[v] >= 0) {
            energy[chamber[v]] += e[v];
            mass[chamber[v]] += voxel_kg;
        }
    double ctdi[5];
    for (int c = 0; c < 5; ++c) {
        if (!(mass[c] > 0.0))
            return fail(DXG_ERR_ARGUMENT, "dxg_calibration_ctdi: phantom too small for its chambers");
        const double dose_mgy = energy[c] * dxg::kKeVToJ / mass[c] * 1e3;
        ctdi[c] = dose_mgy * (2.0 * kChamberHalf) / collimation_cm;
    }
    const double sim = ctdi[0] / 3.0 + (2.0 / 3.0) * 0.25 * (ctdi[1] + ctdi[2] + ctdi[3] + ctdi[4]);
    if (!(sim > 0.0))
        return fail(DXG_ERR_DATA, "dxg_calibration_ctdi: no dose reached the chambers (check the beam geometry)");
    if (ctdiw_sim)
        *ctdiw_sim = sim;
    *factor = ctdi_w_mgy / sim * static_cast<double>(n_c) / static_cast<double>(run_particles_per_exposure);
    return DXG_OK;
}
