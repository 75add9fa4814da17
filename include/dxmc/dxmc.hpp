// C++ facade of the MI355X transport engine with DXMClib's call surface, as
// OpenDXMC uses it (simulationpipeline.cpp:124-235, dxmc_specialization.hpp:19-77,
// beamsettingsmodel.cpp:257-1837).  Header-only; everything forwards to the
// C-ABI of libdxmc_hip.so (dxmc_hip.h, dxmc_beams.h).  Like the reference, no
// function throws: failures are reported through std::optional, through the
// TransportProgress stop flag, and through lastError().
#pragma once

#include "dxmc_beams.h"
#include "dxmc_hip.h"

#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <optional>
#include <random>
#include <string>
#include <utility>
#include <vector>

namespace dxmc {

template <typename T = double>
constexpr T PI_VAL() { return T { 3.14159265358979323846 }; }
template <typename T = double>
constexpr T DEG_TO_RAD() { return PI_VAL<T>() / T { 180 }; }
template <typename T = double>
constexpr T RAD_TO_DEG() { return T { 180 } / PI_VAL<T>(); }

inline std::string lastError() { return dxg_last_error(); }

// ------------------------------------------------------------------ material
struct AttenuationValues {
    double photoelectric = 0, incoherent = 0, coherent = 0;
    double sum() const { return photoelectric + incoherent + coherent; }
};

template <std::size_t NMaterialShells = 5>
class Material {
public:
    // Material<5>::byWeight (simulationpipeline.cpp:136): std::nullopt when the
    // composition cannot be built.
    static std::optional<Material> byWeight(const std::map<std::uint64_t, double>& composition)
    {
        Material m;
        for (const auto& [z, w] : composition) {
            m.m_Z.push_back(static_cast<int32_t>(z));
            m.m_w.push_back(w);
        }
        const double e = 60.0;
        double out[4];
        if (dxg_material_attenuation(&m.view(), &e, 1, out) != DXG_OK)
            return std::nullopt;
        return m;
    }
    // Material::attenuationValues(E) (ctsegmentationpipeline.cpp:73-98), cm2/g
    AttenuationValues attenuationValues(double energy_kev) const
    {
        double out[4] = { 0, 0, 0, 0 };
        dxg_material_attenuation(&view(), &energy_kev, 1, out);
        return { out[0], out[1], out[2] };
    }
    const dxg_material& view() const
    {
        m_view = { m_Z.data(), m_w.data(), static_cast<uint64_t>(m_Z.size()) };
        return m_view;
    }

private:
    std::vector<int32_t> m_Z;
    std::vector<double> m_w;
    mutable dxg_material m_view {};
};

// ------------------------------------------------------------------ progress
class TransportProgress {
public:
    TransportProgress() { dxg_progress_create(&m_p); }
    ~TransportProgress() { dxg_progress_destroy(m_p); }
    TransportProgress(const TransportProgress&) = delete;
    TransportProgress& operator=(const TransportProgress&) = delete;
    std::pair<std::uint64_t, std::uint64_t> progress() const
    {
        uint64_t n = 0, t = 0;
        dxg_progress_get(m_p, &n, &t);
        return { n, t };
    }
    std::string message() const
    {
        char buf[256];
        dxg_progress_message(m_p, buf, sizeof buf);
        return buf;
    }
    bool continueSimulation() const { return dxg_progress_continue(m_p) != 0; }
    void setStopSimulation() { dxg_progress_request_stop(m_p); }
    dxg_progress* handle() const { return m_p; }

private:
    dxg_progress* m_p = nullptr;
};

// ------------------------------------------------------------------- scores
class EnergyScore {
public:
    double energyImparted() const { return m_e; }
    double energyImpartedSquared() const { return m_e2; }
    std::uint64_t numberOfEvents() const { return m_n; }
    double m_e = 0, m_e2 = 0;
    std::uint64_t m_n = 0;
};

class DoseScore {
public:
    double dose() const { return m_dose; }
    double variance() const { return m_variance; }
    std::uint64_t numberOfEvents() const { return m_n; }
    double m_dose = 0, m_variance = 0;
    std::uint64_t m_n = 0;
};

// --------------------------------------------------------------- voxel grid
// AAVoxelGrid<NMaterialShells, LOWENERGYCORRECTION, TRANSPARENTVOXELS>
// (simulationpipeline.cpp:127).  Material indices are validated below the
// number of materials, so the transparent id never occurs.
template <int NMaterialShells = 5, int LOWENERGYCORRECTION = 1, std::uint8_t TRANSPARENTVOXELS = 255>
class AAVoxelGrid {
public:
    AAVoxelGrid() { m_status = dxg_world_create(&m_w, LOWENERGYCORRECTION); }
    ~AAVoxelGrid() { dxg_world_destroy(m_w); }
    AAVoxelGrid(const AAVoxelGrid&) = delete;
    AAVoxelGrid& operator=(const AAVoxelGrid&) = delete;

    bool setData(const std::array<std::size_t, 3>& dims, const std::vector<double>& density,
                 const std::vector<std::uint8_t>& materialIdx, const std::vector<Material<NMaterialShells>>& materials)
    {
        if (m_status != DXG_OK)
            return false;
        const uint64_t d[3] = { dims[0], dims[1], dims[2] };
        std::vector<dxg_material> mats;
        for (const auto& m : materials)
            mats.push_back(m.view());
        const uint64_t n = d[0] * d[1] * d[2];
        if (density.size() != n || materialIdx.size() != n)
            return false;
        m_status = dxg_grid_set_data(m_w, d, density.data(), materialIdx.data(), mats.data(), mats.size());
        m_dims = dims;
        return m_status == DXG_OK;
    }
    bool setSpacing(const std::array<double, 3>& spacing)
    {
        if (m_status != DXG_OK)
            return false;
        m_status = dxg_grid_set_spacing(m_w, spacing.data());
        return m_status == DXG_OK;
    }
    std::size_t size() const { return static_cast<std::size_t>(dxg_grid_size(m_w)); }
    const std::array<std::size_t, 3>& dimensions() const { return m_dims; }

    const DoseScore& doseScored(std::size_t i) const
    {
        refresh();
        return m_dose[i];
    }
    const EnergyScore& energyScored(std::size_t i) const
    {
        refresh();
        return m_energy[i];
    }
    void clearDoseScored()
    {
        dxg_world_clear_scores(m_w);
        m_stale = true;
    }
    int status() const { return m_status; }
    dxg_world* handle() { return m_w; }
    void markStale() { m_stale = true; }
    bool build()
    {
        if (m_status == DXG_OK)
            m_status = dxg_world_build(m_w);
        m_stale = true;
        return m_status == DXG_OK;
    }

private:
    void refresh() const
    {
        if (!m_stale)
            return;
        const std::size_t n = size();
        std::vector<double> d(n), v(n), e(n), e2(n);
        std::vector<uint64_t> c(n);
        m_dose.assign(n, DoseScore {});
        m_energy.assign(n, EnergyScore {});
        if (dxg_read_dose_scores(m_w, d.data(), v.data(), c.data()) == DXG_OK
            && dxg_read_energy_scores(m_w, e.data(), e2.data(), nullptr) == DXG_OK) {
            for (std::size_t i = 0; i < n; ++i) {
                m_dose[i].m_dose = d[i];
                m_dose[i].m_variance = v[i];
                m_dose[i].m_n = c[i];
                m_energy[i].m_e = e[i];
                m_energy[i].m_e2 = e2[i];
                m_energy[i].m_n = c[i];
            }
        }
        m_stale = false;
    }
    dxg_world* m_w = nullptr;
    int m_status = DXG_OK;
    std::array<std::size_t, 3> m_dims { 0, 0, 0 };
    mutable bool m_stale = true;
    mutable std::vector<DoseScore> m_dose;
    mutable std::vector<EnergyScore> m_energy;
};

// -------------------------------------------------------------------- world
// World<Items...>: owns its items (simulationpipeline.cpp:128-131).  The
// engine transports through its voxel grid item.
template <typename... Us>
class World {
public:
    template <typename U>
    U& addItem()
    {
        auto item = std::make_unique<U>();
        U& ref = *item;
        m_items.push_back(std::shared_ptr<void>(item.release(), [](void* p) { delete static_cast<U*>(p); }));
        m_build = [&ref]() { return ref.build(); };
        m_grid = [&ref]() { return ref.handle(); };
        m_stale = [&ref]() { ref.markStale(); };
        return ref;
    }
    bool build() { return m_build ? m_build() : false; }
    dxg_world* gridHandle() { return m_grid ? m_grid() : nullptr; }
    void scoresChanged()
    {
        if (m_stale)
            m_stale();
    }

private:
    std::vector<std::shared_ptr<void>> m_items;
    std::function<bool()> m_build;
    std::function<dxg_world*()> m_grid;
    std::function<void()> m_stale;
};

// ---------------------------------------------------------------- filters
class BowtieFilter {
public:
    BowtieFilter() = default;
    explicit BowtieFilter(const std::vector<std::pair<double, double>>& data)
    {
        for (const auto& [a, w] : data) {
            m_angle.push_back(a);
            m_weight.push_back(w);
        }
    }
    const std::vector<double>& angles() const { return m_angle; }
    const std::vector<double>& weights() const { return m_weight; }
    bool empty() const { return m_angle.empty(); }

private:
    std::vector<double> m_angle, m_weight;
};

class CTAECFilter {
public:
    CTAECFilter() = default;
    CTAECFilter(const std::array<double, 3>& start, const std::array<double, 3>& stop, const std::vector<double>& weights)
        : m_start(start)
        , m_stop(stop)
        , m_weights(weights)
    {
    }
    void setData(const std::array<double, 3>& start, const std::array<double, 3>& stop, const std::vector<double>& weights)
    {
        m_start = start;
        m_stop = stop;
        m_weights = weights;
    }
    bool isEmpty() const { return m_weights.empty(); }
    dxg_aec view() const
    {
        dxg_aec a {};
        for (int i = 0; i < 3; ++i) {
            a.start[i] = m_start[i];
            a.stop[i] = m_stop[i];
        }
        a.weights = m_weights.data();
        a.n = m_weights.size();
        return a;
    }

private:
    std::array<double, 3> m_start { 0, 0, 0 }, m_stop { 0, 0, 1 };
    std::vector<double> m_weights;
};

// The organ AEC filter of CT spiral beams (organAECFilter(); settings at
// beamsettingsmodel.cpp:349-437): weight by gantry angle, dxg_organ_aec.
class CTOrganAECFilter {
public:
    bool useFilter() const { return m_f.use != 0; }
    void setUseFilter(bool on) { m_f.use = on ? 1 : 0; }
    bool compensateOutside() const { return m_f.compensate_outside != 0; }
    void setCompensateOutside(bool on) { m_f.compensate_outside = on ? 1 : 0; }
    double startAngle() const { return m_f.start_angle; }
    double stopAngle() const { return m_f.stop_angle; }
    double rampAngle() const { return m_f.ramp_angle; }
    void setStartAngle(double rad) { m_f.start_angle = rad; }
    void setStopAngle(double rad) { m_f.stop_angle = rad; }
    void setRampAngle(double rad) { m_f.ramp_angle = rad; }
    double startAngleDeg() const { return m_f.start_angle * 57.29577951308232; }
    double stopAngleDeg() const { return m_f.stop_angle * 57.29577951308232; }
    double rampAngleDeg() const { return m_f.ramp_angle * 57.29577951308232; }
    void setStartAngleDeg(double d) { m_f.start_angle = d / 57.29577951308232; }
    void setStopAngleDeg(double d) { m_f.stop_angle = d / 57.29577951308232; }
    void setRampAngleDeg(double d) { m_f.ramp_angle = d / 57.29577951308232; }
    double lowWeight() const { return m_f.low_weight; }
    void setLowWeightFactor(double w) { m_f.low_weight = w; }
    double maxWeight() const
    {
        double w = 1.0;
        return dxg_organ_aec_max_weight(&m_f, &w) == DXG_OK ? w : 1.0;
    }
    const dxg_organ_aec& view() const { return m_f; }

private:
    dxg_organ_aec m_f { 0, 0, 0.0, 0.0, 0.0, 1.0 };
};

// DataContainer::calculateAECfilterFromWaterEquivalentDiameter
// (datacontainer.cpp:42-100) for an image of dims (index x + y nx + z nx ny)
// and spacing in cm: a CT image in HU, or densities in g/cm3.  Empty on bad input.
inline CTAECFilter calculateAECfilterFromWaterEquivalentDiameter(const std::array<std::size_t, 3>& dims,
                                                                 const std::array<double, 3>& spacing,
                                                                 const std::vector<double>& image, bool isCT)
{
    const std::uint64_t d[3] = { dims[0], dims[1], dims[2] };
    if (image.size() != dims[0] * dims[1] * dims[2])
        return CTAECFilter {};
    std::vector<double> w(dims[2]);
    std::array<double, 3> start {}, stop {};
    if (dxg_aec_from_wed(d, spacing.data(), image.data(), isCT ? 1 : 0, w.data(), start.data(), stop.data()) != DXG_OK)
        return CTAECFilter {};
    return CTAECFilter(start, stop, w);
}

// --------------------------------------------------------------------- tube
class Tube {
public:
    Tube(double voltage = 120.0, double anodeAngleDeg = 12.0, const std::map<std::size_t, double>& filtration = {})
        : m_voltage(voltage)
        , m_anode(anodeAngleDeg)
        , m_filtration(filtration)
    {
    }
    double voltage() const { return m_voltage; }
    void setVoltage(double kv) { m_voltage = kv; }
    double anodeAngleDeg() const { return m_anode; }
    void setAnodeAngleDeg(double a) { m_anode = a; }
    void addFiltrationMaterial(std::size_t Z, double mm) { m_filtration[Z] = std::abs(mm); }
    double filtration(std::size_t Z) const
    {
        auto it = m_filtration.find(Z);
        return it == m_filtration.end() ? 0.0 : it->second;
    }
    // Tube::getSpecter: (energies, normalised weights)
    std::pair<std::vector<double>, std::vector<double>> getSpecter() const
    {
        auto t = view();
        uint64_t n = 0;
        std::vector<double> e, w;
        if (dxg_tube_spectrum(&t.first, nullptr, nullptr, 0, &n) != DXG_OK)
            return {};
        e.resize(n);
        w.resize(n);
        dxg_tube_spectrum(&t.first, e.data(), w.data(), n, &n);
        return { e, w };
    }
    double mmAlHalfValueLayer() const
    {
        auto t = view();
        double v = 0;
        dxg_tube_hvl_mm_al(&t.first, &v);
        return v;
    }
    double meanSpecterEnergy() const
    {
        auto t = view();
        double v = 0;
        dxg_tube_mean_energy(&t.first, &v);
        return v;
    }

private:
    struct Arrays {
        std::vector<int32_t> z;
        std::vector<double> mm;
    };
    std::pair<dxg_tube, std::shared_ptr<Arrays>> view() const
    {
        auto a = std::make_shared<Arrays>();
        for (const auto& [z, mm] : m_filtration) {
            a->z.push_back(static_cast<int32_t>(z));
            a->mm.push_back(mm);
        }
        dxg_tube t { m_voltage, m_anode, a->z.data(), a->mm.data(), a->z.size(), 0.5 };
        return { t, a };
    }
    double m_voltage, m_anode;
    std::map<std::size_t, double> m_filtration;
};

// -------------------------------------------------------------------- beams
// What Transport needs from any beam: exposures, spectra and bowties.
struct BeamTables {
    std::vector<dxg_exposure> exposures;
    std::vector<std::vector<double>> energies, weights;
    std::vector<BowtieFilter> bowties;

    // A dxg_beam over these tables, valid while both objects live.
    struct View {
        std::vector<dxg_spectrum> spectra;
        std::vector<dxg_bowtie> bowties;
        dxg_beam beam {};
    };
    std::unique_ptr<View> view(double calibration = 0.0) const
    {
        auto v = std::make_unique<View>();
        for (std::size_t i = 0; i < energies.size(); ++i)
            v->spectra.push_back({ energies[i].data(), weights[i].data(), energies[i].size() });
        for (const auto& b : bowties)
            v->bowties.push_back({ b.angles().data(), b.weights().data(), b.angles().size() });
        v->beam = { exposures.data(), exposures.size(), v->spectra.data(), v->spectra.size(),
                    v->bowties.empty() ? nullptr : v->bowties.data(), v->bowties.size(), calibration };
        return v;
    }
};

namespace detail {
// Calibration factor of a beam's tables from a closed-form entry point (DAP, air KERMA).
inline double closedFormCalibration(const BeamTables& t, double value, int (*fn)(const dxg_beam*, double, double*))
{
    if (!(value > 0))
        return 0.0;
    const auto v = t.view();
    double f = 0.0;
    return fn(&v->beam, value, &f) == DXG_OK ? f : 0.0;
}

// One axial rotation at the origin of a CT tube (the CTDI phantom run).
struct RotationTube {
    const BeamTables* tube;   // spectrum (and bowtie) of this tube in entries 0
    double start_angle, fov, mas;
    bool bowtie;
};

inline double ctdiCalibrationRun(const std::vector<RotationTube>& tubes, double step, double coll, double sdd, double ctdi_w,
                                 double diameter, std::uint64_t run_ppe);

inline double ctdiCalibration(const std::vector<RotationTube>& tubes, double step, double coll, double sdd, double ctdi_w,
                              double diameter, std::uint64_t run_ppe)
{
    if (!(ctdi_w > 0))
        return 0.0;
    // The factor depends on these settings only: a CTDI phantom run per
    // distinct setting, cached for the process (every Transport call asks).
    std::vector<double> key { step, coll, sdd, ctdi_w, diameter, static_cast<double>(run_ppe) };
    for (const auto& rt : tubes) {
        key.insert(key.end(), { rt.start_angle, rt.fov, rt.mas, rt.bowtie ? 1.0 : 0.0 });
        key.insert(key.end(), rt.tube->energies[0].begin(), rt.tube->energies[0].end());
        key.insert(key.end(), rt.tube->weights[0].begin(), rt.tube->weights[0].end());
        if (rt.bowtie) {
            key.insert(key.end(), rt.tube->bowties[0].angles().begin(), rt.tube->bowties[0].angles().end());
            key.insert(key.end(), rt.tube->bowties[0].weights().begin(), rt.tube->bowties[0].weights().end());
        }
        key.push_back(-1.0);   // separator between tubes
    }
    static std::mutex cache_mutex;
    static std::map<std::vector<double>, double> cache;
    {
        std::lock_guard<std::mutex> lock(cache_mutex);
        if (auto it = cache.find(key); it != cache.end())
            return it->second;
    }
    const double factor = ctdiCalibrationRun(tubes, step, coll, sdd, ctdi_w, diameter, run_ppe);
    if (factor > 0) {
        std::lock_guard<std::mutex> lock(cache_mutex);
        cache.emplace(std::move(key), factor);
    }
    return factor;
}

// The CTDI phantom run behind ctdiCalibration (uncached).
inline double ctdiCalibrationRun(const std::vector<RotationTube>& tubes, double step, double coll, double sdd, double ctdi_w,
                                 double diameter, std::uint64_t run_ppe)
{
    constexpr std::uint64_t kPpe = 100000;   // particles per rotation exposure
    BeamTables rot;
    for (const auto& rt : tubes) {
        dxg_ct_sequential_beam p {};
        p.scan_normal[2] = 1.0;
        p.number_of_slices = 1;
        p.slice_spacing_cm = 1.0;
        p.start_angle = rt.start_angle;
        p.step_angle = step;
        p.collimation_cm = coll;
        p.sdd_cm = sdd;
        p.fov_cm = rt.fov;
        p.particles_per_exposure = kPpe;
        p.has_bowtie = rt.bowtie ? 1 : 0;
        uint64_t n = 0;
        if (dxg_ct_sequential_exposures(&p, nullptr, 0, &n) != DXG_OK)
            return 0.0;
        std::vector<dxg_exposure> ex(n);
        dxg_ct_sequential_exposures(&p, ex.data(), n, &n);
        for (auto& e : ex) {
            e.spectrum_index = static_cast<uint32_t>(rot.energies.size());
            e.bowtie_index = rt.bowtie ? static_cast<int32_t>(rot.bowties.size()) : -1;
            e.weight *= rt.mas;
            rot.exposures.push_back(e);
        }
        rot.energies.push_back(rt.tube->energies[0]);
        rot.weights.push_back(rt.tube->weights[0]);
        if (rt.bowtie)
            rot.bowties.push_back(rt.tube->bowties[0]);
    }
    const auto v = rot.view();
    double f = 0.0;
    return dxg_calibration_ctdi(&v->beam, ctdi_w, diameter, coll, run_ppe, 12345, &f, nullptr) == DXG_OK ? f : 0.0;
}
}  // namespace detail

class BeamBase {
public:
    std::uint64_t numberOfParticlesPerExposure() const { return m_ppe; }
    void setNumberOfParticlesPerExposure(std::uint64_t n) { m_ppe = n; }
    std::uint64_t numberOfExposures() const { return exposures().size(); }
    std::uint64_t numberOfParticles() const { return numberOfExposures() * m_ppe; }
    virtual std::vector<dxg_exposure> exposures() const = 0;
    virtual BeamTables tables() const
    {
        BeamTables t;
        t.exposures = exposures();
        return t;
    }
    // mGy per (dose summed over a run's histories) for the beam's stated output
    // (DAP, air KERMA, CTDI); 0 leaves the run uncalibrated.
    virtual double calibrationFactor() const { return 0.0; }
    virtual ~BeamBase() = default;

protected:
    template <typename P, typename F>
    static std::vector<dxg_exposure> collect(const P& params, F builder)
    {
        uint64_t n = 0;
        if (builder(&params, nullptr, 0, &n) != DXG_OK)
            return {};
        std::vector<dxg_exposure> ex(n);
        builder(&params, ex.data(), n, &n);
        return ex;
    }
    std::uint64_t m_ppe = 1000000;
};

class TubeBeam : public BeamBase {
public:
    explicit TubeBeam(const std::map<std::size_t, double>& filtration = {}, double kv = 120.0)
        : m_tube(kv, 12.0, filtration)
    {
    }
    const Tube& tube() const { return m_tube; }
    void setTubeVoltage(double kv) { m_tube.setVoltage(kv); }
    void setTubeAnodeAngleDeg(double a) { m_tube.setAnodeAngleDeg(a); }
    void addTubeFiltrationMaterial(std::size_t Z, double mm) { m_tube.addFiltrationMaterial(Z, mm); }
    double tubeAlHalfValueLayer() const { return m_tube.mmAlHalfValueLayer(); }
    double tubeMeanSpecterEnergy() const { return m_tube.meanSpecterEnergy(); }
    BeamTables tables() const override
    {
        BeamTables t;
        t.exposures = exposures();
        auto [e, w] = m_tube.getSpecter();
        t.energies.push_back(e);
        t.weights.push_back(w);
        return t;
    }

protected:
    Tube m_tube;
};

template <bool ENABLETRACKING = false>
class PencilBeam : public BeamBase {
public:
    PencilBeam(const std::array<double, 3>& pos = { 0, 0, 0 }, const std::array<double, 3>& dir = { 0, 0, 1 }, double energy = 60.0)
        : m_pos(pos)
        , m_dir(dir)
        , m_energy(energy)
    {
    }
    const std::array<double, 3>& position() const { return m_pos; }
    void setPosition(const std::array<double, 3>& p) { m_pos = p; }
    const std::array<double, 3>& direction() const { return m_dir; }
    void setDirection(const std::array<double, 3>& d) { m_dir = d; }
    double energy() const { return m_energy; }
    void setEnergy(double e) { m_energy = std::clamp(e, 20.0, 150.0); }   // beamsettingsmodel.cpp:664
    void setNumberOfExposures(std::uint64_t n) { m_nexp = n; }
    double airKerma() const { return m_airKerma; }
    void setAirKerma(double k) { m_airKerma = k; }
    std::vector<dxg_exposure> exposures() const override
    {
        dxg_pencil_beam p {};
        for (int i = 0; i < 3; ++i) {
            p.position[i] = m_pos[i];
            p.direction[i] = m_dir[i];
        }
        p.n_exposures = m_nexp;
        p.particles_per_exposure = m_ppe;
        return collect(p, dxg_pencil_exposures);
    }
    BeamTables tables() const override
    {
        BeamTables t;
        t.exposures = exposures();
        t.energies.push_back({ m_energy });
        t.weights.push_back({ 1.0 });
        return t;
    }
    double calibrationFactor() const override { return detail::closedFormCalibration(tables(), m_airKerma, dxg_calibration_air_kerma); }

private:
    std::array<double, 3> m_pos, m_dir;
    double m_energy;
    std::uint64_t m_nexp = 1;
    double m_airKerma = 0;   // 0: uncalibrated (the Python beams' default too)
};

template <bool ENABLETRACKING = false>
class DXBeam : public TubeBeam {
public:
    DXBeam(const std::array<double, 3>& pos = { 0, 0, 0 },
           const std::array<std::array<double, 3>, 2>& cosines = { { { 1, 0, 0 }, { 0, 1, 0 } } },
           const std::map<std::size_t, double>& filtration = {})
        : TubeBeam(filtration, 70.0)
        , m_pos(pos)
        , m_cos(cosines)
    {
    }
    const std::array<double, 3>& position() const { return m_pos; }
    void setPosition(const std::array<double, 3>& p) { m_pos = p; }
    const std::array<std::array<double, 3>, 2>& directionCosines() const { return m_cos; }
    void setDirectionCosines(const std::array<std::array<double, 3>, 2>& c) { m_cos = c; }
    const std::array<double, 2>& collimationHalfAngles() const { return m_half; }
    void setCollimationHalfAngles(const std::array<double, 2>& a) { m_half = { std::abs(a[0]), std::abs(a[1]) }; }
    void setCollimationHalfAnglesDeg(const std::array<double, 2>& a)
    {
        setCollimationHalfAngles({ a[0] * DEG_TO_RAD(), a[1] * DEG_TO_RAD() });
    }
    void setNumberOfExposures(std::uint64_t n) { m_nexp = n; }
    double DAPvalue() const { return m_dap; }
    void setDAPvalue(double d) { m_dap = d; }
    std::vector<dxg_exposure> exposures() const override
    {
        dxg_dx_beam p {};
        for (int i = 0; i < 3; ++i) {
            p.position[i] = m_pos[i];
            p.cosines[i] = m_cos[0][i];
            p.cosines[3 + i] = m_cos[1][i];
        }
        p.collimation_half_angles[0] = m_half[0];
        p.collimation_half_angles[1] = m_half[1];
        p.n_exposures = m_nexp;
        p.particles_per_exposure = m_ppe;
        return collect(p, dxg_dx_exposures);
    }

    double calibrationFactor() const override { return detail::closedFormCalibration(tables(), m_dap, dxg_calibration_dap); }

private:
    std::array<double, 3> m_pos;
    std::array<std::array<double, 3>, 2> m_cos;
    std::array<double, 2> m_half { 0.1, 0.1 };
    std::uint64_t m_nexp = 1;
    double m_dap = 0;   // 0: uncalibrated (the Python beams' default too)
};

template <bool ENABLETRACKING = false>
class CTSpiralBeam : public TubeBeam {
public:
    CTSpiralBeam(const std::array<double, 3>& start = { 0, 0, -20 }, const std::array<double, 3>& stop = { 0, 0, 20 },
                 const std::map<std::size_t, double>& filtration = {})
        : TubeBeam(filtration, 120.0)
        , m_start(start)
        , m_stop(stop)
    {
    }
    const std::array<double, 3>& startPosition() const { return m_start; }
    void setStartPosition(const std::array<double, 3>& p) { m_start = p; }
    const std::array<double, 3>& stopPosition() const { return m_stop; }
    void setStopPosition(const std::array<double, 3>& p) { m_stop = p; }
    double scanFieldOfView() const { return m_fov; }
    void setScanFieldOfView(double f) { m_fov = std::abs(f); }
    double sourceDetectorDistance() const { return m_sdd; }
    void setSourceDetectorDistance(double d) { m_sdd = std::abs(d); }
    double collimation() const { return m_coll; }
    void setCollimation(double c) { m_coll = std::abs(c); }
    double startAngleDeg() const { return m_startAngle * RAD_TO_DEG(); }
    void setStartAngleDeg(double a) { m_startAngle = a * DEG_TO_RAD(); }
    double stepAngleDeg() const { return m_stepAngle * RAD_TO_DEG(); }
    void setStepAngleDeg(double a) { m_stepAngle = std::abs(a) * DEG_TO_RAD(); }
    double pitch() const { return m_pitch; }
    void setPitch(double p) { m_pitch = std::abs(p); }
    double CTDIvol() const { return m_ctdivol; }
    void setCTDIvol(double c) { m_ctdivol = c; }
    double CTDIdiameter() const { return m_ctdiDiameter; }
    void setCTDIdiameter(double d) { m_ctdiDiameter = d; }
    const BowtieFilter& bowtieFilter() const { return m_bowtie; }
    void setBowtieFilter(const BowtieFilter& b) { m_bowtie = b; }
    const CTAECFilter& AECFilter() const { return m_aec; }
    void setAECFilter(const CTAECFilter& a) { m_aec = a; }
    CTOrganAECFilter& organAECFilter() { return m_organAec; }
    const CTOrganAECFilter& organAECFilter() const { return m_organAec; }

    std::vector<dxg_exposure> exposures() const override { return collect(params(), dxg_ct_spiral_exposures); }
    BeamTables tables() const override
    {
        BeamTables t = TubeBeam::tables();
        if (!m_bowtie.empty())
            t.bowties.push_back(m_bowtie);
        return t;
    }
    // CTDIvol of a spiral = CTDIw / pitch
    double calibrationFactor() const override
    {
        const BeamTables a = tables();
        return detail::ctdiCalibration({ { &a, m_startAngle, m_fov, 1.0, !m_bowtie.empty() } }, m_stepAngle, m_coll, m_sdd,
                                       m_ctdivol * m_pitch, m_ctdiDiameter, m_ppe);
    }

protected:
    dxg_ct_spiral_beam params() const
    {
        dxg_ct_spiral_beam p {};
        for (int i = 0; i < 3; ++i) {
            p.start[i] = m_start[i];
            p.stop[i] = m_stop[i];
        }
        p.start_angle = m_startAngle;
        p.step_angle = m_stepAngle;
        p.pitch = m_pitch;
        p.collimation_cm = m_coll;
        p.sdd_cm = m_sdd;
        p.fov_cm = m_fov;
        p.particles_per_exposure = m_ppe;
        p.has_bowtie = m_bowtie.empty() ? 0 : 1;
        p.relative_mas_a = 1;
        p.relative_mas_b = 1;
        p.fov_b_cm = m_fov;
        p.aec = m_aec.view();
        p.organ_aec = m_organAec.view();
        return p;
    }
    std::array<double, 3> m_start, m_stop;
    double m_fov = 50, m_sdd = 119, m_coll = 3.84, m_startAngle = 0, m_stepAngle = 5 * DEG_TO_RAD(), m_pitch = 1;
    double m_ctdivol = 0, m_ctdiDiameter = 32;   // CTDIvol 0: uncalibrated (the Python beams' default too)
    BowtieFilter m_bowtie;
    CTAECFilter m_aec;
    CTOrganAECFilter m_organAec;
};

template <bool ENABLETRACKING = false>
class CTSpiralDualEnergyBeam : public CTSpiralBeam<ENABLETRACKING> {
public:
    using Base = CTSpiralBeam<ENABLETRACKING>;
    CTSpiralDualEnergyBeam(const std::array<double, 3>& start = { 0, 0, -20 }, const std::array<double, 3>& stop = { 0, 0, 20 },
                           const std::map<std::size_t, double>& filtration = {})
        : Base(start, stop, filtration)
        , m_tubeB(140.0, 12.0, filtration)
    {
    }
    const Tube& tubeA() const { return this->m_tube; }
    const Tube& tubeB() const { return m_tubeB; }
    void setTubeAVoltage(double kv) { this->m_tube.setVoltage(kv); }
    void setTubeBVoltage(double kv) { m_tubeB.setVoltage(kv); }
    void setTubesAnodeAngleDeg(double a)
    {
        this->m_tube.setAnodeAngleDeg(a);
        m_tubeB.setAnodeAngleDeg(a);
    }
    void addTubeFiltrationMaterial(std::size_t Z, double mm)
    {
        this->m_tube.addFiltrationMaterial(Z, mm);
        m_tubeB.addFiltrationMaterial(Z, mm);
    }
    double scanFieldOfViewA() const { return this->m_fov; }
    void setScanFieldOfViewA(double f) { this->m_fov = std::abs(f); }
    double scanFieldOfViewB() const { return m_fovB; }
    void setScanFieldOfViewB(double f) { m_fovB = std::abs(f); }
    double tubeBoffsetAngleDeg() const { return m_offset * RAD_TO_DEG(); }
    void setTubeBoffsetAngleDeg(double a) { m_offset = a * DEG_TO_RAD(); }
    double relativeMasTubeA() const { return m_masA; }
    void setRelativeMasTubeA(double m) { m_masA = std::abs(m); }
    double relativeMasTubeB() const { return m_masB; }
    void setRelativeMasTubeB(double m) { m_masB = std::abs(m); }
    void setBowtieFilterA(const BowtieFilter& b) { this->m_bowtie = b; }
    void setBowtieFilterB(const BowtieFilter& b) { m_bowtieB = b; }

    std::vector<dxg_exposure> exposures() const override
    {
        auto p = this->params();
        p.dual_source = 1;
        p.fov_b_cm = m_fovB;
        p.tube_b_offset_angle = m_offset;
        p.relative_mas_a = m_masA;
        p.relative_mas_b = m_masB;
        p.has_bowtie = (this->m_bowtie.empty() && m_bowtieB.empty()) ? 0 : 1;
        return BeamBase::collect(p, dxg_ct_spiral_exposures);
    }
    BeamTables tables() const override
    {
        BeamTables t;
        t.exposures = exposures();
        for (const Tube* tube : { &this->m_tube, &m_tubeB }) {
            auto [e, w] = tube->getSpecter();
            t.energies.push_back(e);
            t.weights.push_back(w);
        }
        if (!this->m_bowtie.empty() || !m_bowtieB.empty()) {
            const BowtieFilter& a = this->m_bowtie.empty() ? m_bowtieB : this->m_bowtie;
            t.bowties.push_back(a);
            t.bowties.push_back(m_bowtieB.empty() ? a : m_bowtieB);
        }
        return t;
    }

    double calibrationFactor() const override
    {
        BeamTables a, b;
        auto [ea, wa] = this->m_tube.getSpecter();
        auto [eb, wb] = m_tubeB.getSpecter();
        a.energies.push_back(ea);
        a.weights.push_back(wa);
        b.energies.push_back(eb);
        b.weights.push_back(wb);
        const bool bow = !this->m_bowtie.empty() || !m_bowtieB.empty();
        if (bow) {
            const BowtieFilter& bA = this->m_bowtie.empty() ? m_bowtieB : this->m_bowtie;
            a.bowties.push_back(bA);
            b.bowties.push_back(m_bowtieB.empty() ? bA : m_bowtieB);
        }
        return detail::ctdiCalibration({ { &a, this->m_startAngle, this->m_fov, m_masA, bow },
                                         { &b, this->m_startAngle + m_offset, m_fovB, m_masB, bow } },
                                       this->m_stepAngle, this->m_coll, this->m_sdd, this->m_ctdivol * this->m_pitch,
                                       this->m_ctdiDiameter, this->m_ppe);
    }

private:
    Tube m_tubeB;
    double m_fovB = 33, m_offset = 90 * DEG_TO_RAD(), m_masA = 1, m_masB = 1;
    BowtieFilter m_bowtieB;
};

template <bool ENABLETRACKING = false>
class CTSequentialBeam : public TubeBeam {
public:
    CTSequentialBeam(const std::array<double, 3>& pos = { 0, 0, 0 }, const std::array<double, 3>& normal = { 0, 0, 1 },
                     const std::map<std::size_t, double>& filtration = {})
        : TubeBeam(filtration, 120.0)
        , m_pos(pos)
        , m_normal(normal)
    {
    }
    const std::array<double, 3>& position() const { return m_pos; }
    void setPosition(const std::array<double, 3>& p) { m_pos = p; }
    const std::array<double, 3>& scanNormal() const { return m_normal; }
    void setScanNormal(const std::array<double, 3>& n) { m_normal = n; }
    std::uint64_t numberOfSlices() const { return m_slices; }
    void setNumberOfSlices(std::uint64_t n) { m_slices = std::max<std::uint64_t>(1, n); }
    double sliceSpacing() const { return m_spacing; }
    void setSliceSpacing(double s) { m_spacing = std::abs(s); }
    double scanFieldOfView() const { return m_fov; }
    void setScanFieldOfView(double f) { m_fov = std::abs(f); }
    double sourceDetectorDistance() const { return m_sdd; }
    void setSourceDetectorDistance(double d) { m_sdd = std::abs(d); }
    double collimation() const { return m_coll; }
    void setCollimation(double c) { m_coll = std::abs(c); }
    double startAngleDeg() const { return m_startAngle * RAD_TO_DEG(); }
    void setStartAngleDeg(double a) { m_startAngle = a * DEG_TO_RAD(); }
    double stepAngleDeg() const { return m_stepAngle * RAD_TO_DEG(); }
    void setStepAngleDeg(double a) { m_stepAngle = std::abs(a) * DEG_TO_RAD(); }
    double CTDIw() const { return m_ctdiw; }
    void setCTDIw(double c) { m_ctdiw = c; }
    double CTDIdiameter() const { return m_ctdiDiameter; }
    void setCTDIdiameter(double d) { m_ctdiDiameter = d; }
    void setBowtieFilter(const BowtieFilter& b) { m_bowtie = b; }
    std::vector<dxg_exposure> exposures() const override
    {
        dxg_ct_sequential_beam p {};
        for (int i = 0; i < 3; ++i) {
            p.position[i] = m_pos[i];
            p.scan_normal[i] = m_normal[i];
        }
        p.number_of_slices = m_slices;
        p.slice_spacing_cm = m_spacing;
        p.start_angle = m_startAngle;
        p.step_angle = m_stepAngle;
        p.collimation_cm = m_coll;
        p.sdd_cm = m_sdd;
        p.fov_cm = m_fov;
        p.particles_per_exposure = m_ppe;
        p.has_bowtie = m_bowtie.empty() ? 0 : 1;
        return collect(p, dxg_ct_sequential_exposures);
    }
    BeamTables tables() const override
    {
        BeamTables t = TubeBeam::tables();
        if (!m_bowtie.empty())
            t.bowties.push_back(m_bowtie);
        return t;
    }
    double calibrationFactor() const override
    {
        const BeamTables a = tables();
        return detail::ctdiCalibration({ { &a, m_startAngle, m_fov, 1.0, !m_bowtie.empty() } }, m_stepAngle, m_coll, m_sdd,
                                       m_ctdiw, m_ctdiDiameter, m_ppe);
    }

private:
    std::array<double, 3> m_pos, m_normal;
    std::uint64_t m_slices = 1;
    double m_spacing = 1, m_fov = 50, m_sdd = 119, m_coll = 3.84, m_startAngle = 0, m_stepAngle = 5 * DEG_TO_RAD();
    double m_ctdiw = 0, m_ctdiDiameter = 32;   // CTDIw 0: uncalibrated (the Python beams' default too)
    BowtieFilter m_bowtie;
};

template <bool ENABLETRACKING = false>
class CBCTBeam : public TubeBeam {
public:
    CBCTBeam(const std::array<double, 3>& isocenter = { 0, 0, 0 }, const std::array<double, 3>& axis = { 0, 0, 1 },
             const std::map<std::size_t, double>& filtration = {})
        : TubeBeam(filtration, 90.0)
        , m_iso(isocenter)
        , m_axis(axis)
    {
    }
    const std::array<double, 3>& isocenter() const { return m_iso; }
    void setIsocenter(const std::array<double, 3>& c) { m_iso = c; }
    const std::array<double, 3>& rotationAxis() const { return m_axis; }
    void setRotationAxis(const std::array<double, 3>& a) { m_axis = a; }
    double sourceDetectorDistance() const { return m_sdd; }
    void setSourceDetectorDistance(double d) { m_sdd = std::abs(d); }
    double startAngleDeg() const { return m_start * RAD_TO_DEG(); }
    void setStartAngleDeg(double a) { m_start = a * DEG_TO_RAD(); }
    double stopAngleDeg() const { return m_stop * RAD_TO_DEG(); }
    void setStopAngleDeg(double a) { m_stop = a * DEG_TO_RAD(); }
    double stepAngleDeg() const { return m_step * RAD_TO_DEG(); }
    void setStepAngleDeg(double a) { m_step = std::abs(a) * DEG_TO_RAD(); }
    void setCollimationHalfAngles(double x, double y) { m_half = { std::abs(x), std::abs(y) }; }
    void setCollimationHalfAnglesDeg(double x, double y) { setCollimationHalfAngles(x * DEG_TO_RAD(), y * DEG_TO_RAD()); }
    const std::array<double, 2>& collimationHalfAngles() const { return m_half; }
    double DAPvalue() const { return m_dap; }
    void setDAPvalue(double d) { m_dap = d; }
    std::vector<dxg_exposure> exposures() const override
    {
        dxg_cbct_beam p {};
        for (int i = 0; i < 3; ++i) {
            p.isocenter[i] = m_iso[i];
            p.rotation_axis[i] = m_axis[i];
        }
        p.start_angle = m_start;
        p.stop_angle = m_stop;
        p.step_angle = m_step;
        p.sdd_cm = m_sdd;
        p.collimation_half_angles[0] = m_half[0];
        p.collimation_half_angles[1] = m_half[1];
        p.particles_per_exposure = m_ppe;
        return collect(p, dxg_cbct_exposures);
    }

    double calibrationFactor() const override { return detail::closedFormCalibration(tables(), m_dap, dxg_calibration_dap); }

private:
    std::array<double, 3> m_iso, m_axis;
    double m_sdd = 120, m_start = -90 * DEG_TO_RAD(), m_stop = 90 * DEG_TO_RAD(), m_step = 1 * DEG_TO_RAD();
    std::array<double, 2> m_half { 10 * DEG_TO_RAD(), 10 * DEG_TO_RAD() };
    double m_dap = 0;   // 0: uncalibrated (the Python beams' default too)
};

// --------------------------------------------------------------- transport
class Transport {
public:
    Transport()
        : m_seed(std::random_device {}())
    {
    }
    void setNumberOfThreads(std::size_t n) { m_threads = n; }   // host threads do not drive the device path
    std::size_t numberOfThreads() const { return m_threads; }
    void setRandomSeed(std::uint64_t seed) { m_seed = seed; }

    // Transport::operator()(world, beam, progress, useBeamCalibration)
    // (simulationpipeline.cpp:163-166).  Blocking.  Returns false on failure or
    // cancellation; on failure the progress stop flag is also set.
    template <typename W, typename B>
    bool operator()(W& world, const B& beam, TransportProgress* progress = nullptr, bool useBeamCalibration = true)
    {
        const BeamTables t = beam.tables();
        const auto v = t.view(useBeamCalibration ? beam.calibrationFactor() : 0.0);
        const int rc = dxg_transport_run(world.gridHandle(), &v->beam, progress ? progress->handle() : nullptr,
                                         useBeamCalibration ? 1 : 0, m_seed++);
        world.scoresChanged();
        if (rc != DXG_OK && rc != DXG_ERR_CANCELLED && progress)
            progress->setStopSimulation();
        return rc == DXG_OK;
    }

private:
    std::uint64_t m_seed;
    std::size_t m_threads = 0;
};

}  // namespace dxmc
