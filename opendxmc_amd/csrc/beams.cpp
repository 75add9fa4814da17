// Host-side history generation: exposure lists of the DXMClib beam families
// and the X-ray tube spectrum model (include/dxmc_beams.h).
#include "dxmc_beams.h"

#include "physics.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

namespace {

constexpr double kPi = 3.14159265358979323846;

int fail(int code, const std::string& msg) { return dxg::setError(code, msg); }

struct V3 {
    double x, y, z;
};
V3 v3(const double* a) { return { a[0], a[1], a[2] }; }
V3 operator+(V3 a, V3 b) { return { a.x + b.x, a.y + b.y, a.z + b.z }; }
V3 operator-(V3 a, V3 b) { return { a.x - b.x, a.y - b.y, a.z - b.z }; }
V3 operator*(double s, V3 a) { return { s * a.x, s * a.y, s * a.z }; }
double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
V3 cross(V3 a, V3 b) { return { a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x }; }
double norm(V3 a) { return std::sqrt(dot(a, a)); }
V3 unit(V3 a) { return (1.0 / norm(a)) * a; }

// Orthonormal gantry frame (e1, e2) about rotation axis n.
void gantryFrame(V3 n, V3& e1, V3& e2)
{
    V3 ref = { 0, 1, 0 };
    if (std::fabs(dot(n, ref)) > 0.999)
        ref = { 1, 0, 0 };
    e1 = unit(cross(ref, n));
    e2 = cross(n, e1);
}

void setExposure(dxg_exposure& e, V3 pos, V3 c0, V3 c1, double hx, double hy, double weight, uint32_t spec, int32_t bow, uint64_t np)
{
    std::memset(&e, 0, sizeof e);
    e.position[0] = pos.x;
    e.position[1] = pos.y;
    e.position[2] = pos.z;
    e.cosines[0] = c0.x;
    e.cosines[1] = c0.y;
    e.cosines[2] = c0.z;
    e.cosines[3] = c1.x;
    e.cosines[4] = c1.y;
    e.cosines[5] = c1.z;
    e.collimation_half_angles[0] = hx;
    e.collimation_half_angles[1] = hy;
    e.weight = weight;
    e.spectrum_index = spec;
    e.bowtie_index = bow;
    e.n_particles = np;
}

// Exposure at gantry angle a about axis n through centre c.
void gantryExposure(dxg_exposure& e, V3 c, V3 n, V3 e1, V3 e2, double a, double R, double hx, double hy, double weight,
                    uint32_t spec, int32_t bow, uint64_t np)
{
    const V3 b = std::cos(a) * e1 + std::sin(a) * e2;
    setExposure(e, c - R * b, cross(n, b), n, hx, hy, weight, spec, bow, np);
}

struct Aec {
    bool on = false;
    V3 s {}, d {};
    double len2 = 0;
    std::vector<double> w;
    bool init(const dxg_aec& a, std::string& err)
    {
        if (a.n == 0)
            return true;
        if (!a.weights) {
            err = "AEC weights missing";
            return false;
        }
        s = v3(a.start);
        d = v3(a.stop) - s;
        len2 = dot(d, d);
        double sum = 0;
        for (uint64_t i = 0; i < a.n; ++i) {
            if (!(a.weights[i] >= 0) || !std::isfinite(a.weights[i])) {
                err = "AEC weights must be finite and non-negative";
                return false;
            }
            w.push_back(a.weights[i]);
            sum += a.weights[i];
        }
        if (!(sum > 0)) {
            err = "AEC weights sum to zero";
            return false;
        }
        const double mean = sum / static_cast<double>(w.size());
        for (auto& v : w)
            v /= mean;
        on = true;
        return true;
    }
    double at(V3 p) const
    {
        if (!on)
            return 1.0;
        if (w.size() == 1 || len2 <= 0)
            return w[0];
        const double t = std::clamp(dot(p - s, d) / len2, 0.0, 1.0) * static_cast<double>(w.size() - 1);
        const size_t k = std::min(static_cast<size_t>(t), w.size() - 2);
        const double f = t - static_cast<double>(k);
        return w[k] + f * (w[k + 1] - w[k]);
    }
};

// Organ AEC: weight by gantry angle (see dxg_organ_aec in dxmc_beams.h).
struct OrganAec {
    bool on = false;
    double start = 0, arc = 0, ramp = 0, low = 1, high = 1;
    bool init(const dxg_organ_aec& f, std::string& err)
    {
        if (!f.use)
            return true;
        if (!(f.low_weight >= 0) || !std::isfinite(f.low_weight) || !std::isfinite(f.start_angle) || !std::isfinite(f.stop_angle)
            || !(f.ramp_angle >= 0) || !std::isfinite(f.ramp_angle)) {
            err = "organ AEC: angles must be finite, ramp and low weight non-negative";
            return false;
        }
        constexpr double two_pi = 2.0 * kPi;
        on = true;
        start = f.start_angle;
        arc = std::fmod(f.stop_angle - f.start_angle, two_pi);
        if (arc < 0)
            arc += two_pi;
        ramp = std::min(f.ramp_angle, 0.5 * (two_pi - arc));
        low = f.low_weight;
        high = 1.0;
        const double rest = two_pi - arc - ramp;   // the turn's weight budget outside the low arc and half the ramps
        if (f.compensate_outside && rest > 1e-12)
            high = (two_pi - low * (arc + ramp)) / rest;
        return true;
    }
    double at(double a) const
    {
        if (!on)
            return 1.0;
        constexpr double two_pi = 2.0 * kPi;
        double x = std::fmod(a - start, two_pi);
        if (x < 0)
            x += two_pi;
        if (x <= arc)
            return low;
        if (ramp > 0 && x < arc + ramp)
            return low + (high - low) * (x - arc) / ramp;
        if (ramp > 0 && x > two_pi - ramp)
            return low + (high - low) * (two_pi - x) / ramp;
        return high;
    }
};

int finish(uint64_t count, uint64_t cap, dxg_exposure* out, uint64_t* n)
{
    if (n)
        *n = count;
    if (out && cap < count)
        return fail(DXG_ERR_ARGUMENT, "exposure buffer too small: need " + std::to_string(count));
    return DXG_OK;
}

// ------------------------------------------------------------------ tube model
struct FilterDensity {
    int z;
    double rho;
};
constexpr FilterDensity kFilterDensity[] = {
    { 4, 1.848 }, { 6, 2.26 }, { 13, 2.699 }, { 22, 4.54 }, { 26, 7.874 }, { 29, 8.96 }, { 42, 10.22 },
    { 45, 12.41 }, { 47, 10.50 }, { 50, 7.31 }, { 64, 7.90 }, { 73, 16.65 }, { 74, 19.30 }, { 79, 19.32 }, { 82, 11.35 },
};

double filterDensity(int z)
{
    for (const auto& f : kFilterDensity)
        if (f.z == z)
            return f.rho;
    return -1;
}

// Kanaya-Okayama electron range in tungsten (cm) at energy e keV.
double rangeTungsten(double e)
{
    const double A = 183.84, Z = 74.0, rho = 19.3;
    return 0.0276 * A * std::pow(e, 1.67) / (std::pow(Z, 0.889) * rho) * 1e-4;
}

double elementMu(const dxg::PhysicsList& pl, int z, double e)   // cm^2/g
{
    const double apg = dxg::kNA / pl.element(z).A * 1e-24;
    return apg * (pl.photo(z, e) + pl.incoherent(z, e) + pl.coherent(z, e));
}

struct TubeSpectrum {
    std::vector<double> e, w;
};

constexpr double kWKEdge = 69.525;
// Tungsten K lines (keV) and relative intensities Ka2 : Ka1 : Kb1 : Kb2.
constexpr double kWLineE[4] = { 57.982, 59.318, 67.244, 69.067 };
constexpr double kWLineI[4] = { 0.58, 1.0, 0.22, 0.08 };

int tubeSpectrum(const dxg_tube* t, TubeSpectrum& out)
{
    if (!t)
        return fail(DXG_ERR_ARGUMENT, "tube: null argument");
    const double kvp = t->voltage_kv;
    if (!(kvp >= 10.0 && kvp <= 200.0))
        return fail(DXG_ERR_ARGUMENT, "tube voltage must be within [10, 200] kV");
    const double ang = t->anode_angle_deg;
    if (!(ang > 0.5 && ang < 90.0))
        return fail(DXG_ERR_ARGUMENT, "anode angle must be within (0.5, 90) degrees");
    const double de = t->energy_step_kev > 0 ? t->energy_step_kev : 0.5;
    if (t->n_filtration && (!t->filtration_Z || !t->filtration_mm))
        return fail(DXG_ERR_ARGUMENT, "tube filtration arrays missing");
    std::string err;
    const dxg::PhysicsList* pl = dxg::sharedPhysics(err);
    if (!pl)
        return fail(DXG_ERR_PHYSICS, err);
    std::vector<std::pair<int, double>> filt;   // (Z, rho*t g/cm2)
    for (uint64_t i = 0; i < t->n_filtration; ++i) {
        const int z = t->filtration_Z[i];
        const double mm = t->filtration_mm[i];
        if (!(mm >= 0) || !std::isfinite(mm))
            return fail(DXG_ERR_ARGUMENT, "filtration thickness must be non-negative");
        if (mm == 0)
            continue;
        const double rho = filterDensity(z);
        if (rho <= 0 || z > pl->zmax())
            return fail(DXG_ERR_ARGUMENT, "no density for filtration element Z=" + std::to_string(z));
        filt.emplace_back(z, rho * mm * 0.1);
    }
    const double sin_a = std::sin(ang * kPi / 180.0);
    const double r_kvp = rangeTungsten(kvp);
    auto transmission = [&](double e) {
        const double tw = 0.5 * (r_kvp - rangeTungsten(e)) / sin_a;
        double att = elementMu(*pl, 74, e) * 19.3 * std::max(tw, 0.0);
        for (const auto& f : filt)
            att += elementMu(*pl, f.first, e) * f.second;
        return std::exp(-att);
    };
    std::vector<std::pair<double, double>> lines;
    double brems_total = 0;
    for (int k = 0;; ++k) {
        const double e = (k + 0.5) * de;
        if (e >= kvp)
            break;
        if (e < dxg::kCutoffKeV)
            continue;
        const double n = (kvp - e) / e * de;
        brems_total += n;
        lines.emplace_back(e, n * transmission(e));
    }
    if (kvp > kWKEdge) {
        const double ktot = 0.1 * std::pow(kvp / kWKEdge - 1.0, 1.63) * brems_total;
        double isum = 0;
        for (double v : kWLineI)
            isum += v;
        for (int i = 0; i < 4; ++i)
            lines.emplace_back(kWLineE[i], ktot * kWLineI[i] / isum * transmission(kWLineE[i]));
    }
    std::sort(lines.begin(), lines.end());
    double sum = 0;
    for (auto& l : lines)
        sum += l.second;
    if (!(sum > 0))
        return fail(DXG_ERR_ARGUMENT, "tube filtration absorbs the whole spectrum");
    out.e.clear();
    out.w.clear();
    for (auto& l : lines) {
        out.e.push_back(l.first);
        out.w.push_back(l.second / sum);
    }
    return DXG_OK;
}

// Mean fraction of the photon energy given to the electron in a free-electron
// (Klein-Nishina) scatter, by midpoint integration over cos(theta).
double knTransferFraction(double e)
{
    const double k = e / dxg::kMeC2KeV;
    constexpr int N = 512;
    double s = 0, st = 0;
    for (int i = 0; i < N; ++i) {
        const double mu = -1.0 + (i + 0.5) * 2.0 / N;
        const double p = 1.0 / (1.0 + k * (1.0 - mu));
        const double dcs = p * p * (p + 1.0 / p - (1.0 - mu * mu));
        s += dcs;
        st += dcs * (1.0 - p);
    }
    return st / s;
}

}  // namespace

extern "C" {

int dxg_tube_spectrum(const dxg_tube* tube, double* energy_kev, double* weight, uint64_t cap, uint64_t* n)
{
    TubeSpectrum s;
    int rc = tubeSpectrum(tube, s);
    if (rc != DXG_OK)
        return rc;
    if (n)
        *n = s.e.size();
    if (!energy_kev && !weight)
        return DXG_OK;
    if (cap < s.e.size())
        return fail(DXG_ERR_ARGUMENT, "spectrum buffer too small: need " + std::to_string(s.e.size()));
    for (size_t i = 0; i < s.e.size(); ++i) {
        if (energy_kev)
            energy_kev[i] = s.e[i];
        if (weight)
            weight[i] = s.w[i];
    }
    return DXG_OK;
}

int dxg_tube_mean_energy(const dxg_tube* tube, double* mean_kev)
{
    TubeSpectrum s;
    int rc = tubeSpectrum(tube, s);
    if (rc != DXG_OK)
        return rc;
    if (!mean_kev)
        return fail(DXG_ERR_ARGUMENT, "null output");
    double m = 0;
    for (size_t i = 0; i < s.e.size(); ++i)
        m += s.e[i] * s.w[i];
    *mean_kev = m;
    return DXG_OK;
}

int dxg_tube_hvl_mm_al(const dxg_tube* tube, double* hvl_mm)
{
    TubeSpectrum s;
    int rc = tubeSpectrum(tube, s);
    if (rc != DXG_OK)
        return rc;
    if (!hvl_mm)
        return fail(DXG_ERR_ARGUMENT, "null output");
    std::string err;
    const dxg::PhysicsList* pl = dxg::sharedPhysics(err);
    if (!pl)
        return fail(DXG_ERR_PHYSICS, err);
    // dry air by mass: C, N, O, Ar
    const int32_t airZ[4] = { 6, 7, 8, 18 };
    const double airW[4] = { 0.000124, 0.755268, 0.231781, 0.012827 };
    dxg::MixMaterial air;
    if (!dxg::makeMaterial(*pl, airZ, airW, 4, air, err))
        return fail(DXG_ERR_MATERIAL, err);
    const size_t n = s.e.size();
    std::vector<double> kerma(n), mu_al(n);
    for (size_t i = 0; i < n; ++i) {
        double pe, inc, coh;
        air.attenuation(*pl, s.e[i], pe, inc, coh);
        kerma[i] = s.w[i] * s.e[i] * (pe + inc * knTransferFraction(s.e[i]));
        mu_al[i] = elementMu(*pl, 13, s.e[i]) * 2.699 * 0.1;   // per mm
    }
    auto k_of = [&](double t) {
        double v = 0;
        for (size_t i = 0; i < n; ++i)
            v += kerma[i] * std::exp(-mu_al[i] * t);
        return v;
    };
    const double half = 0.5 * k_of(0.0);
    double lo = 0, hi = 1.0;
    while (k_of(hi) > half && hi < 1e3)
        hi *= 2.0;
    for (int it = 0; it < 100; ++it) {
        const double mid = 0.5 * (lo + hi);
        (k_of(mid) > half ? lo : hi) = mid;
    }
    *hvl_mm = 0.5 * (lo + hi);
    return DXG_OK;
}

int dxg_pencil_exposures(const dxg_pencil_beam* b, dxg_exposure* out, uint64_t cap, uint64_t* n)
{
    if (!b)
        return fail(DXG_ERR_ARGUMENT, "pencil beam: null argument");
    const V3 d = v3(b->direction);
    if (!(norm(d) > 0))
        return fail(DXG_ERR_ARGUMENT, "pencil beam direction is zero");
    const V3 dir = unit(d);
    const uint64_t count = b->n_exposures;
    int rc = finish(count, cap, out, n);
    if (rc != DXG_OK || !out)
        return rc;
    V3 ref = { 0, 0, 1 };
    if (std::fabs(dot(dir, ref)) > 0.9)
        ref = { 1, 0, 0 };
    const V3 c1 = unit(cross(dir, ref));
    const V3 c0 = cross(c1, dir);
    for (uint64_t i = 0; i < count; ++i)
        setExposure(out[i], v3(b->position), c0, c1, 0.0, 0.0, 1.0, 0, -1, b->particles_per_exposure);
    return DXG_OK;
}

int dxg_dx_exposures(const dxg_dx_beam* b, dxg_exposure* out, uint64_t cap, uint64_t* n)
{
    if (!b)
        return fail(DXG_ERR_ARGUMENT, "DX beam: null argument");
    const V3 c0 = v3(b->cosines), c1 = v3(b->cosines + 3);
    if (!(norm(cross(c0, c1)) > 1e-12))
        return fail(DXG_ERR_ARGUMENT, "DX beam direction cosines are parallel");
    const uint64_t count = b->n_exposures;
    int rc = finish(count, cap, out, n);
    if (rc != DXG_OK || !out)
        return rc;
    for (uint64_t i = 0; i < count; ++i)
        setExposure(out[i], v3(b->position), unit(c0), unit(c1), std::fabs(b->collimation_half_angles[0]),
                    std::fabs(b->collimation_half_angles[1]), 1.0, 0, -1, b->particles_per_exposure);
    return DXG_OK;
}

int dxg_organ_aec_weight(const dxg_organ_aec* f, double angle, double* weight)
{
    if (!f || !weight)
        return fail(DXG_ERR_ARGUMENT, "organ AEC: null argument");
    std::string err;
    OrganAec o;
    if (!o.init(*f, err))
        return fail(DXG_ERR_ARGUMENT, err);
    *weight = o.at(angle);
    return DXG_OK;
}

int dxg_organ_aec_max_weight(const dxg_organ_aec* f, double* weight)
{
    if (!f || !weight)
        return fail(DXG_ERR_ARGUMENT, "organ AEC: null argument");
    std::string err;
    OrganAec o;
    if (!o.init(*f, err))
        return fail(DXG_ERR_ARGUMENT, err);
    *weight = o.on ? o.high : 1.0;
    return DXG_OK;
}

int dxg_aec_from_wed(const uint64_t dims[3], const double spacing[3], const double* image, int image_is_ct, double* weights,
                     double start[3], double stop[3])
{
    if (!dims || !spacing || !image || !weights || !start || !stop)
        return fail(DXG_ERR_ARGUMENT, "AEC from WED: null argument");
    if (dims[0] == 0 || dims[1] == 0 || dims[2] == 0 || !(spacing[0] > 0) || !(spacing[1] > 0) || !(spacing[2] > 0))
        return fail(DXG_ERR_ARGUMENT, "AEC from WED: empty image or non-positive spacing");
    constexpr double kMuWater = 0.2;   // 1/cm, about that of water at 60-70 keV
    const uint64_t step = dims[0] * dims[1];
    const double pixel_cm2 = spacing[0] * spacing[1];
    for (uint64_t z = 0; z < dims[2]; ++z) {
        const double* slice = image + z * step;
        double sum = 0;
        for (uint64_t i = 0; i < step; ++i)
            sum += slice[i];
        const double area = image_is_ct ? (sum / static_cast<double>(step) / 1000.0 + 1.0) * static_cast<double>(step) * pixel_cm2
                                        : sum * pixel_cm2;
        const double diameter = 2.0 * std::sqrt(std::max(area, 0.0) / kPi);
        weights[z] = std::exp(kMuWater * diameter);
    }
    const double half = 0.5 * spacing[2] * static_cast<double>(dims[2]);
    start[0] = start[1] = stop[0] = stop[1] = 0.0;
    start[2] = -half;
    stop[2] = half;
    return DXG_OK;
}

int dxg_ct_spiral_exposures(const dxg_ct_spiral_beam* b, dxg_exposure* out, uint64_t cap, uint64_t* n)
{
    if (!b)
        return fail(DXG_ERR_ARGUMENT, "CT spiral beam: null argument");
    if (!(b->step_angle > 0) || !(b->pitch > 0) || !(b->collimation_cm > 0) || !(b->sdd_cm > 0) || !(b->fov_cm > 0))
        return fail(DXG_ERR_ARGUMENT, "CT spiral beam: step angle, pitch, collimation, SDD and FOV must be positive");
    if (b->fov_cm >= b->sdd_cm || (b->dual_source && !(b->fov_b_cm > 0 && b->fov_b_cm < b->sdd_cm)))
        return fail(DXG_ERR_ARGUMENT, "CT spiral beam: field of view must be below SDD");
    std::string err;
    Aec aec;
    if (!aec.init(b->aec, err))
        return fail(DXG_ERR_ARGUMENT, err);
    OrganAec organ;
    if (!organ.init(b->organ_aec, err))
        return fail(DXG_ERR_ARGUMENT, err);
    const V3 s = v3(b->start), e = v3(b->stop);
    const double len = norm(e - s);
    const V3 axis = len > 0 ? unit(e - s) : V3 { 0, 0, 1 };
    const double advance = b->pitch * b->collimation_cm;   // per turn
    const uint64_t turns_steps = static_cast<uint64_t>(std::floor(len / advance * 2.0 * kPi / b->step_angle + 1e-9)) + 1;
    const uint64_t per = b->dual_source ? 2 : 1;
    const uint64_t count = turns_steps * per;
    int rc = finish(count, cap, out, n);
    if (rc != DXG_OK || !out)
        return rc;
    V3 e1, e2;
    gantryFrame(axis, e1, e2);
    const double R = 0.5 * b->sdd_cm;
    const double hy = std::atan(b->collimation_cm / b->sdd_cm);
    const double hxa = std::asin(b->fov_cm / b->sdd_cm);
    const double hxb = b->dual_source ? std::asin(b->fov_b_cm / b->sdd_cm) : hxa;
    for (uint64_t i = 0; i < turns_steps; ++i) {
        const double a = b->start_angle + static_cast<double>(i) * b->step_angle;
        const V3 c = s + (static_cast<double>(i) * b->step_angle / (2.0 * kPi) * advance) * axis;
        const double w = aec.at(c);
        if (!b->dual_source) {
            gantryExposure(out[i], c, axis, e1, e2, a, R, hxa, hy, w * organ.at(a), 0, b->has_bowtie ? 0 : -1, b->particles_per_exposure);
        } else {
            const double ab = a + b->tube_b_offset_angle;
            gantryExposure(out[2 * i], c, axis, e1, e2, a, R, hxa, hy, w * b->relative_mas_a * organ.at(a), 0, b->has_bowtie ? 0 : -1,
                           b->particles_per_exposure);
            gantryExposure(out[2 * i + 1], c, axis, e1, e2, ab, R, hxb, hy, w * b->relative_mas_b * organ.at(ab), 1,
                           b->has_bowtie ? 1 : -1, b->particles_per_exposure);
        }
    }
    return DXG_OK;
}

int dxg_ct_sequential_exposures(const dxg_ct_sequential_beam* b, dxg_exposure* out, uint64_t cap, uint64_t* n)
{
    if (!b)
        return fail(DXG_ERR_ARGUMENT, "CT sequential beam: null argument");
    if (!(b->step_angle > 0) || !(b->collimation_cm > 0) || !(b->sdd_cm > 0) || !(b->fov_cm > 0) || b->fov_cm >= b->sdd_cm)
        return fail(DXG_ERR_ARGUMENT, "CT sequential beam: bad step angle, collimation, SDD or FOV");
    const V3 nrm = v3(b->scan_normal);
    if (!(norm(nrm) > 0))
        return fail(DXG_ERR_ARGUMENT, "CT sequential beam: scan normal is zero");
    const V3 axis = unit(nrm);
    const uint64_t per_slice = std::max<uint64_t>(1, static_cast<uint64_t>(std::ceil(2.0 * kPi / b->step_angle - 1e-9)));
    const uint64_t count = per_slice * std::max<uint64_t>(b->number_of_slices, 0);
    int rc = finish(count, cap, out, n);
    if (rc != DXG_OK || !out)
        return rc;
    V3 e1, e2;
    gantryFrame(axis, e1, e2);
    const double R = 0.5 * b->sdd_cm;
    const double hy = std::atan(b->collimation_cm / b->sdd_cm);
    const double hx = std::asin(b->fov_cm / b->sdd_cm);
    uint64_t k = 0;
    for (uint64_t sl = 0; sl < b->number_of_slices; ++sl) {
        const V3 c = v3(b->position) + (static_cast<double>(sl) * b->slice_spacing_cm) * axis;
        for (uint64_t j = 0; j < per_slice; ++j)
            gantryExposure(out[k++], c, axis, e1, e2, b->start_angle + static_cast<double>(j) * b->step_angle, R, hx, hy, 1.0, 0,
                           b->has_bowtie ? 0 : -1, b->particles_per_exposure);
    }
    return DXG_OK;
}

int dxg_cbct_exposures(const dxg_cbct_beam* b, dxg_exposure* out, uint64_t cap, uint64_t* n)
{
    if (!b)
        return fail(DXG_ERR_ARGUMENT, "CBCT beam: null argument");
    if (!(b->step_angle > 0) || !(b->sdd_cm > 0) || b->stop_angle < b->start_angle)
        return fail(DXG_ERR_ARGUMENT, "CBCT beam: bad angles or SDD");
    const V3 ra = v3(b->rotation_axis);
    if (!(norm(ra) > 0))
        return fail(DXG_ERR_ARGUMENT, "CBCT beam: rotation axis is zero");
    const V3 axis = unit(ra);
    const uint64_t count = static_cast<uint64_t>(std::floor((b->stop_angle - b->start_angle) / b->step_angle + 1e-9)) + 1;
    int rc = finish(count, cap, out, n);
    if (rc != DXG_OK || !out)
        return rc;
    V3 e1, e2;
    gantryFrame(axis, e1, e2);
    for (uint64_t i = 0; i < count; ++i)
        gantryExposure(out[i], v3(b->isocenter), axis, e1, e2, b->start_angle + static_cast<double>(i) * b->step_angle,
                       0.5 * b->sdd_cm, std::fabs(b->collimation_half_angles[0]), std::fabs(b->collimation_half_angles[1]), 1.0, 0,
                       -1, b->particles_per_exposure);
    return DXG_OK;
}

}  // extern "C"

namespace {

// Air-KERMA-area product (mGy cm2) of the simulated photons of a beam: sum over
// exposures of n x weight x <E mu_tr/rho(air)(E)> of the exposure's spectrum.
int simulatedKermaArea(const dxg_beam* beam, double& kap, const char* fn)
{
    if (!beam || !beam->exposures || beam->n_exposures == 0 || !beam->spectra || beam->n_spectra == 0)
        return fail(DXG_ERR_ARGUMENT, std::string(fn) + ": beam needs exposures and spectra");
    std::string err;
    const dxg::PhysicsList* pl = dxg::sharedPhysics(err);
    if (!pl)
        return fail(DXG_ERR_PHYSICS, err);
    const int32_t airZ[4] = { 6, 7, 8, 18 };
    const double airW[4] = { 0.000124, 0.755268, 0.231781, 0.012827 };
    dxg::MixMaterial air;
    if (!dxg::makeMaterial(*pl, airZ, airW, 4, air, err))
        return fail(DXG_ERR_MATERIAL, err);
    std::vector<double> mean(beam->n_spectra, 0.0);
    for (uint64_t s = 0; s < beam->n_spectra; ++s) {
        const dxg_spectrum& sp = beam->spectra[s];
        double wsum = 0.0, acc = 0.0;
        for (uint64_t k = 0; k < sp.n; ++k) {
            const double e = sp.energy_kev[k];
            if (!(e > 0.0) || !(sp.weight[k] >= 0.0))
                return fail(DXG_ERR_ARGUMENT, std::string(fn) + ": invalid spectrum");
            double pe, inc, coh;
            air.attenuation(*pl, e, pe, inc, coh);
            acc += sp.weight[k] * e * (pe + inc * knTransferFraction(e));
            wsum += sp.weight[k];
        }
        if (!(wsum > 0.0))
            return fail(DXG_ERR_ARGUMENT, std::string(fn) + ": spectrum weights sum to zero");
        mean[s] = acc / wsum;
    }
    kap = 0.0;
    for (uint64_t i = 0; i < beam->n_exposures; ++i) {
        const dxg_exposure& e = beam->exposures[i];
        if (e.spectrum_index >= beam->n_spectra)
            return fail(DXG_ERR_ARGUMENT, std::string(fn) + ": exposure references a missing spectrum");
        kap += static_cast<double>(e.n_particles) * e.weight * mean[e.spectrum_index];
    }
    kap *= dxg::kKeVToJ * 1e3 * 1e3;   // keV cm2/g -> mGy cm2
    if (!(kap > 0.0))
        return fail(DXG_ERR_ARGUMENT, std::string(fn) + ": beam has no particles");
    return DXG_OK;
}

}  // namespace

extern "C" {

int dxg_calibration_dap(const dxg_beam* beam, double dap_mgycm2, double* factor)
{
    if (!factor)
        return fail(DXG_ERR_ARGUMENT, "dxg_calibration_dap: null output");
    if (!(dap_mgycm2 > 0.0))
        return fail(DXG_ERR_ARGUMENT, "dxg_calibration_dap: the dose-area product must be positive");
    double kap = 0.0;
    if (int rc = simulatedKermaArea(beam, kap, "dxg_calibration_dap"); rc != DXG_OK)
        return rc;
    *factor = dap_mgycm2 / kap;
    return DXG_OK;
}

int dxg_calibration_air_kerma(const dxg_beam* beam, double air_kerma_mgy, double* factor)
{
    if (!factor)
        return fail(DXG_ERR_ARGUMENT, "dxg_calibration_air_kerma: null output");
    if (!(air_kerma_mgy > 0.0))
        return fail(DXG_ERR_ARGUMENT, "dxg_calibration_air_kerma: the air KERMA must be positive");
    double kap = 0.0;
    if (int rc = simulatedKermaArea(beam, kap, "dxg_calibration_air_kerma"); rc != DXG_OK)
        return rc;
    *factor = air_kerma_mgy * 1.0 / kap;   // KERMA over a 1 cm2 cross-section
    return DXG_OK;
}

}  // extern "C"

