// Physics list reader, mixture rule and device table builder.
#include "physics.hpp"

#include <dlfcn.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <memory>
#include <mutex>

namespace dxg {

namespace {

template <typename T>
bool readVec(std::ifstream& f, std::vector<T>& v, size_t n)
{
    v.resize(n);
    f.read(reinterpret_cast<char*>(v.data()), static_cast<std::streamsize>(n * sizeof(T)));
    return static_cast<size_t>(f.gcount()) == n * sizeof(T);
}

template <typename T>
bool readOne(std::ifstream& f, T& v)
{
    f.read(reinterpret_cast<char*>(&v), sizeof(T));
    return f.gcount() == sizeof(T);
}

// Interpolate log(v) linearly in log(e); `le` ascending, duplicates mark edges.
double interpLogLog(const std::vector<double>& le, const std::vector<double>& lv, double logE)
{
    if (logE <= le.front())
        return std::exp(lv.front());
    if (logE >= le.back())
        return std::exp(lv.back());
    const auto it = std::upper_bound(le.begin(), le.end(), logE);
    const size_t b = static_cast<size_t>(it - le.begin());
    const size_t a = b - 1;
    const double t = (logE - le[a]) / (le[b] - le[a]);
    return std::exp(lv[a] + t * (lv[b] - lv[a]));
}

double interpLinear(const std::vector<double>& xs, const std::vector<double>& v, double x)
{
    if (x <= xs.front())
        return v.front();
    if (x >= xs.back())
        return v.back();
    const auto it = std::upper_bound(xs.begin(), xs.end(), x);
    const size_t b = static_cast<size_t>(it - xs.begin());
    const size_t a = b - 1;
    return v[a] + (x - xs[a]) / (xs[b] - xs[a]) * (v[b] - v[a]);
}

double safeLog(double v) { return std::log(std::max(v, 1e-300)); }

float roundUp(double v)
{
    float f = static_cast<float>(v);
    if (static_cast<double>(f) < v)
        f = std::nextafter(f, 3.0e38f);
    return f;
}

}  // namespace

bool PhysicsList::load(const std::string& path, std::string& err)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) {
        err = "cannot open physics list '" + path + "'";
        return false;
    }
    char magic[8] = {};
    f.read(magic, 8);
    uint32_t zmax = 0, ne = 0, nx = 0;
    if (std::memcmp(magic, "DXGPHYS2", 8) != 0 || !readOne(f, zmax) || !readOne(f, ne) || !readOne(f, nx)) {
        err = "physics list '" + path + "' has a bad header";
        return false;
    }
    std::vector<double> e;
    if (!readVec(f, e, ne) || !readVec(f, x_, nx)) {
        err = "physics list truncated";
        return false;
    }
    log_e_.resize(ne);
    for (uint32_t i = 0; i < ne; ++i)
        log_e_[i] = std::log(e[i]);
    std::vector<ElementData> els(zmax);
    for (auto& el : els) {
        std::vector<double> inc, coh, pe_e, pe, shells, fl;
        uint32_t npe = 0, nsh = 0;
        bool ok = readOne(f, el.A) && readVec(f, inc, ne) && readVec(f, coh, ne) && readVec(f, el.F, nx)
            && readVec(f, el.S, nx) && readOne(f, npe) && readVec(f, pe_e, npe) && readVec(f, pe, npe)
            && readVec(f, el.pe_k, npe) && readOne(f, nsh) && readVec(f, shells, static_cast<size_t>(nsh) * 6)
            && readVec(f, fl, 7);
        if (!ok || el.A <= 0 || nsh == 0) {
            err = "physics list element block truncated";
            return false;
        }
        for (uint32_t s = 0; s < nsh; ++s) {
            const double* r = &shells[6 * s];   // n, l, occupancy, binding, zeff, n*
            el.shells.push_back({ r[2], r[3], r[4] / r[5] });
        }
        el.k_edge = shells[3];
        el.omega_k = fl[0];
        for (int i = 0; i < 3; ++i) {
            el.line_e[i] = fl[1 + i];
            el.line_p[i] = fl[4 + i];
        }
        el.log_inc.resize(ne);
        el.log_coh.resize(ne);
        for (uint32_t i = 0; i < ne; ++i) {
            el.log_inc[i] = safeLog(inc[i]);
            el.log_coh[i] = safeLog(coh[i]);
        }
        el.pe_log_e.resize(npe);
        el.pe_log_v.resize(npe);
        for (uint32_t i = 0; i < npe; ++i) {
            el.pe_log_e[i] = std::log(pe_e[i]);
            el.pe_log_v[i] = safeLog(pe[i]);
        }
    }
    elements_ = std::move(els);
    return true;
}

double PhysicsList::photo(int z, double e) const
{
    const auto& el = element(z);
    return interpLogLog(el.pe_log_e, el.pe_log_v, std::log(e));
}
double PhysicsList::photoK(int z, double e) const
{
    const auto& el = element(z);
    if (e <= el.k_edge)
        return 0.0;
    // the K part is zero below the edge; above it interpolate like the total
    const auto it = std::upper_bound(el.pe_log_e.begin(), el.pe_log_e.end(), std::log(e));
    const size_t b = std::min(static_cast<size_t>(it - el.pe_log_e.begin()), el.pe_log_e.size() - 1);
    const size_t a = b == 0 ? 0 : b - 1;
    const double va = el.pe_k[a], vb = el.pe_k[b];
    if (va <= 0 || vb <= 0 || b == a)
        return std::max(vb, 0.0);
    const double t = (std::log(e) - el.pe_log_e[a]) / (el.pe_log_e[b] - el.pe_log_e[a]);
    return std::exp(std::log(va) + t * (std::log(vb) - std::log(va)));
}

double PhysicsList::incoherent(int z, double e) const { return interpLogLog(log_e_, element(z).log_inc, std::log(e)); }
double PhysicsList::coherent(int z, double e) const { return interpLogLog(log_e_, element(z).log_coh, std::log(e)); }
double PhysicsList::formFactor(int z, double x) const { return interpLinear(x_, element(z).F, x); }
double PhysicsList::scatterFunction(int z, double x) const { return interpLinear(x_, element(z).S, x); }

void MixMaterial::attenuation(const PhysicsList& pl, double e, double& pe, double& inc, double& coh) const
{
    pe = inc = coh = 0;
    for (size_t i = 0; i < Z.size(); ++i) {
        const double n = atoms_per_gram[i] * 1e-24;   // barn -> cm^2
        pe += n * pl.photo(Z[i], e);
        inc += n * pl.incoherent(Z[i], e);
        coh += n * pl.coherent(Z[i], e);
    }
}

double MixMaterial::S(const PhysicsList& pl, double x) const
{
    double s = 0;
    for (size_t i = 0; i < Z.size(); ++i)
        s += atoms_per_gram[i] * pl.scatterFunction(Z[i], x);
    return s;
}

double MixMaterial::F2(const PhysicsList& pl, double x) const
{
    double s = 0;
    for (size_t i = 0; i < Z.size(); ++i) {
        const double f = pl.formFactor(Z[i], x);
        s += atoms_per_gram[i] * f * f;
    }
    return s;
}

bool makeMaterial(const PhysicsList& pl, const int32_t* Z, const double* w, uint64_t n, MixMaterial& out, std::string& err)
{
    if (n == 0 || Z == nullptr || w == nullptr) {
        err = "material has no elements";
        return false;
    }
    double total = 0;
    for (uint64_t i = 0; i < n; ++i) {
        if (Z[i] < 1 || Z[i] > pl.zmax()) {
            err = "material element Z=" + std::to_string(Z[i]) + " outside the physics list";
            return false;
        }
        if (!std::isfinite(w[i]) || w[i] < 0) {
            err = "material mass fraction must be finite and non-negative";
            return false;
        }
        total += w[i];
    }
    if (!(total > 0)) {
        err = "material mass fractions sum to zero";
        return false;
    }
    MixMaterial m;
    for (uint64_t i = 0; i < n; ++i) {
        if (w[i] == 0)
            continue;
        const double apg = (w[i] / total) * kNA / pl.element(Z[i]).A;
        m.Z.push_back(Z[i]);
        m.atoms_per_gram.push_back(apg);
        m.electrons_per_gram += apg * Z[i];
    }
    out = std::move(m);
    return true;
}

// Klein-Nishina per unit cos(theta), up to a constant: P^2 (P + 1/P - sin^2)
// with P = E'/E = 1 / (1 + k (1 - cos)), k = E / m_e c^2.  The cumulative
// integral over cos in [-1, 1] is taken by Simpson's rule on kFine intervals,
// and each node u_i is inverted by linear interpolation inside its interval
// (interval widths of 2/kFine put the inversion error near 1e-8 in cos).
void buildComptonTable(std::vector<float>& out)
{
    constexpr int kFine = 16384;
    out.assign(static_cast<size_t>(kKnRows) * kKnN * 2, 0.f);
    std::vector<double> cum(kFine + 1);
    std::vector<double> node(kKnN);
    for (int r = 0; r < kKnRows; ++r) {
        const double k = (kKnE0 + r * kKnDE) / kMeC2KeV;
        const auto pdf = [k](double c) {
            const double P = 1.0 / (1.0 + k * (1.0 - c));
            return P * P * (P + 1.0 / P - (1.0 - c * c));
        };
        const double h = 2.0 / kFine;
        cum[0] = 0;
        for (int b = 1; b <= kFine; ++b) {
            const double c0 = -1.0 + (b - 1) * h;
            cum[static_cast<size_t>(b)] = cum[static_cast<size_t>(b) - 1] + h / 6.0 * (pdf(c0) + 4.0 * pdf(c0 + 0.5 * h) + pdf(c0 + h));
        }
        const double tot = cum[kFine];
        node[0] = -1.0;
        node[kKnN - 1] = 1.0;
        size_t b = 1;
        for (int i = 1; i < kKnN - 1; ++i) {
            const double target = tot * i / (kKnN - 1);
            while (b < static_cast<size_t>(kFine) && cum[b] < target)
                ++b;
            const double span = cum[b] - cum[b - 1];
            const double f = span > 0 ? (target - cum[b - 1]) / span : 0.0;
            node[static_cast<size_t>(i)] = -1.0 + (static_cast<double>(b - 1) + f) * h;
        }
        for (int i = 0; i < kKnN; ++i) {
            float* e = &out[(static_cast<size_t>(r) * kKnN + i) * 2];
            e[0] = static_cast<float>(node[static_cast<size_t>(i)]);
            e[1] = i + 1 < kKnN ? static_cast<float>(node[static_cast<size_t>(i) + 1]) - e[0] : 0.f;
        }
    }
}

// S(x)/Z on a fine x grid (FP64, interpolated linearly: S is smooth), then
// per row the cumulative integral of KN(cos) S(x(cos))/Z over cos by
// Simpson's rule on kFine intervals, inverted at the table's nodes.
void buildComptonSTable(const PhysicsList& pl, const MixMaterial& m, std::vector<float>& out)
{
    constexpr int kFine = 16384, kSx = 8192;
    const double x_top = std::exp2(static_cast<double>(kKnsRows - 1) / kKnsPerOctave) / kHcKeVAng;   // x of backscatter at the top row
    std::vector<double> sx(kSx + 1);
    for (int k = 0; k <= kSx; ++k)
        sx[static_cast<size_t>(k)] = std::min(1.0, m.S(pl, x_top * k / kSx) / m.electrons_per_gram);
    const auto sOf = [&](double x) {
        const double f = std::min(x / x_top, 1.0) * kSx;
        const int k = std::min(static_cast<int>(f), kSx - 1);
        return sx[static_cast<size_t>(k)] + (f - k) * (sx[static_cast<size_t>(k) + 1] - sx[static_cast<size_t>(k)]);
    };
    std::vector<double> cum(kFine + 1);
    std::vector<double> node(kKnN);
    const size_t base = out.size();
    out.resize(base + static_cast<size_t>(kKnsRows) * kKnN * 2);
    for (int r = 0; r < kKnsRows; ++r) {
        const double e = std::exp2(static_cast<double>(r) / kKnsPerOctave), k = e / kMeC2KeV, xe = e / kHcKeVAng;
        const auto pdf = [&](double c) {
            const double P = 1.0 / (1.0 + k * (1.0 - c));
            return P * P * (P + 1.0 / P - (1.0 - c * c)) * sOf(xe * std::sqrt(std::max(0.0, 0.5 * (1.0 - c))));
        };
        const double h = 2.0 / kFine;
        cum[0] = 0;
        for (int b = 1; b <= kFine; ++b) {
            const double c0 = -1.0 + (b - 1) * h;
            cum[static_cast<size_t>(b)] = cum[static_cast<size_t>(b) - 1] + h / 6.0 * (pdf(c0) + 4.0 * pdf(c0 + 0.5 * h) + pdf(c0 + h));
        }
        const double tot = cum[kFine];
        node[0] = -1.0;
        node[kKnN - 1] = 1.0;
        size_t b = 1;
        for (int i = 1; i < kKnN - 1; ++i) {
            const double w = 1.0 - static_cast<double>(i) / (kKnN - 1);
            const double target = tot * (1.0 - w * w);
            while (b < static_cast<size_t>(kFine) && cum[b] < target)
                ++b;
            const double span = cum[b] - cum[b - 1];
            const double f = span > 0 ? (target - cum[b - 1]) / span : 0.0;
            node[static_cast<size_t>(i)] = -1.0 + (static_cast<double>(b - 1) + f) * h;
        }
        for (int i = 0; i < kKnN; ++i) {
            float* o = &out[base + (static_cast<size_t>(r) * kKnN + i) * 2];
            o[0] = static_cast<float>(node[static_cast<size_t>(i)]);
            o[1] = i + 1 < kKnN ? static_cast<float>(node[static_cast<size_t>(i) + 1]) - o[0] : 0.f;
        }
    }
}

const std::vector<float>& comptonTable()
{
    static const std::vector<float> t = [] {
        std::vector<float> v;
        buildComptonTable(v);
        return v;
    }();
    return t;
}

void buildTables(const PhysicsList& pl, const std::vector<MixMaterial>& mats, const std::vector<double>& rho_max, int level,
                 DeviceTables& t)
{
    const int nm = static_cast<int>(mats.size());
    t.n_mat = nm;
    t.mu_tot.assign(static_cast<size_t>(nm) * kLutN, 0.f);
    t.mu_parts.assign(static_cast<size_t>(nm) * kLutN * 4, 0.f);
    t.majorant.assign(kLutN, 0.f);
    std::vector<double> tot(static_cast<size_t>(nm) * kLutN);
    for (int m = 0; m < nm; ++m) {
        for (int j = 0; j < kLutN; ++j) {
            const double e = kLutE0 + j * kLutDE;
            double pe, inc, coh;
            mats[static_cast<size_t>(m)].attenuation(pl, e, pe, inc, coh);
            const size_t o = static_cast<size_t>(m) * kLutN + j;
            tot[o] = pe + inc + coh;
            t.mu_tot[o] = static_cast<float>(pe + inc + coh);
            t.mu_parts[4 * o + 0] = static_cast<float>(pe);
            t.mu_parts[4 * o + 1] = static_cast<float>(pe + inc);
            t.mu_parts[4 * o + 2] = static_cast<float>(pe + inc + coh);
            t.mu_parts[4 * o + 3] = 0.f;
        }
    }
    // Bin majorant: linear interpolation never exceeds the larger end node, so
    // max over materials of max(node j, node j+1) * rho_max bounds every voxel
    // over [E_j, E_j+1].  The FP32 value is rounded up and padded by 1e-6 to
    // cover the FP32 evaluation of mu * rho in the kernel.
    for (int j = 0; j < kLutN; ++j) {
        double mx = 0;
        for (int m = 0; m < nm; ++m) {
            const size_t o = static_cast<size_t>(m) * kLutN + j;
            const double hi = j + 1 < kLutN ? std::max(tot[o], tot[o + 1]) : tot[o];
            mx = std::max(mx, hi * rho_max[static_cast<size_t>(m)]);
        }
        // a vacuum grid still needs a positive majorant: photons then cross it in one step
        t.majorant[static_cast<size_t>(j)] = roundUp(std::max(mx * (1.0 + 1e-6), 1e-20));
    }
    t.region_gain.assign(static_cast<size_t>(nm), 0.0);
    for (int m = 0; m < nm; ++m)
        for (int j = 0; j < kLutN; ++j) {
            const size_t o = static_cast<size_t>(m) * kLutN + j;
            const double hi = j + 1 < kLutN ? std::max(tot[o], tot[o + 1]) : tot[o];
            t.region_gain[static_cast<size_t>(m)] =
                std::max(t.region_gain[static_cast<size_t>(m)], hi / static_cast<double>(t.majorant[static_cast<size_t>(j)]));
        }
    t.kn_inv = comptonTable();
    t.kns_inv.clear();
    if (level >= 1)
        for (int m = 0; m < nm; ++m)
            buildComptonSTable(pl, mats[static_cast<size_t>(m)], t.kns_inv);
    // Rayleigh sampling tables (correction level 1 and 2).
    constexpr int kFine = 16384;
    t.ray_fwd.assign(static_cast<size_t>(nm) * kSfN, 0.f);
    t.ray_inv.assign(static_cast<size_t>(nm) * kRayInvN, 0.f);
    std::vector<double> ucum(kFine + 1), acum(kFine + 1);
    for (int m = 0; m < nm; ++m) {
        const MixMaterial& mat = mats[static_cast<size_t>(m)];
        // cumulative A(u) = int_0^u F^2(sqrt(u')) du' by the trapezoid rule on a fine x grid
        ucum[0] = 0;
        acum[0] = 0;
        double fprev = mat.F2(pl, 0.0);
        for (int k = 1; k <= kFine; ++k) {
            const double x = kXMax * k / kFine;
            const double u = x * x;
            const double f = mat.F2(pl, x);
            ucum[k] = u;
            acum[k] = acum[k - 1] + 0.5 * (f + fprev) * (u - ucum[k - 1]);
            fprev = f;
        }
        const double atot = acum[kFine];
        for (int k = 0; k < kSfN; ++k) {
            const double x = kXMax * k / (kSfN - 1);
            const double u = x * x;
            // locate u in the fine grid
            const auto it = std::upper_bound(ucum.begin(), ucum.end(), u);
            size_t b = std::min(static_cast<size_t>(it - ucum.begin()), static_cast<size_t>(kFine));
            size_t a = b - 1;
            const double tt = ucum[b] > ucum[a] ? (u - ucum[a]) / (ucum[b] - ucum[a]) : 0.0;
            const double av = std::min(atot, acum[a] + tt * (acum[b] - acum[a]));
            t.ray_fwd[static_cast<size_t>(m) * kSfN + k] = static_cast<float>(av / atot);
        }
        t.ray_fwd[static_cast<size_t>(m) * kSfN + kSfN - 1] = 1.0f;
        for (int k = 0; k < kRayInvN; ++k) {
            const double target = atot * k / (kRayInvN - 1);
            const auto it = std::lower_bound(acum.begin(), acum.end(), target);
            size_t b = std::min(static_cast<size_t>(it - acum.begin()), static_cast<size_t>(kFine));
            double u;
            if (b == 0) {
                u = 0;
            } else {
                const size_t a = b - 1;
                const double span = acum[b] - acum[a];
                u = span > 0 ? ucum[a] + (target - acum[a]) / span * (ucum[b] - ucum[a]) : ucum[b];
            }
            t.ray_inv[static_cast<size_t>(m) * kRayInvN + k] = static_cast<float>(u / (kXMax * kXMax));
        }
    }
    if (level < 2)
        return;

    // ---- impulse approximation: per-material shells and the unit profile inverse CDF
    t.ia_shell.assign(static_cast<size_t>(nm) * kIaShells * 4, 0.f);
    for (int m = 0; m < nm; ++m) {
        const auto sh = impulseShells(pl, mats[static_cast<size_t>(m)]);
        double cum = 0;
        for (int s = 0; s < kIaShells; ++s) {
            float* r = &t.ia_shell[(static_cast<size_t>(m) * kIaShells + s) * 4];
            if (s < static_cast<int>(sh.size())) {
                cum += sh[static_cast<size_t>(s)].electron_fraction;
                r[1] = static_cast<float>(sh[static_cast<size_t>(s)].binding_kev);
                r[2] = static_cast<float>(sh[static_cast<size_t>(s)].p0_mec);
            }
            r[0] = static_cast<float>(std::min(cum, 1.0));
        }
        t.ia_shell[(static_cast<size_t>(m) * kIaShells + kIaShells - 1) * 4] = 1.0f;
    }
    t.ia_inv.assign(kIaInvN, 0.f);
    for (int k = 0; k < kIaInvN; ++k) {
        // invert G by bisection; the ends are clamped at +-8 (G(-8) ~ 1e-5)
        const double u = std::clamp(static_cast<double>(k) / (kIaInvN - 1), 1e-6, 1.0 - 1e-6);
        double lo = -8, hi = 8;
        for (int it = 0; it < 80; ++it) {
            const double mid = 0.5 * (lo + hi);
            (comptonProfileCdf(mid) < u ? lo : hi) = mid;
        }
        t.ia_inv[static_cast<size_t>(k)] = static_cast<float>(0.5 * (lo + hi));
    }

    // ---- K fluorescence: which element's K shell absorbed, per energy node
    t.fl_cum.assign(static_cast<size_t>(nm) * kLutN * kFlMax, 0.f);
    t.fl_line.assign(static_cast<size_t>(nm) * kFlMax * 8, 0.f);
    for (int m = 0; m < nm; ++m) {
        const MixMaterial& mat = mats[static_cast<size_t>(m)];
        std::vector<size_t> fl;   // element slots with a trackable K line
        for (size_t i = 0; i < mat.Z.size() && fl.size() < static_cast<size_t>(kFlMax); ++i) {
            const ElementData& el = pl.element(mat.Z[i]);
            if (el.omega_k > 0 && el.line_e[0] >= kCutoffKeV)
                fl.push_back(i);
        }
        for (size_t k = 0; k < fl.size(); ++k) {
            const ElementData& el = pl.element(mat.Z[fl[k]]);
            float* r = &t.fl_line[(static_cast<size_t>(m) * kFlMax + k) * 8];
            r[0] = static_cast<float>(el.omega_k);
            r[1] = static_cast<float>(el.line_e[0]);
            r[2] = static_cast<float>(el.line_e[1]);
            r[3] = static_cast<float>(el.line_e[2]);
            r[4] = static_cast<float>(el.line_p[0]);
            r[5] = static_cast<float>(el.line_p[0] + el.line_p[1]);
        }
        for (int j = 0; j < kLutN; ++j) {
            const double e = kLutE0 + j * kLutDE;
            double total = 0;
            for (size_t i = 0; i < mat.Z.size(); ++i)
                total += mat.atoms_per_gram[i] * pl.photo(mat.Z[i], e);
            double cum = 0;
            for (int k = 0; k < kFlMax; ++k) {
                if (k < static_cast<int>(fl.size()) && total > 0)
                    cum += mat.atoms_per_gram[fl[static_cast<size_t>(k)]] * pl.photoK(mat.Z[fl[static_cast<size_t>(k)]], e) / total;
                t.fl_cum[(static_cast<size_t>(m) * kLutN + j) * kFlMax + k] = static_cast<float>(std::min(cum, 1.0));
            }
        }
    }
}

double comptonProfileCdf(double x)
{
    const double q = 1.0 + x * x;
    const double prim = x / (4.0 * q * q) + 3.0 * x / (8.0 * q) + 0.375 * std::atan(x);
    return 0.5 + prim / (0.375 * 3.14159265358979323846);
}

std::vector<IaShell> impulseShells(const PhysicsList& pl, const MixMaterial& m)
{
    struct S {
        double e, b, p0;
    };
    std::vector<S> all;
    double total = 0;
    for (size_t i = 0; i < m.Z.size(); ++i)
        for (const auto& s : pl.element(m.Z[i]).shells) {
            const double e = m.atoms_per_gram[i] * s.occupancy;
            all.push_back({ e, s.binding_kev, s.p0_au * kFineStructure });
            total += e;
        }
    std::stable_sort(all.begin(), all.end(), [](const S& a, const S& b) { return a.b > b.b; });
    std::vector<IaShell> out;
    for (size_t i = 0; i < all.size() && i + 1 < static_cast<size_t>(kIaShells); ++i)
        out.push_back({ all[i].e / total, all[i].b, all[i].p0 });
    if (all.size() >= static_cast<size_t>(kIaShells)) {
        double e = 0, p2 = 0, bmin = 1e300;
        for (size_t i = kIaShells - 1; i < all.size(); ++i) {
            e += all[i].e;
            p2 += all[i].e * all[i].p0 * all[i].p0;
            bmin = std::min(bmin, all[i].b);
        }
        out.push_back({ e / total, bmin, std::sqrt(p2 / e) });
    }
    return out;
}

const PhysicsList* sharedPhysics(std::string& err)
{
    static std::mutex mtx;
    static std::unique_ptr<PhysicsList> list;
    std::lock_guard<std::mutex> lk(mtx);
    if (!list) {
        auto pl = std::make_unique<PhysicsList>();
        if (!pl->load(defaultPhysicsListPath(), err))
            return nullptr;
        list = std::move(pl);
    }
    return list.get();
}

std::string defaultPhysicsListPath()
{
    if (const char* env = std::getenv("DXG_PHYSICS_LIST"))
        return env;
    Dl_info info {};
    if (dladdr(reinterpret_cast<void*>(&defaultPhysicsListPath), &info) && info.dli_fname) {
        std::string so(info.dli_fname);
        const auto slash = so.find_last_of('/');
        const std::string dir = slash == std::string::npos ? "." : so.substr(0, slash);
        return dir + "/physics/physics_list.bin";
    }
    return "physics/physics_list.bin";
}

std::vector<float> cellRatios(const DeviceTables& t, const uint64_t dims[3], const std::vector<double>& density,
                              const std::vector<uint8_t>& material, int cell, int cn[3])
{
    for (int a = 0; a < 3; ++a)
        cn[a] = static_cast<int>((dims[a] + static_cast<uint64_t>(cell) - 1) / static_cast<uint64_t>(cell));
    const size_t ncell = static_cast<size_t>(cn[0]) * cn[1] * cn[2];
    const int nm = t.n_mat;
    // densest voxel of every material in every cell
    std::vector<double> rho(ncell * static_cast<size_t>(nm), 0.0);
    for (uint64_t z = 0; z < dims[2]; ++z)
        for (uint64_t y = 0; y < dims[1]; ++y) {
            const size_t row = (static_cast<size_t>(z / cell) * cn[1] + y / cell) * cn[0];
            const size_t v0 = (z * dims[1] + y) * dims[0];
            for (uint64_t x = 0; x < dims[0]; ++x) {
                const size_t c = row + x / cell;
                double& r = rho[c * nm + material[v0 + x]];
                r = std::max(r, density[v0 + x]);
            }
        }
    // max over j and m commute: the ratio is max_m g_m rho_m(c); padded so
    // that the FP32 product ratio * majorant still bounds the FP32 mu * rho
    std::vector<float> ratio(ncell, 0.0f);
    for (size_t c = 0; c < ncell; ++c) {
        double r = 0.0;
        for (int m = 0; m < nm; ++m)
            r = std::max(r, t.region_gain[static_cast<size_t>(m)] * rho[c * nm + m]);
        ratio[c] = r > 0.0 ? roundUp(r * (1.0 + 1e-5)) : 0.0f;
    }
    return ratio;
}

}  // namespace dxg
