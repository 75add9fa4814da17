// Host-side physics for the transport engine: the physics list reader, the
// Material<5>::byWeight mixture rule and the FP32 device tables built by
// World::build() (simulationpipeline.cpp:133-153).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace dxg {

constexpr double kNA = 6.02214076e23;
constexpr double kHcKeVAng = 12.398419843320026;   // h c [keV Angstrom]
constexpr double kMeC2KeV = 510.99895000;          // electron rest energy [keV]
constexpr double kKeVToJ = 1.602176634e-16;

// Attenuation table energy grid: uniform, E_j = kLutE0 + j * kLutDE.
constexpr double kLutE0 = 1.0;
constexpr double kLutDE = 0.25;
constexpr double kLutEMax = 200.0;
constexpr int kLutN = 797;   // (kLutEMax - kLutE0) / kLutDE + 1

// Form-factor sampling tables: x = sin(theta/2)/lambda in [0, kXMax] (1/Angstrom).
constexpr double kXMax = 16.0;
constexpr int kSfN = 1024;   // nodes of the F^2 forward table
constexpr int kRayInvN = 1024;   // nodes of the inverse F^2 table

// Klein-Nishina inverse-CDF table: rows at E_r = kKnE0 + r kKnDE (keV), r <
// kKnRows; nodes at u_i = i / (kKnN - 1).  Entry (r, i) holds cos(theta) with
// P(cos <= c) = u_i under the Klein-Nishina distribution at E_r, and the step
// to node i + 1 (0 at the last node).
constexpr double kKnE0 = 1.0;
constexpr double kKnDE = 4.0;
constexpr int kKnRows = 51;   // E_r up to kLutEMax + 1
constexpr int kKnN = 128;
constexpr int kKnsPerOctave = 12;   // rows per octave of the Klein-Nishina x S(x) tables
constexpr int kKnsRows = 93;        // E_r = 2^(r / 12) keV up to 2^(92/12) = 203 keV

// Photons below this energy deposit what is left and stop.
constexpr double kCutoffKeV = 1.0;

struct ElementData {
    double A = 0;
    std::vector<double> log_inc, log_coh;     // on the common log-energy grid
    std::vector<double> F, S;                 // on the x grid
    std::vector<double> pe_log_e, pe_log_v;   // edge-resolved photoelectric table
    std::vector<double> pe_k;                 // K-shell part at the same nodes (barn/atom)
    struct Shell {
        double occupancy, binding_kev, p0_au;   // p0 = zeff / n*: Compton profile momentum scale
    };
    std::vector<Shell> shells;
    // K fluorescence: yield, line energies (Ka1, Ka2, Kb) and their probabilities
    double omega_k = 0, line_e[3] = { 0, 0, 0 }, line_p[3] = { 0, 0, 0 };
    double k_edge = 0;
};

class PhysicsList {
public:
    bool load(const std::string& path, std::string& err);
    bool loaded() const { return !elements_.empty(); }
    int zmax() const { return static_cast<int>(elements_.size()); }
    const ElementData& element(int z) const { return elements_[static_cast<size_t>(z - 1)]; }

    // barn/atom at energy E (keV), log-log interpolation
    double photo(int z, double e) const;
    double photoK(int z, double e) const;   // K-shell part of photo()
    double incoherent(int z, double e) const;
    double coherent(int z, double e) const;
    double formFactor(int z, double x) const;
    double scatterFunction(int z, double x) const;

private:
    std::vector<double> log_e_, x_;
    std::vector<ElementData> elements_;
};

// A material built from element mass fractions (Material<5>::byWeight).
struct MixMaterial {
    std::vector<int> Z;
    std::vector<double> atoms_per_gram;
    double electrons_per_gram = 0;

    // cm^2/g
    void attenuation(const PhysicsList& pl, double e, double& pe, double& inc, double& coh) const;
    double S(const PhysicsList& pl, double x) const;    // per gram
    double F2(const PhysicsList& pl, double x) const;   // per gram
};

bool makeMaterial(const PhysicsList& pl, const int32_t* Z, const double* w, uint64_t n, MixMaterial& out, std::string& err);

// FP32 tables in the device layout (see transport.hip for the consumers).
struct DeviceTables {
    int n_mat = 0;
    std::vector<float> mu_tot;     // [m][j] total mass attenuation (cm^2/g), kLutN per material
    std::vector<float> mu_parts;   // [m][j][4]: pe, pe+inc, total, 0 (cm^2/g)
    std::vector<float> majorant;   // [j] 1/cm, valid over [E_j, E_j+1]
    std::vector<float> ray_fwd;    // [m][k] A(x_k)/A(kXMax), A(x) = int_0^{x^2} F^2 du
    std::vector<float> ray_inv;    // [m][k] u(y_k)/kXMax^2, y_k = k/(kRayInvN-1)
    std::vector<float> kn_inv;     // [r][i][2] Klein-Nishina inverse CDF (buildComptonTable)
    std::vector<float> kns_inv;    // [m][r][i][2] Klein-Nishina x S(x) inverse CDF per material, levels 1 and 2 (buildComptonSTable)
    // correction level 2
    std::vector<float> ia_shell;   // [m][kIaShells][4]: cumulative electron fraction, binding (keV), p0 (m_e c), 0
    std::vector<float> ia_inv;     // [k] x(u_k) of the unit Compton profile (8/3pi)(1+x^2)^-3, u_k = k/(kIaInvN-1)
    std::vector<float> fl_cum;     // [m][j][kFlMax]: cumulative P(K-shell absorption in fluorescing element) at E_j
    std::vector<float> fl_line;    // [m][kFlMax][8]: omega_K, E_Ka1, E_Ka2, E_Kb, P(Ka1), P(Ka1)+P(Ka2), 0, 0
    // host only: g_m = max over bins j of binmax_m(j) / majorant_j, so that for
    // a region whose material m reaches density rho_m, max_m g_m rho_m times
    // majorant_j bounds mu * rho of every voxel of the region over bin j
    std::vector<double> region_gain;
};

// Macro-cell majorants: the grid is tiled by cells of `cell` voxels per axis;
// ratio[c] * majorant_j bounds mu * rho over bin j for every voxel of cell c
// (x fastest: c = cx + cnx (cy + cny cz)).  density/material in grid order.
std::vector<float> cellRatios(const DeviceTables& t, const uint64_t dims[3], const std::vector<double>& density,
                              const std::vector<uint8_t>& material, int cell, int cn[3]);

// Impulse approximation: the shells of a material (Material<5>: five).
constexpr int kIaShells = 5;
constexpr int kIaInvN = 2048;
// Fluorescing elements per material whose K lines are tracked.
constexpr int kFlMax = 12;
constexpr double kFineStructure = 7.2973525693e-3;

// Level-2 shell selection of one material: the four most tightly bound
// shells and one shell lumping the rest (minimum binding, rms momentum scale).
struct IaShell {
    double electron_fraction, binding_kev, p0_mec;
};
std::vector<IaShell> impulseShells(const PhysicsList& pl, const MixMaterial& m);

// Unit Compton profile CDF, G(x) = int_-inf^x (8/3pi)(1+t^2)^-3 dt.
double comptonProfileCdf(double x);

// The Klein-Nishina inverse-CDF table, kKnRows x kKnN pairs (see kKnE0):
// material-independent, integrated and inverted in FP64.
void buildComptonTable(std::vector<float>& out);
// The same table, built once per process.
const std::vector<float>& comptonTable();

// Klein-Nishina x S(x)/Z inverse-CDF table of one material (the angular
// distribution of incoherent scattering at correction levels 1 and 2): rows
// at E_r = 2^(r / kKnsPerOctave) keV, r < kKnsRows (S(x) scales with x ~ E sin(theta/2),
// so the distribution's shape changes by the same amount per octave), nodes
// at u_i = 1 - (1 - i/(kKnN-1))^2 (denser towards u = 1, where S(x) -> 0 makes
// cos(u) steep); entry (r, i) = {cos(u_i), cos(u_{i+1}) - cos(u_i)}.  Appends
// kKnsRows x kKnN pairs to out.
void buildComptonSTable(const PhysicsList& pl, const MixMaterial& m, std::vector<float>& out);

void buildTables(const PhysicsList& pl, const std::vector<MixMaterial>& mats, const std::vector<double>& rho_max,
                 int level, DeviceTables& t);

// Locate the physics list next to the shared library (physics/physics_list.bin),
// or at $DXG_PHYSICS_LIST.
std::string defaultPhysicsListPath();

// Process-wide physics list, loaded on first use; nullptr and `err` on failure.
const PhysicsList* sharedPhysics(std::string& err);

// Thread-local message behind dxg_last_error(); returns `code`.
int setError(int code, const std::string& msg);

}  // namespace dxg
