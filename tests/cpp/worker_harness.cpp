// Headless reproduction of OpenDXMC's SimulationPipeline::worker<CORRECTION>
// (simulationpipeline.cpp:124-235) on the C++ facade: same World /
// AAVoxelGrid / Transport / TransportProgress calls, same post-processing
// (air dose zeroed, mGy -> uGy switch below 1 mGy, variance scaled with it).
//
// usage: worker_harness <phantom.bin> <bowtie.txt|-> <out.bin> <correction 0|1> <seed> <particles per exposure>
// phantom.bin: u64 dims[3], f64 spacing[3], u64 n_mats, per material u64 n then
//              n x (u64 Z, f64 mass fraction), f64 density[N], u8 material[N]
// bowtie.txt:  "angle weight" lines (or "-" for none)
// out.bin:     f64 dose[N], f64 count[N], f64 variance[N], then the unit string
// The beam is a 120 kV CT spiral (6 mm Al) over the grid's z extent, 4 cm
// collimation, pitch 1, 15 degree steps.
#include <dxmc/beams/ctspiralbeam.hpp>
#include <dxmc/transport.hpp>
#include <dxmc/world/world.hpp>
#include <dxmc/world/worlditems/aavoxelgrid.hpp>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <map>
#include <memory>
#include <string>
#include <vector>

using Material = dxmc::Material<5>;
using CTSpiralBeam = dxmc::CTSpiralBeam<false>;

struct Data {
    std::array<std::size_t, 3> dims {};
    std::array<double, 3> spacing {};
    std::vector<std::map<std::uint64_t, double>> materials;
    std::vector<double> density;
    std::vector<std::uint8_t> material;
    std::vector<double> dose, count, variance;
    std::string units;
};

template <typename T>
static bool rd(std::ifstream& f, T& v)
{
    f.read(reinterpret_cast<char*>(&v), sizeof(T));
    return f.good();
}

static bool readPhantom(const char* path, Data& d)
{
    std::ifstream f(path, std::ios::binary);
    std::uint64_t dims[3], nm;
    if (!rd(f, dims) || !rd(f, d.spacing) || !rd(f, nm))
        return false;
    for (int i = 0; i < 3; ++i)
        d.dims[i] = dims[i];
    for (std::uint64_t m = 0; m < nm; ++m) {
        std::uint64_t n;
        rd(f, n);
        std::map<std::uint64_t, double> comp;
        for (std::uint64_t k = 0; k < n; ++k) {
            std::uint64_t z;
            double w;
            rd(f, z);
            rd(f, w);
            comp[z] = w;
        }
        d.materials.push_back(comp);
    }
    const std::size_t N = d.dims[0] * d.dims[1] * d.dims[2];
    d.density.resize(N);
    d.material.resize(N);
    f.read(reinterpret_cast<char*>(d.density.data()), static_cast<std::streamsize>(N * sizeof(double)));
    f.read(reinterpret_cast<char*>(d.material.data()), static_cast<std::streamsize>(N));
    return f.good();
}

// SimulationPipeline::worker<CORRECTION>, minus Qt.
template <int CORRECTION = 1>
void worker(bool deleteAirDose, int nthreads, Data& data, const std::vector<std::shared_ptr<CTSpiralBeam>>& beams,
            dxmc::TransportProgress* progress, std::uint64_t seed)
{
    using VoxelGrid = dxmc::AAVoxelGrid<5, CORRECTION, 255>;
    using World = dxmc::World<VoxelGrid>;

    World world;
    auto& vgrid = world.template addItem<VoxelGrid>();
    {
        std::vector<Material> materials;
        for (const auto& t : data.materials) {
            auto material = Material::byWeight(t);
            if (!material) {
                progress->setStopSimulation();
                return;
            }
            materials.push_back(material.value());
        }
        vgrid.setData(data.dims, data.density, data.material, materials);
        vgrid.setSpacing(data.spacing);
    }
    world.build();

    dxmc::Transport transport;
    transport.setRandomSeed(seed);
    if (nthreads > 0)
        transport.setNumberOfThreads(nthreads);
    for (const auto& beam : beams) {
        transport(world, *beam, progress, true);
        if (!progress->continueSimulation())
            return;
    }

    const auto N = vgrid.size();
    {
        std::vector<double> dose(N);
        for (std::size_t i = 0; i < N; ++i)
            dose[i] = vgrid.doseScored(i).dose();
        if (deleteAirDose)
            std::transform(dose.cbegin(), dose.cend(), data.material.cbegin(), dose.begin(), [](auto d, auto m) { return m > 0 ? d : 0.0; });
        const auto mx = *std::max_element(dose.cbegin(), dose.cend());
        if (mx < 1) {
            std::transform(dose.cbegin(), dose.cend(), dose.begin(), [](auto d) { return d * 1e3; });
            data.units = "uGy";
        } else {
            data.units = "mGy";
        }
        data.dose = dose;
    }
    {
        std::vector<double> c(N, 0);
        for (std::size_t i = 0; i < N; ++i)
            c[i] = static_cast<double>(vgrid.doseScored(i).numberOfEvents());
        if (deleteAirDose)
            std::transform(c.cbegin(), c.cend(), data.material.cbegin(), c.begin(), [](auto d, auto m) { return m > 0 ? d : 0; });
        data.count = c;
    }
    {
        std::vector<double> v(N, 0.0);
        for (std::size_t i = 0; i < N; ++i)
            v[i] = vgrid.doseScored(i).variance();
        if (deleteAirDose)
            std::transform(v.cbegin(), v.cend(), data.material.cbegin(), v.begin(), [](auto d, auto m) { return m > 0 ? d : 0.0; });
        if (data.units[0] == 'u')
            std::for_each(v.begin(), v.end(), [](auto& x) { x *= 1e6; });
        data.variance = v;
    }
    progress->setStopSimulation();
}

int main(int argc, char** argv)
{
    if (argc != 7 && argc != 8) {
        std::cerr << "usage: worker_harness phantom.bin bowtie.txt|- out.bin correction seed ppe [ctdivol_mgy]\n";
        return 2;
    }
    Data data;
    if (!readPhantom(argv[1], data)) {
        std::cerr << "cannot read phantom\n";
        return 2;
    }
    const int correction = std::atoi(argv[4]);
    const std::uint64_t seed = std::strtoull(argv[5], nullptr, 10);
    const std::uint64_t ppe = std::strtoull(argv[6], nullptr, 10);

    const double hz = 0.5 * data.dims[2] * data.spacing[2];
    auto beam = std::make_shared<CTSpiralBeam>(std::array<double, 3> { 0, 0, -hz }, std::array<double, 3> { 0, 0, hz },
                                               std::map<std::size_t, double> { { 13, 6.0 } });
    beam->setTubeVoltage(120);
    beam->setCollimation(4.0);
    beam->setPitch(1.0);
    beam->setStepAngleDeg(15);
    beam->setNumberOfParticlesPerExposure(ppe);
    if (argc == 8)
        beam->setCTDIvol(std::strtod(argv[7], nullptr));   // default: 0, uncalibrated
    if (std::string(argv[2]) != "-") {
        std::ifstream bf(argv[2]);
        std::vector<std::pair<double, double>> nodes;
        double a, w;
        while (bf >> a >> w)
            nodes.emplace_back(a, w);
        beam->setBowtieFilter(dxmc::BowtieFilter(nodes));
    }
    dxmc::TransportProgress progress;
    if (correction == 0)
        worker<0>(true, 0, data, { beam }, &progress, seed);
    else
        worker<1>(true, 0, data, { beam }, &progress, seed);
    if (data.dose.empty()) {
        std::cerr << "simulation failed: " << dxmc::lastError() << "\n";
        return 1;
    }
    std::ofstream out(argv[3], std::ios::binary);
    out.write(reinterpret_cast<const char*>(data.dose.data()), static_cast<std::streamsize>(data.dose.size() * 8));
    out.write(reinterpret_cast<const char*>(data.count.data()), static_cast<std::streamsize>(data.count.size() * 8));
    out.write(reinterpret_cast<const char*>(data.variance.data()), static_cast<std::streamsize>(data.variance.size() * 8));
    out << data.units;
    std::cout << "histories " << beam->numberOfParticles() << " units " << data.units << "\n";
    return 0;
}
