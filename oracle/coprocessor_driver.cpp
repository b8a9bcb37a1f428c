// Palabos's own co-processor path (coProcessors/, the reference's plug-in seam)
// driven with the MI355X engine behind it.
//
// AmdCoProcessor3D<T> derives from the reference's real plb::CoProcessor3D<T>
// (coProcessors/coProcessor3D.h:59-116) and forwards its four methods to the
// engine's C ABI (include/palabos_amd.h) only. Palabos reaches its co-processor
// through the singleton global::defaultCoProcessor3D<T>() (coProcessor3D.h:
// 180-186); the explicit specializations below are the switch a Palabos build
// makes to use the engine, without touching the reference's sources.
//
// The program builds two periodic MultiBlockLattice3D with the same uneven
// blocks and the same seeded populations: one stepped by Palabos on the CPU,
// one whose pure-BGK blocks are offloaded by Palabos's initiateCoProcessors /
// transferToCoProcessors / transferFromCoProcessors
// (coProcessorCommunication3D.hh:36-123) and MultiBlockLattice3D::collideAndStream
// (multiBlockLattice3D.hh:318-331). The co-processor updates each block with its
// envelope; the plug-in contract leaves the outer layer undefined, so every step
// is one transfer cycle followed by the envelope refresh Palabos does itself
// (MultiBlock3D::duplicateOverlaps). Prints one JSON line: whether every bulk
// population of the offloaded lattice equals the CPU one bit for bit.
//
//   coprocessor_driver double|float nx ny nz xcut steps omega
//
// Test infrastructure (built by oracle/Makefile.ref where the reference sources
// are, linked to palabos_amd/libpalabos_amd.so); never part of the product.
#include "core/globalDefs.h"
#include "parallelism/headers3D.h"
#include "latticeBoltzmann/headers3D.h"
#include "core/headers3D.h"
#include "basicDynamics/headers3D.h"
#include "boundaryCondition/headers3D.h"
#include "io/headers3D.h"
#include "atomicBlock/headers3D.h"
#include "multiBlock/headers3D.h"
#include "algorithm/headers3D.h"
#include "dataProcessors/headers3D.h"
#include "coProcessors/headers3D.h"
#include "parallelism/headers3D.hh"
#include "latticeBoltzmann/headers3D.hh"
#include "core/headers3D.hh"
#include "basicDynamics/headers3D.hh"
#include "boundaryCondition/headers3D.hh"
#include "io/headers3D.hh"
#include "atomicBlock/headers3D.hh"
#include "multiBlock/headers3D.hh"
#include "algorithm/headers3D.hh"
#include "dataProcessors/headers3D.hh"
#include "coProcessors/headers3D.hh"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "palabos_amd.h"

using namespace plb;

template<typename T> struct AmdStorage;
template<> struct AmdStorage<double> { static const int code = PLBAMD_F64; };
template<> struct AmdStorage<float> { static const int code = PLBAMD_F32; };

static plbamd_box amdBox(Box3D const& b) {
    plbamd_box c;
    c.x0 = (int32_t)b.x0; c.x1 = (int32_t)b.x1; c.y0 = (int32_t)b.y0;
    c.y1 = (int32_t)b.y1; c.z0 = (int32_t)b.z0; c.z1 = (int32_t)b.z1;
    return c;
}

// D3Q19 BGK domains on the engine; 1 = success, 0 = failure (coProcessor3D.h:73-75).
template<typename T>
class AmdCoProcessor3D : public CoProcessor3D<T> {
public:
    virtual int addDomain(plint nx, plint ny, plint nz, T omega, int& domainHandle) {
        static const bool ready = plbamd_init(0) == PLBAMD_OK;
        if (!ready) return 0;
        plbamd_lattice_desc d;
        std::memset(&d, 0, sizeof(d));
        d.nx = (int32_t)nx; d.ny = (int32_t)ny; d.nz = (int32_t)nz;
        d.storage = AmdStorage<T>::code;
        d.dynamics = PLBAMD_BGK;
        d.omega = (double)omega;
        if (plbamd_lattice_create(&d, &domainHandle) != PLBAMD_OK) return 0;
        return plbamd_toggle_statistics(domainHandle, 0) == PLBAMD_OK ? 1 : 0;
    }
    virtual int send(int domainHandle, Box3D const& subDomain, std::vector<char> const& data) {
        plbamd_box b = amdBox(subDomain);
        return plbamd_upload(domainHandle, &b, AmdStorage<T>::code, &data[0], data.size()) == PLBAMD_OK ? 1 : 0;
    }
    virtual int receive(int domainHandle, Box3D const& subDomain, std::vector<char>& data) const {
        data.resize((size_t)subDomain.nCells() * 19 * sizeof(T));
        plbamd_box b = amdBox(subDomain);
        return plbamd_download(domainHandle, &b, AmdStorage<T>::code, &data[0], data.size()) == PLBAMD_OK ? 1 : 0;
    }
    virtual int collideAndStream(int domainHandle) {
        return plbamd_collide_and_stream(domainHandle, 1) == PLBAMD_OK ? 1 : 0;
    }
};

namespace plb {
namespace global {
template<> inline CoProcessor3D<double>& defaultCoProcessor3D<double>() {
    static AmdCoProcessor3D<double> singleton;
    return singleton;
}
template<> inline CoProcessor3D<float>& defaultCoProcessor3D<float>() {
    static AmdCoProcessor3D<float> singleton;
    return singleton;
}
}  // namespace global
}  // namespace plb

namespace {

typedef descriptors::D3Q19Descriptor<double> Dd;

// seeded non-equilibrium state: equilibrium of a smooth field plus a small hash-based perturbation
double seeded(plint x, plint y, plint z, int i) {
    unsigned long long h = (unsigned long long)(((x * 1009 + y) * 1013 + z) * 19 + i) * 0x9E3779B97F4A7C15ull;
    h ^= h >> 31;
    h *= 0xBF58476D1CE4E5B9ull;
    h ^= h >> 29;
    return ((double)(h >> 11) * (1.0 / 9007199254740992.0) - 0.5) * 1e-3;
}

template<typename T>
MultiBlockLattice3D<T, descriptors::D3Q19Descriptor>* makeLattice(plint nx, plint ny, plint nz, plint xcut, T omega,
                                                                   bool offload) {
    typedef MultiBlockLattice3D<T, descriptors::D3Q19Descriptor> Lattice;
    SparseBlockStructure3D sparse(Box3D(0, nx - 1, 0, ny - 1, 0, nz - 1));
    sparse.addBlock(Box3D(0, xcut - 1, 0, ny - 1, 0, nz - 1), 0);
    sparse.addBlock(Box3D(xcut, nx - 1, 0, ny - 1, 0, nz - 1), 1);
    ThreadAttribution* attribution;
    if (offload) {
        ExplicitThreadAttribution* a = new ExplicitThreadAttribution;
        a->addBlock(0, 0);
        a->addBlock(1, 0);
        attribution = a;
    } else {
        attribution = defaultMultiBlockPolicy3D().getThreadAttribution();
    }
    MultiBlockManagement3D management(sparse, attribution, 1);
    Lattice* lat = new Lattice(management, defaultMultiBlockPolicy3D().getBlockCommunicator(),
                               defaultMultiBlockPolicy3D().getCombinedStatistics(),
                               defaultMultiBlockPolicy3D().getMultiCellAccess<T, descriptors::D3Q19Descriptor>(),
                               new BGKdynamics<T, descriptors::D3Q19Descriptor>(omega));
    lat->periodicity().toggleAll(true);
    for (plint x = 0; x < nx; ++x)
        for (plint y = 0; y < ny; ++y)
            for (plint z = 0; z < nz; ++z) {
                const double rho = 1.0 + 0.01 * std::sin(0.3 * x + 0.2 * y);
                const double u[3] = {0.02 * std::cos(0.25 * z), -0.01 * std::sin(0.35 * x), 0.015};
                double j[3] = {rho * u[0], rho * u[1], rho * u[2]};
                const double jSqr = j[0] * j[0] + j[1] * j[1] + j[2] * j[2];
                Cell<T, descriptors::D3Q19Descriptor>& cell = lat->get(x, y, z);
                for (int i = 0; i < 19; ++i) {
                    Array<double, 3> jj(j[0], j[1], j[2]);
                    const double feq = dynamicsTemplates<double, descriptors::D3Q19Descriptor>::bgk_ma2_equilibrium(
                        i, rho - 1.0, 1.0 / rho, jj, jSqr);
                    cell[i] = (T)(feq + seeded(x, y, z, i));
                }
            }
    lat->initialize();
    return lat;
}

template<typename T>
int run(plint nx, plint ny, plint nz, plint xcut, plint steps, T omega) {
    typedef MultiBlockLattice3D<T, descriptors::D3Q19Descriptor> Lattice;
    Lattice* cpu = makeLattice<T>(nx, ny, nz, xcut, omega, false);
    Lattice* dev = makeLattice<T>(nx, ny, nz, xcut, omega, true);
    initiateCoProcessors(*dev, BGKdynamics<T, descriptors::D3Q19Descriptor>(omega).getId(), false);
    const ThreadAttribution& ta = dev->getMultiBlockManagement().getThreadAttribution();
    int handles = 0;
    for (plint id = 0; id < 2; ++id) handles += ta.getCoProcessorHandle(id) >= 0 ? 1 : 0;
    for (plint s = 0; s < steps; ++s) {
        cpu->collideAndStream();
        transferToCoProcessors(*dev);
        dev->collideAndStream();
        transferFromCoProcessors(*dev);
        dev->duplicateOverlaps(modif::staticVariables);
    }
    plint differ = 0;
    double maxDiff = 0;
    for (plint x = 0; x < nx; ++x)
        for (plint y = 0; y < ny; ++y)
            for (plint z = 0; z < nz; ++z)
                for (int i = 0; i < 19; ++i) {
                    const T a = cpu->get(x, y, z)[i], b = dev->get(x, y, z)[i];
                    if (std::memcmp(&a, &b, sizeof(T)) != 0) {
                        ++differ;
                        maxDiff = std::max(maxDiff, std::fabs((double)a - (double)b));
                    }
                }
    std::printf("{\"handles\": %d, \"steps\": %ld, \"populations\": %ld, \"differ\": %ld, \"max_abs_diff\": %.3e}\n",
                handles, (long)steps, (long)(nx * ny * nz * 19), (long)differ, maxDiff);
    delete cpu;
    delete dev;
    return 0;
}

}  // namespace

int main(int argc, char* argv[]) {
    plbInit(&argc, &argv);
    if (argc < 8) {
        std::fprintf(stderr, "usage: coprocessor_driver double|float nx ny nz xcut steps omega\n");
        return 2;
    }
    const plint nx = std::atol(argv[2]), ny = std::atol(argv[3]), nz = std::atol(argv[4]);
    const plint xcut = std::atol(argv[5]), steps = std::atol(argv[6]);
    const double omega = std::atof(argv[7]);
    return std::string(argv[1]) == "float" ? run<float>(nx, ny, nz, xcut, steps, (float)omega)
                                           : run<double>(nx, ny, nz, xcut, steps, omega);
}
