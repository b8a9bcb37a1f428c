// The 3D regression case of the reference's jlabos doctest (cylinderFlow3D,
// jlabos/src/testing/test_cylinder.py:78-130), written against the Palabos API
// and running on the MI355X engine: a D3Q19 BGK cavity with regularized velocity
// conditions (createLocalBoundaryCondition3D) on every face, edge and corner,
// zero velocity except the lid, u = (0.02, 0, 0); Re = 100 on ny.
//
// usage: cylinder3d nx ny nz steps out.bin [float|double]
// Writes the final populations (T, [nx][ny][nz][19]) to out.bin and prints the
// averages as one JSON line.
#include <cstdio>
#include <cstdlib>
#include <string>

#include "palabos_amd.hh"

using namespace plb;

#define D descriptors::D3Q19Descriptor

template<typename T>
int run(plint nx, plint ny, plint nz, plint steps, const char* out) {
    // as the doctest: nu and tau in double, omega = floatT(1 / tau)
    const double uMax = 0.02, Re = 100.;
    const double nu = uMax * (double)ny / Re;
    const double tau = 3 * nu + 0.5;
    MultiBlockLattice3D<T, D> lattice(nx, ny, nz, new BGKdynamics<T, D>((T)(1. / tau)));
    OnLatticeBoundaryCondition3D<T, D>* bc = createLocalBoundaryCondition3D<T, D>();
    bc->setVelocityConditionOnBlockBoundaries(lattice);
    setBoundaryVelocity(lattice, lattice.getBoundingBox(), Array<T, 3>((T)0, (T)0, (T)0));
    setBoundaryVelocity(lattice, Box3D(1, nx - 2, 1, ny - 2, nz - 1, nz - 1), Array<T, 3>((T)uMax, (T)0, (T)0));
    initializeAtEquilibrium(lattice, lattice.getBoundingBox(), (T)1, Array<T, 3>((T)0, (T)0, (T)0));
    lattice.initialize();
    for (plint i = 0; i < steps; ++i) lattice.collideAndStream();

    FILE* fo = std::fopen(out, "wb");
    if (!fo) return 2;
    for (plint x = 0; x < nx; ++x)
        for (plint y = 0; y < ny; ++y)
            for (plint z = 0; z < nz; ++z) {
                Cell<T, D> cell = lattice.get(x, y, z);
                std::fwrite(cell.getRawPopulations(), sizeof(T), 19, fo);
            }
    std::fclose(fo);
    std::printf("{\"avg_energy\": %.17g, \"avg_density\": %.17g, \"steps\": %ld}\n",
                (double)computeAverageEnergy(lattice), (double)computeAverageDensity(lattice), (long)steps);
    delete bc;
    return 0;
}

int main(int argc, char* argv[]) {
    if (argc < 6) {
        std::fprintf(stderr, "usage: cylinder3d nx ny nz steps out.bin [float|double]\n");
        return 2;
    }
    const plint nx = std::atol(argv[1]), ny = std::atol(argv[2]), nz = std::atol(argv[3]);
    const plint steps = std::atol(argv[4]);
    const bool useFloat = argc > 6 && std::string(argv[6]) == "float";
    try {
        return useFloat ? run<float>(nx, ny, nz, steps, argv[5]) : run<double>(nx, ny, nz, steps, argv[5]);
    } catch (PlbException const& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
}
