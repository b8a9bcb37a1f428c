// Lid-driven cavity written against the Palabos API (compare
// jlabos/examples/Cavity3d and the regression case of
// jlabos/src/testing/test_cylinder.py:78-130), running on the MI355X engine.
//
// D3Q19 BGK, omega = 1/0.6, rho0 = 1, f at rest; Zou-He lid on the top z face
// interior with u = (0.05, 0, 0); the other walls are the domain edges.
//
// usage: cavity3d N steps out.bin [float|double] [wrapped]
// Writes the final populations (T, [N][N][N][19]) to out.bin and prints the
// averages and statistics as one JSON line.
#include <cstdio>
#include <cstdlib>
#include <string>

#include "palabos_amd.hh"

using namespace plb;

template<typename T>
int run(plint n, plint steps, const char* out, bool wrapped) {
    typedef MultiBlockLattice3D<T, descriptors::D3Q19Descriptor> Lattice;
    const T omega = (T)1 / (T)0.6;
    Lattice lattice(n, n, n, new BGKdynamics<T, descriptors::D3Q19Descriptor>(omega));
    OnLatticeBoundaryCondition3D<T, descriptors::D3Q19Descriptor>* bc =
        wrapped ? createZouHeBoundaryCondition3D<T, descriptors::D3Q19Descriptor>()
                : createDynamicsBasedZouHeBoundaryCondition3D<T, descriptors::D3Q19Descriptor>();
    Box3D lid(1, n - 2, 1, n - 2, n - 1, n - 1);
    bc->addVelocityBoundary2P(lid, lattice);
    setBoundaryVelocity(lattice, lid, Array<T, 3>((T)0.05, (T)0, (T)0));
    initializeAtEquilibrium(lattice, lattice.getBoundingBox(), (T)1, Array<T, 3>((T)0, (T)0, (T)0));
    lattice.initialize();
    for (plint i = 0; i < steps; ++i) lattice.collideAndStream();

    FILE* fo = std::fopen(out, "wb");
    if (!fo) return 2;
    for (plint x = 0; x < n; ++x)
        for (plint y = 0; y < n; ++y)
            for (plint z = 0; z < n; ++z) {
                Cell<T, descriptors::D3Q19Descriptor> cell = lattice.get(x, y, z);
                std::fwrite(cell.getRawPopulations(), sizeof(T), 19, fo);
            }
    std::fclose(fo);
    Array<T, 3> u;
    lattice.get(n / 2, n / 2, n / 2).computeVelocity(u);
    // analysis fields (dataAnalysisWrapper3D.h): kinetic energy and velocity norm
    // next to the lid's centre
    std::unique_ptr<MultiScalarField3D<T>> energy = computeKineticEnergy(lattice);
    std::unique_ptr<MultiScalarField3D<T>> norm = computeVelocityNorm(lattice);
    std::unique_ptr<MultiTensorField3D<T, 3>> velocity = computeVelocity(lattice);
    const plint c = n / 2, top = n - 2;
    std::printf("{\"avg_density\": %.17g, \"avg_energy\": %.17g, \"stored_density\": %.17g, "
                "\"stored_energy\": %.17g, \"center_ux\": %.17g, \"steps\": %ld, "
                "\"energy_below_lid\": %.17g, \"norm_below_lid\": %.17g, \"ux_below_lid\": %.17g}\n",
                (double)computeAverageDensity(lattice), (double)computeAverageEnergy(lattice),
                getStoredAverageDensity<T>(lattice), getStoredAverageEnergy<T>(lattice), (double)u[0], (long)steps,
                (double)energy->get(c, c, top), (double)norm->get(c, c, top), (double)velocity->get(c, c, top)[0]);
    delete bc;
    return 0;
}

int main(int argc, char* argv[]) {
    if (argc < 4) {
        std::fprintf(stderr, "usage: cavity3d N steps out.bin [float|double] [wrapped]\n");
        return 2;
    }
    plint n = std::atol(argv[1]), steps = std::atol(argv[2]);
    bool useFloat = argc > 4 && std::string(argv[4]) == "float";
    bool wrapped = argc > 5 && std::string(argv[5]) == "wrapped";
    try {
        return useFloat ? run<float>(n, steps, argv[3], wrapped) : run<double>(n, steps, argv[3], wrapped);
    } catch (PlbException const& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
}
