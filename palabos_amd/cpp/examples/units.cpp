// Prints the IncomprFlowParam quantities of the facade (palabos_amd.hh) in the
// format of the reference driver's units mode, for the parity test of the
// class against Palabos's core/units.h:39-95.
//   units physU latU Re physLen N lx ly lz t   (physU = 0: the six-argument constructor)
#include <cstdio>
#include <cstdlib>

#include "palabos_amd.hh"

using namespace plb;

int main(int argc, char* argv[]) {
    if (argc < 10) {
        std::fprintf(stderr, "usage: units physU latU Re physLen N lx ly lz t\n");
        return 2;
    }
    const double physU = std::atof(argv[1]), latU = std::atof(argv[2]), Re = std::atof(argv[3]);
    const double physLen = std::atof(argv[4]);
    const plint N = std::atol(argv[5]);
    const double lx = std::atof(argv[6]), ly = std::atof(argv[7]), lz = std::atof(argv[8]), t = std::atof(argv[9]);
    IncomprFlowParam<double> p = physU != 0 ? IncomprFlowParam<double>(physU, latU, Re, physLen, N, lx, ly, lz)
                                            : IncomprFlowParam<double>(latU, Re, N, lx, ly, lz);
    std::printf("%ld %ld %ld %ld %ld %ld %.17g %.17g %.17g %.17g %.17g %ld %.17g %.17g %.17g %.17g %.17g\n",
                (long)p.getNx(), (long)p.getNy(), (long)p.getNz(), (long)p.getNx(true), (long)p.getNy(true),
                (long)p.getNz(true), p.getLatticeNu(), p.getTau(), p.getOmega(), p.getDeltaX(), p.getDeltaT(),
                (long)p.nStep(t), p.getLx(), p.getLy(), p.getLz(), p.getPhysicalU(), p.getPhysicalLength());
    return 0;
}
