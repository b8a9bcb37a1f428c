// Drives the compiled Palabos reference (built from the sources under the
// reference tree by oracle/Makefile.ref) on a case described by a small text
// file, and writes the post-step populations and statistics as raw binary.
//
// This program is test infrastructure: it produces the golden vectors that pin
// the oracle restatement (oracle/d3q19_oracle.cpp). It is never linked into the
// product library.
//
// Case file, one "key values..." per line:
//   T double|float
//   size nx ny nz
//   dyn bgk|rbgk                      base bulk dynamics
//   omega w
//   periodic px py pz                 0/1 per axis
//   bcstyle dynamics|wrapped          createDynamicsBased...() or wrapped factory
//   bcfamily zouhe|regularized        createZouHe... or createLocal...
//   init file                         double f-bar, C order [nx][ny][nz][19]
//   tags file                         uint8 [nx][ny][nz]: 0 bulk 1 BounceBack 2 NoDynamics 3 MomentumExchangeBounceBack
//   bc vel|press dir ori x0 x1 y0 y1 z0 z1 file [bctype]
//                                     vel: 3 doubles per box cell, press: 1 double rho per box cell,
//                                     box cells in x,y,z C order; bctype: boundary::BcType as an
//                                     integer (boundaryCondition.h:40; default 0 = dirichlet)
//   edge plane n1 n2 x0 x1 y0 y1 z0 z1 file [bctype]
//                                     addExternalVelocityEdge<plane><n1><n2> (n = -1 | 1), velocities
//   corner xn yn zn x y z file [bctype]
//                                     addExternalVelocityCorner<xn><yn><zn>, one velocity
//   blockbc vel|press bctype file     set{Velocity,Pressure}ConditionOnBlockBoundaries on the whole
//                                     lattice; file: per lattice cell, C order, 3 velocity doubles
//                                     (vel) or 1 density (press), defined on every boundary cell
//   block x0 x1 y0 y1 z0 z1           optional; any number -> SparseBlockStructure3D multi-block
//   dump s1 s2 ...                    steps after which populations are written (0 = after initialize)
//   force fx fy fz                    optional; ForcedD3Q19 + GuoExternalForceBGKdynamics (bgk, no planes)
//   fields 1                          optional; also write the analysis fields at every dump
//   external a                        optional; every step is Palabos's external-moment cycle instead
//                                     of collideAndStream: BoxRhoBarJfunctional3D fills rhoBar and j
//                                     (dataAnalysisFunctional3D.hh:528-561), j_x += a * (1 + rhoBar) on
//                                     every cell (in T), ExternalRhoJcollideAndStream3D collides with them
//                                     and streams (dynamicsProcessor3D.hh:41-209); the field after the
//                                     last coupling is written as prefix_rhobarj_s<step>.bin
//                                     (T, [nx][ny][nz][4]: rhoBar jx jy jz)
//   out prefix
// Outputs: prefix_s<step>.bin (T, [nx][ny][nz][19]) and prefix_stats.txt lines;
// with "fields", prefix_fields_s<step>.bin (T, [nx][ny][nz][7]: computeDensity,
// computeRhoBar, computeVelocity x y z, computeKineticEnergy, computeVelocityNorm)
//   step storedAvgRho storedAvgEnergy avgRho(computeAverageDensity) avgEnergy force_x force_y force_z
//
// Bench mode: ref_driver --bench T n steps  -> periodic BGK(1.8) n^3 lattice, prints MLUPS.
//
// Checkpoint mode (the reference's saveBinaryBlock / loadBinaryBlock):
//   ref_driver --checkpoint save T plain|forced nx ny nz records.bin out.dat uint|size
//       records.bin: per cell, C order, the 19 f-bar (+ the 3 force externals for
//       "forced", ForcedD3Q19Descriptor) as T; they are set into a
//       MultiBlockLattice3D, which saveBinaryBlock writes to out.dat
//   ref_driver --checkpoint load T plain|forced nx ny nz in.dat records.bin uint|size
//       loadBinaryBlock of in.dat into a fresh lattice; its cell records are written
//   "uint" passes enforceUint = true (32-bit size header), "size" false.
//
// Units mode (the reference's IncomprFlowParam, core/units.h:39-95):
//   ref_driver --units physU latU Re physLen N lx ly lz t
//       IncomprFlowParam<double>(physU, latU, Re, physLen, N, lx, ly, lz), or the
//       six-argument constructor (latU, Re, N, lx, ly, lz) when physU is 0; prints
//       one line: getNx getNy getNz getNx(true) getNy(true) getNz(true)
//       getLatticeNu getTau getOmega getDeltaX getDeltaT nStep(t) getLx getLy getLz
//       getPhysicalU getPhysicalLength (reals as %.17g).

// The umbrella palabos3D.hh does not compile with g++ 11 (multiGrid), so the
// lattice-core subset is included instead.
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

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>
#include <ctime>

using namespace plb;

namespace {

struct BcSpec {
    std::string kind;        // vel | press | edge | corner
    int dir = 0, ori = 0;    // planes; edges: dir = plane, ori unused
    int normal[3] = {0, 0, 0};  // edges: n1, n2 in normal[1], normal[2]; corners: x, y, z normals
    Box3D box;
    std::string file;
    int bcType = 0;
};

struct BlockBcSpec {
    std::string kind;        // vel | press
    int bcType = 0;
    std::string file;
};

struct CaseSpec {
    std::string T = "double";
    plint nx = 0, ny = 0, nz = 0;
    std::string dyn = "bgk";
    double omega = 1.0;
    int periodic[3] = {0, 0, 0};
    std::string bcstyle = "dynamics";
    std::string bcfamily = "zouhe";
    std::string init, tags, out;
    std::vector<BcSpec> bcs;
    std::vector<BlockBcSpec> blockBcs;
    std::vector<Box3D> blocks;
    std::vector<plint> dumps;
    bool forced = false;
    double force[3] = {0, 0, 0};
    bool fields = false;
    bool external = false;
    double externalA = 0;
};

std::vector<char> readFile(std::string const& name) {
    std::ifstream f(name.c_str(), std::ios::binary);
    if (!f) { std::fprintf(stderr, "cannot open %s\n", name.c_str()); std::exit(2); }
    return std::vector<char>((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
}

CaseSpec parseCase(std::string const& name) {
    CaseSpec c;
    std::ifstream f(name.c_str());
    if (!f) { std::fprintf(stderr, "cannot open %s\n", name.c_str()); std::exit(2); }
    std::string line;
    while (std::getline(f, line)) {
        std::istringstream s(line);
        std::string key;
        if (!(s >> key) || key[0] == '#') continue;
        if (key == "T") s >> c.T;
        else if (key == "size") s >> c.nx >> c.ny >> c.nz;
        else if (key == "dyn") s >> c.dyn;
        else if (key == "omega") s >> c.omega;
        else if (key == "periodic") s >> c.periodic[0] >> c.periodic[1] >> c.periodic[2];
        else if (key == "bcstyle") s >> c.bcstyle;
        else if (key == "bcfamily") s >> c.bcfamily;
        else if (key == "init") s >> c.init;
        else if (key == "tags") s >> c.tags;
        else if (key == "out") s >> c.out;
        else if (key == "force") { s >> c.force[0] >> c.force[1] >> c.force[2]; c.forced = true; }
        else if (key == "external") { s >> c.externalA; c.external = true; }
        else if (key == "bc") {
            BcSpec b;
            s >> b.kind >> b.dir >> b.ori >> b.box.x0 >> b.box.x1 >> b.box.y0 >> b.box.y1
              >> b.box.z0 >> b.box.z1 >> b.file;
            if (!(s >> b.bcType)) b.bcType = 0;
            c.bcs.push_back(b);
        } else if (key == "edge") {
            BcSpec b;
            b.kind = "edge";
            s >> b.dir >> b.normal[1] >> b.normal[2] >> b.box.x0 >> b.box.x1 >> b.box.y0 >> b.box.y1
              >> b.box.z0 >> b.box.z1 >> b.file;
            if (!(s >> b.bcType)) b.bcType = 0;
            c.bcs.push_back(b);
        } else if (key == "corner") {
            BcSpec b;
            b.kind = "corner";
            plint x, y, z;
            s >> b.normal[0] >> b.normal[1] >> b.normal[2] >> x >> y >> z >> b.file;
            b.box = Box3D(x, x, y, y, z, z);
            if (!(s >> b.bcType)) b.bcType = 0;
            c.bcs.push_back(b);
        } else if (key == "blockbc") {
            BlockBcSpec b;
            s >> b.kind >> b.bcType >> b.file;
            c.blockBcs.push_back(b);
        } else if (key == "block") {
            Box3D b;
            s >> b.x0 >> b.x1 >> b.y0 >> b.y1 >> b.z0 >> b.z1;
            c.blocks.push_back(b);
        } else if (key == "fields") {
            int v = 0;
            s >> v;
            c.fields = v != 0;
        } else if (key == "dump") {
            plint v;
            while (s >> v) c.dumps.push_back(v);
        } else {
            std::fprintf(stderr, "unknown key %s\n", key.c_str());
            std::exit(2);
        }
    }
    return c;
}

template<typename T, template<typename U> class D>
void addVelocityPlane(OnLatticeBoundaryCondition3D<T,D>& bc,
                      MultiBlockLattice3D<T,D>& lat,
                      int dir, int ori, Box3D box, boundary::BcType t) {
    if (dir == 0 && ori < 0) bc.addVelocityBoundary0N(box, lat, t);
    else if (dir == 0) bc.addVelocityBoundary0P(box, lat, t);
    else if (dir == 1 && ori < 0) bc.addVelocityBoundary1N(box, lat, t);
    else if (dir == 1) bc.addVelocityBoundary1P(box, lat, t);
    else if (dir == 2 && ori < 0) bc.addVelocityBoundary2N(box, lat, t);
    else bc.addVelocityBoundary2P(box, lat, t);
}

template<typename T, template<typename U> class D>
void addPressurePlane(OnLatticeBoundaryCondition3D<T,D>& bc,
                      MultiBlockLattice3D<T,D>& lat,
                      int dir, int ori, Box3D box, boundary::BcType t) {
    if (dir == 0 && ori < 0) bc.addPressureBoundary0N(box, lat, t);
    else if (dir == 0) bc.addPressureBoundary0P(box, lat, t);
    else if (dir == 1 && ori < 0) bc.addPressureBoundary1N(box, lat, t);
    else if (dir == 1) bc.addPressureBoundary1P(box, lat, t);
    else if (dir == 2 && ori < 0) bc.addPressureBoundary2N(box, lat, t);
    else bc.addPressureBoundary2P(box, lat, t);
}

// addExternalVelocityEdge<plane><n1><n2> (boundaryCondition3D.h), n = -1 | +1
template<typename T, template<typename U> class D>
void addVelocityEdge(OnLatticeBoundaryCondition3D<T,D>& bc, MultiBlockLattice3D<T,D>& lat,
                     int plane, int n1, int n2, Box3D box, boundary::BcType t) {
    const int k = plane * 4 + (n1 > 0 ? 2 : 0) + (n2 > 0 ? 1 : 0);
    switch (k) {
    case 0: bc.addExternalVelocityEdge0NN(box, lat, t); break;
    case 1: bc.addExternalVelocityEdge0NP(box, lat, t); break;
    case 2: bc.addExternalVelocityEdge0PN(box, lat, t); break;
    case 3: bc.addExternalVelocityEdge0PP(box, lat, t); break;
    case 4: bc.addExternalVelocityEdge1NN(box, lat, t); break;
    case 5: bc.addExternalVelocityEdge1NP(box, lat, t); break;
    case 6: bc.addExternalVelocityEdge1PN(box, lat, t); break;
    case 7: bc.addExternalVelocityEdge1PP(box, lat, t); break;
    case 8: bc.addExternalVelocityEdge2NN(box, lat, t); break;
    case 9: bc.addExternalVelocityEdge2NP(box, lat, t); break;
    case 10: bc.addExternalVelocityEdge2PN(box, lat, t); break;
    default: bc.addExternalVelocityEdge2PP(box, lat, t); break;
    }
}

// addExternalVelocityCorner<xn><yn><zn>
template<typename T, template<typename U> class D>
void addVelocityCorner(OnLatticeBoundaryCondition3D<T,D>& bc, MultiBlockLattice3D<T,D>& lat,
                       const int n[3], plint x, plint y, plint z, boundary::BcType t) {
    const int k = (n[0] > 0 ? 4 : 0) + (n[1] > 0 ? 2 : 0) + (n[2] > 0 ? 1 : 0);
    switch (k) {
    case 0: bc.addExternalVelocityCornerNNN(x, y, z, lat, t); break;
    case 1: bc.addExternalVelocityCornerNNP(x, y, z, lat, t); break;
    case 2: bc.addExternalVelocityCornerNPN(x, y, z, lat, t); break;
    case 3: bc.addExternalVelocityCornerNPP(x, y, z, lat, t); break;
    case 4: bc.addExternalVelocityCornerPNN(x, y, z, lat, t); break;
    case 5: bc.addExternalVelocityCornerPNP(x, y, z, lat, t); break;
    case 6: bc.addExternalVelocityCornerPPN(x, y, z, lat, t); break;
    default: bc.addExternalVelocityCornerPPP(x, y, z, lat, t); break;
    }
}

template<typename T, template<typename U> class D>
int runCase(CaseSpec const& c) {
    typedef D<T> L;
    typedef MultiBlockLattice3D<T,D> Lattice;
    const int Q = 19;

    Dynamics<T,D>* bulk = 0;
    if (c.forced) bulk = new GuoExternalForceBGKdynamics<T,D>((T)c.omega);
    else if (c.dyn == "bgk") bulk = new BGKdynamics<T,D>((T)c.omega);
    else if (c.dyn == "rbgk") bulk = new RegularizedBGKdynamics<T,D>((T)c.omega);
    else { std::fprintf(stderr, "unknown dyn\n"); return 2; }

    Lattice* lattice = 0;
    if (c.blocks.empty()) {
        lattice = new Lattice(c.nx, c.ny, c.nz, bulk);
    } else {
        SparseBlockStructure3D sparse(Box3D(0, c.nx-1, 0, c.ny-1, 0, c.nz-1));
        for (pluint i = 0; i < c.blocks.size(); ++i) sparse.addBlock(c.blocks[i], (plint)i);
        MultiBlockManagement3D management(sparse, defaultMultiBlockPolicy3D().getThreadAttribution(), 1);
        lattice = new Lattice(management,
                              defaultMultiBlockPolicy3D().getBlockCommunicator(),
                              defaultMultiBlockPolicy3D().getCombinedStatistics(),
                              defaultMultiBlockPolicy3D().getMultiCellAccess<T,D>(),
                              bulk);
    }
    Lattice& lat = *lattice;
    lat.periodicity().toggle(0, c.periodic[0] != 0);
    lat.periodicity().toggle(1, c.periodic[1] != 0);
    lat.periodicity().toggle(2, c.periodic[2] != 0);

    // Guo body force, uniform (external scalars of ForcedD3Q19).
    if (c.forced) {
        Array<T,3> F((T)c.force[0], (T)c.force[1], (T)c.force[2]);
        setExternalVector(lat, lat.getBoundingBox(), D<T>::ExternalField::forceBeginsAt, F);
    }

    // Solid cells.
    Array<plint,3> forceIds;
    bool haveMomentumExchange = false;
    if (!c.tags.empty()) {
        std::vector<char> tags = readFile(c.tags);
        for (plint x = 0; x < c.nx; ++x) for (plint y = 0; y < c.ny; ++y) for (plint z = 0; z < c.nz; ++z)
            if (tags[(x*c.ny + y)*c.nz + z] == 3) haveMomentumExchange = true;
        if (haveMomentumExchange) {
            forceIds[0] = lat.internalStatSubscription().subscribeSum();
            forceIds[1] = lat.internalStatSubscription().subscribeSum();
            forceIds[2] = lat.internalStatSubscription().subscribeSum();
        }
        for (plint x = 0; x < c.nx; ++x) for (plint y = 0; y < c.ny; ++y) for (plint z = 0; z < c.nz; ++z) {
            char t = tags[(x*c.ny + y)*c.nz + z];
            Box3D cell(x, x, y, y, z, z);
            if (t == 1) defineDynamics(lat, cell, new BounceBack<T,D>((T)1));
            else if (t == 2) defineDynamics(lat, cell, new NoDynamics<T,D>((T)1));
            else if (t == 3) defineDynamics(lat, cell,
                     new MomentumExchangeBounceBack<T,D>(forceIds));
        }
        if (haveMomentumExchange) initializeMomentumExchange(lat, lat.getBoundingBox());
    }

    // Boundary planes.
    OnLatticeBoundaryCondition3D<T,D>* bc = 0;
    if (!c.bcs.empty() || !c.blockBcs.empty()) {
        if (c.bcfamily == "zouhe")
            bc = (c.bcstyle == "dynamics") ? createDynamicsBasedZouHeBoundaryCondition3D<T,D>()
                                           : createZouHeBoundaryCondition3D<T,D>();
        else
            bc = (c.bcstyle == "dynamics") ? createDynamicsBasedLocalBoundaryCondition3D<T,D>()
                                           : createLocalBoundaryCondition3D<T,D>();
        for (pluint i = 0; i < c.bcs.size(); ++i) {
            BcSpec const& b = c.bcs[i];
            const boundary::BcType t = (boundary::BcType)b.bcType;
            if (b.kind == "vel") addVelocityPlane<T,D>(*bc, lat, b.dir, b.ori, b.box, t);
            else if (b.kind == "press") addPressurePlane<T,D>(*bc, lat, b.dir, b.ori, b.box, t);
            else if (b.kind == "edge") addVelocityEdge<T,D>(*bc, lat, b.dir, b.normal[1], b.normal[2], b.box, t);
            else addVelocityCorner<T,D>(*bc, lat, b.normal, b.box.x0, b.box.y0, b.box.z0, t);
        }
        for (pluint i = 0; i < c.blockBcs.size(); ++i) {
            const boundary::BcType t = (boundary::BcType)c.blockBcs[i].bcType;
            if (c.blockBcs[i].kind == "vel") bc->setVelocityConditionOnBlockBoundaries(lat, t);
            else bc->setPressureConditionOnBlockBoundaries(lat, t);
        }
    }

    // Initial populations.
    {
        std::vector<char> raw = readFile(c.init);
        const double* f = reinterpret_cast<const double*>(&raw[0]);
        for (plint x = 0; x < c.nx; ++x) for (plint y = 0; y < c.ny; ++y) for (plint z = 0; z < c.nz; ++z) {
            Cell<T,D>& cell = lat.get(x, y, z);
            for (int i = 0; i < Q; ++i) cell[i] = (T)f[((x*c.ny + y)*c.nz + z)*Q + i];
        }
    }
    // Boundary values (after the composite dynamics exist).
    for (pluint i = 0; i < c.bcs.size(); ++i) {
        BcSpec const& b = c.bcs[i];
        std::vector<char> raw = readFile(b.file);
        const double* v = reinterpret_cast<const double*>(&raw[0]);
        plint k = 0;
        for (plint x = b.box.x0; x <= b.box.x1; ++x)
            for (plint y = b.box.y0; y <= b.box.y1; ++y)
                for (plint z = b.box.z0; z <= b.box.z1; ++z, ++k) {
                    Cell<T,D>& cell = lat.get(x, y, z);
                    if (b.kind != "press") {
                        Array<T,3> u((T)v[3*k], (T)v[3*k+1], (T)v[3*k+2]);
                        cell.defineVelocity(u);
                    } else {
                        cell.defineDensity((T)v[k]);
                    }
                }
    }
    // Boundary values of the block-boundary conditions: every cell of the outer
    // surface (defineVelocity / defineDensity leave other dynamics unchanged).
    for (pluint i = 0; i < c.blockBcs.size(); ++i) {
        std::vector<char> raw = readFile(c.blockBcs[i].file);
        const double* v = reinterpret_cast<const double*>(&raw[0]);
        for (plint x = 0; x < c.nx; ++x) for (plint y = 0; y < c.ny; ++y) for (plint z = 0; z < c.nz; ++z) {
            if (x != 0 && x != c.nx - 1 && y != 0 && y != c.ny - 1 && z != 0 && z != c.nz - 1) continue;
            const plint k = (x * c.ny + y) * c.nz + z;
            Cell<T,D>& cell = lat.get(x, y, z);
            if (c.blockBcs[i].kind == "vel") cell.defineVelocity(Array<T,3>((T)v[3*k], (T)v[3*k+1], (T)v[3*k+2]));
            else cell.defineDensity((T)v[k]);
        }
    }
    lat.initialize();

    std::string statsName = c.out + "_stats.txt";
    FILE* stats = std::fopen(statsName.c_str(), "w");
    plint maxStep = 0;
    for (pluint i = 0; i < c.dumps.size(); ++i) if (c.dumps[i] > maxStep) maxStep = c.dumps[i];
    std::vector<T> buf((size_t)(c.nx*c.ny*c.nz*Q));
    // the external-moment fields: the lattice's own block management, its periodicity
    MultiScalarField3D<T> rhoBarField((MultiBlock3D&)lat);
    MultiTensorField3D<T,3> jField((MultiBlock3D&)lat);
    for (int a = 0; a < 3; ++a) {
        rhoBarField.periodicity().toggle(a, c.periodic[a] != 0);
        jField.periodicity().toggle(a, c.periodic[a] != 0);
    }
    std::vector<MultiBlock3D*> rhoBarJargs;
    rhoBarJargs.push_back(&lat);
    rhoBarJargs.push_back(&rhoBarField);
    rhoBarJargs.push_back(&jField);
    for (plint step = 0; step <= maxStep; ++step) {
        if (step > 0 && c.external) {
            applyProcessingFunctional(new BoxRhoBarJfunctional3D<T,D>(), lat.getBoundingBox(), rhoBarJargs);
            for (plint x = 0; x < c.nx; ++x) for (plint y = 0; y < c.ny; ++y) for (plint z = 0; z < c.nz; ++z)
                jField.get(x, y, z)[0] += (T)c.externalA * ((T)1 + rhoBarField.get(x, y, z));
            jField.duplicateOverlaps(modif::staticVariables);
            applyProcessingFunctional(new ExternalRhoJcollideAndStream3D<T,D>(), lat.getBoundingBox(), rhoBarJargs);
        } else if (step > 0) {
            lat.collideAndStream();
        }
        bool dumpNow = false;
        for (pluint i = 0; i < c.dumps.size(); ++i) if (c.dumps[i] == step) dumpNow = true;
        if (!dumpNow) continue;
        for (plint x = 0; x < c.nx; ++x) for (plint y = 0; y < c.ny; ++y) for (plint z = 0; z < c.nz; ++z) {
            Cell<T,D>& cell = lat.get(x, y, z);
            for (int i = 0; i < Q; ++i) buf[((x*c.ny + y)*c.nz + z)*Q + i] = cell[i];
        }
        char name[512];
        std::snprintf(name, sizeof(name), "%s_s%ld.bin", c.out.c_str(), (long)step);
        FILE* o = std::fopen(name, "wb");
        std::fwrite(&buf[0], sizeof(T), buf.size(), o);
        std::fclose(o);
        if (c.external && step > 0) {
            std::vector<T> fb((size_t)(c.nx*c.ny*c.nz*4));
            size_t k = 0;
            for (plint x = 0; x < c.nx; ++x) for (plint y = 0; y < c.ny; ++y) for (plint z = 0; z < c.nz; ++z) {
                fb[k++] = rhoBarField.get(x, y, z);
                for (int a = 0; a < 3; ++a) fb[k++] = jField.get(x, y, z)[a];
            }
            std::snprintf(name, sizeof(name), "%s_rhobarj_s%ld.bin", c.out.c_str(), (long)step);
            FILE* fo = std::fopen(name, "wb");
            std::fwrite(&fb[0], sizeof(T), fb.size(), fo);
            std::fclose(fo);
        }
        if (c.fields) {
            std::auto_ptr<MultiScalarField3D<T> > rho = computeDensity(lat);
            std::auto_ptr<MultiScalarField3D<T> > rhoBar = computeRhoBar(lat);
            std::auto_ptr<MultiTensorField3D<T,3> > u = computeVelocity(lat);
            std::auto_ptr<MultiScalarField3D<T> > energy = computeKineticEnergy(lat);
            std::auto_ptr<MultiScalarField3D<T> > norm = computeVelocityNorm(lat);
            std::vector<T> fb((size_t)(c.nx*c.ny*c.nz*7));
            size_t k = 0;
            for (plint x = 0; x < c.nx; ++x) for (plint y = 0; y < c.ny; ++y) for (plint z = 0; z < c.nz; ++z) {
                fb[k++] = rho->get(x, y, z);
                fb[k++] = rhoBar->get(x, y, z);
                for (int a = 0; a < 3; ++a) fb[k++] = u->get(x, y, z)[a];
                fb[k++] = energy->get(x, y, z);
                fb[k++] = norm->get(x, y, z);
            }
            std::snprintf(name, sizeof(name), "%s_fields_s%ld.bin", c.out.c_str(), (long)step);
            FILE* fo = std::fopen(name, "wb");
            std::fwrite(&fb[0], sizeof(T), fb.size(), fo);
            std::fclose(fo);
        }
        double fx = 0, fy = 0, fz = 0;
        if (haveMomentumExchange && step > 0) {
            fx = lat.getInternalStatistics().getSum(forceIds[0]);
            fy = lat.getInternalStatistics().getSum(forceIds[1]);
            fz = lat.getInternalStatistics().getSum(forceIds[2]);
        }
        double sRho = step > 0 ? getStoredAverageDensity<T>(lat) : 0.0;
        double sEn = step > 0 ? getStoredAverageEnergy<T>(lat) : 0.0;
        std::fprintf(stats, "%ld %.17g %.17g %.17g %.17g %.17g %.17g %.17g\n", (long)step,
                     sRho, sEn, (double)computeAverageDensity(lat), (double)computeAverageEnergy(lat),
                     fx, fy, fz);
    }
    std::fclose(stats);
    delete bc;
    delete lattice;
    return 0;
}

template<typename T>
int runBench(plint n, plint steps) {
    typedef MultiBlockLattice3D<T,descriptors::D3Q19Descriptor> Lattice;
    Lattice lat(n, n, n, new BGKdynamics<T,descriptors::D3Q19Descriptor>((T)1.8));
    lat.periodicity().toggleAll(true);
    lat.toggleInternalStatistics(false);
    Array<T,3> u((T)0.01, (T)0.02, (T)0.0);
    initializeAtEquilibrium(lat, lat.getBoundingBox(), (T)1, u);
    lat.initialize();
    lat.collideAndStream();  // warm-up
    timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    for (plint s = 0; s < steps; ++s) lat.collideAndStream();
    clock_gettime(CLOCK_MONOTONIC, &t1);
    double sec = (t1.tv_sec - t0.tv_sec) + 1e-9 * (t1.tv_nsec - t0.tv_nsec);
    double mlups = (double)n*n*n*steps / sec / 1e6;
    std::printf("{\"mlups\": %.6f, \"seconds\": %.6f, \"n\": %ld, \"steps\": %ld, \"T\": \"%s\"}\n",
                mlups, sec, (long)n, (long)steps, sizeof(T) == 8 ? "double" : "float");
    return 0;
}

template<typename T, template<typename U> class Desc>
int runCheckpoint(const std::string& mode, plint nx, plint ny, plint nz, const std::string& a, const std::string& b,
                  bool enforceUint) {
    typedef MultiBlockLattice3D<T,Desc> Lattice;
    const plint numPop = Desc<T>::numPop, numExt = Desc<T>::ExternalField::numScalars;
    const plint rec = numPop + numExt;
    Lattice lat(nx, ny, nz, new BGKdynamics<T,Desc>((T)1));
    std::vector<T> records((size_t)(nx * ny * nz * rec));
    if (mode == "save") {
        std::ifstream in(a.c_str(), std::ios::binary);
        in.read(reinterpret_cast<char*>(records.data()), (std::streamsize)(records.size() * sizeof(T)));
        if (!in) { std::fprintf(stderr, "cannot read %s\n", a.c_str()); return 1; }
        size_t k = 0;
        for (plint x = 0; x < nx; ++x)
            for (plint y = 0; y < ny; ++y)
                for (plint z = 0; z < nz; ++z, k += (size_t)rec) {
                    Cell<T,Desc>& cell = lat.get(x, y, z);
                    for (plint i = 0; i < numPop; ++i) cell[i] = records[k + i];
                    for (plint e = 0; e < numExt; ++e) *cell.getExternal(e) = records[k + numPop + e];
                }
        saveBinaryBlock(lat, b, enforceUint);
    } else {
        loadBinaryBlock(lat, a, enforceUint);
        size_t k = 0;
        for (plint x = 0; x < nx; ++x)
            for (plint y = 0; y < ny; ++y)
                for (plint z = 0; z < nz; ++z, k += (size_t)rec) {
                    Cell<T,Desc>& cell = lat.get(x, y, z);
                    for (plint i = 0; i < numPop; ++i) records[k + i] = cell[i];
                    for (plint e = 0; e < numExt; ++e) records[k + numPop + e] = *cell.getExternal(e);
                }
        std::ofstream out(b.c_str(), std::ios::binary);
        out.write(reinterpret_cast<const char*>(records.data()), (std::streamsize)(records.size() * sizeof(T)));
    }
    return 0;
}

}  // namespace

int runUnits(char* argv[]) {
    const double physU = std::atof(argv[2]), latU = std::atof(argv[3]), Re = std::atof(argv[4]);
    const double physLen = std::atof(argv[5]);
    const plint N = std::atol(argv[6]);
    const double lx = std::atof(argv[7]), ly = std::atof(argv[8]), lz = std::atof(argv[9]), t = std::atof(argv[10]);
    IncomprFlowParam<double> p = physU != 0 ? IncomprFlowParam<double>(physU, latU, Re, physLen, N, lx, ly, lz)
                                            : IncomprFlowParam<double>(latU, Re, N, lx, ly, lz);
    std::printf("%ld %ld %ld %ld %ld %ld %.17g %.17g %.17g %.17g %.17g %ld %.17g %.17g %.17g %.17g %.17g\n",
                (long)p.getNx(), (long)p.getNy(), (long)p.getNz(), (long)p.getNx(true), (long)p.getNy(true),
                (long)p.getNz(true), p.getLatticeNu(), p.getTau(), p.getOmega(), p.getDeltaX(), p.getDeltaT(),
                (long)p.nStep(t), p.getLx(), p.getLy(), p.getLz(), p.getPhysicalU(), p.getPhysicalLength());
    return 0;
}

int main(int argc, char* argv[]) {
    plbInit(&argc, &argv);
    if (argc >= 2 && std::string(argv[1]) == "--units") {
        if (argc < 11) { std::fprintf(stderr, "usage: --units physU latU Re physLen N lx ly lz t\n"); return 2; }
        return runUnits(argv);
    }
    if (argc >= 2 && std::string(argv[1]) == "--checkpoint") {
        if (argc < 11) {
            std::fprintf(stderr, "usage: --checkpoint save|load double|float plain|forced nx ny nz in out uint|size\n");
            return 2;
        }
        const std::string mode = argv[2], T = argv[3], desc = argv[4];
        const plint nx = std::atol(argv[5]), ny = std::atol(argv[6]), nz = std::atol(argv[7]);
        const bool enforceUint = std::string(argv[10]) == "uint";
        if (desc == "forced")
            return T == "float" ? runCheckpoint<float, descriptors::ForcedD3Q19Descriptor>(mode, nx, ny, nz, argv[8], argv[9], enforceUint)
                                : runCheckpoint<double, descriptors::ForcedD3Q19Descriptor>(mode, nx, ny, nz, argv[8], argv[9], enforceUint);
        return T == "float" ? runCheckpoint<float, descriptors::D3Q19Descriptor>(mode, nx, ny, nz, argv[8], argv[9], enforceUint)
                            : runCheckpoint<double, descriptors::D3Q19Descriptor>(mode, nx, ny, nz, argv[8], argv[9], enforceUint);
    }
    if (argc >= 2 && std::string(argv[1]) == "--bench") {
        if (argc < 5) { std::fprintf(stderr, "usage: --bench double|float n steps\n"); return 2; }
        plint n = std::atol(argv[3]), steps = std::atol(argv[4]);
        return std::string(argv[2]) == "float" ? runBench<float>(n, steps) : runBench<double>(n, steps);
    }
    if (argc < 2) { std::fprintf(stderr, "usage: ref_driver case.txt | --bench T n steps\n"); return 2; }
    CaseSpec c = parseCase(argv[1]);
    if (c.forced) {
        if (!c.bcs.empty() || c.dyn != "bgk") { std::fprintf(stderr, "force: bgk without planes only\n"); return 2; }
        return c.T == "float" ? runCase<float, descriptors::ForcedD3Q19Descriptor>(c)
                              : runCase<double, descriptors::ForcedD3Q19Descriptor>(c);
    }
    return c.T == "float" ? runCase<float, descriptors::D3Q19Descriptor>(c)
                          : runCase<double, descriptors::D3Q19Descriptor>(c);
}
