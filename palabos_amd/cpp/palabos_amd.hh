// Palabos-compatible C++ API for the D3Q19 collide-and-stream path, running on
// the MI355X engine through the C ABI of include/palabos_amd.h.
//
// A Palabos program that builds a D3Q19 lattice with BGK or regularized BGK
// dynamics, BounceBack / NoDynamics / MomentumExchangeBounceBack nodes,
// Zou-He or regularized velocity/pressure planes and a uniform Guo force, and
// steps it with collideAndStream(), compiles against this header instead of
// palabos3D.h/.hh with the same class and function names:
//
//   MultiBlockLattice3D<T,DESCRIPTOR> lattice(nx, ny, nz, new BGKdynamics<T,DESCRIPTOR>(omega));
//   OnLatticeBoundaryCondition3D<T,DESCRIPTOR>* bc = createZouHeBoundaryCondition3D<T,DESCRIPTOR>();
//   bc->addVelocityBoundary2P(Box3D(1, nx-2, 1, ny-2, nz-1, nz-1), lattice);
//   setBoundaryVelocity(lattice, Box3D(1, nx-2, 1, ny-2, nz-1, nz-1), Array<T,3>(u, 0, 0));
//   initializeAtEquilibrium(lattice, lattice.getBoundingBox(), (T)1, Array<T,3>(0, 0, 0));
//   lattice.initialize();
//   for (plint i = 0; i < n; ++i) lattice.collideAndStream();
//   pcout << computeAverageEnergy(lattice) << getStoredAverageDensity<T>(lattice);
//
// Populations live on the device. lattice.get(x,y,z) gives a cell whose
// populations are read from / written to a host mirror that is downloaded on
// demand and uploaded before the next step, as Palabos's cell access is.
//
// Differences from Palabos, by design of the accelerated path:
//  * one process-wide device; all cells of a lattice live in one device block
//  * boundary planes are the face interiors given to add*Boundary; edge and
//    corner cells keep the same-step edge reflection (Palabos's edge/corner
//    finite-difference processors are not part of this path)
//  * both the wrapped and the dynamics-based factories run the completion inside
//    the collision; for the wrapped factories the device completes the boundary
//    populations of every downloaded snapshot, as Palabos's post-stream
//    processor leaves them (the next step is unaffected: completion is idempotent)
//  * all MomentumExchangeBounceBack cells report into one force (the first three
//    subscribed sums)
#pragma once

#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <iostream>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "palabos_amd.h"

namespace plb {

typedef long plint;
typedef unsigned long pluint;

class PlbException : public std::runtime_error {
public:
    explicit PlbException(std::string const& m) : std::runtime_error(m) {}
};

namespace detail {
inline void check(int rc, const char* what) {
    if (rc != PLBAMD_OK) throw PlbException(std::string(what) + ": " + plbamd_last_error());
}
inline void ensureInit() {
    static bool done = false;
    if (!done) {
        check(plbamd_init(0), "plbamd_init");
        done = true;
    }
}
}  // namespace detail

// core/geometry3D.h
struct Box3D {
    plint x0, x1, y0, y1, z0, z1;
    Box3D() : x0(0), x1(0), y0(0), y1(0), z0(0), z1(0) {}
    Box3D(plint x0_, plint x1_, plint y0_, plint y1_, plint z0_, plint z1_)
        : x0(x0_), x1(x1_), y0(y0_), y1(y1_), z0(z0_), z1(z1_) {}
    plint getNx() const { return x1 - x0 + 1; }
    plint getNy() const { return y1 - y0 + 1; }
    plint getNz() const { return z1 - z0 + 1; }
    plint nCells() const { return getNx() * getNy() * getNz(); }
    bool contains(plint x, plint y, plint z) const {
        return x >= x0 && x <= x1 && y >= y0 && y <= y1 && z >= z0 && z <= z1;
    }
    plbamd_box toC() const {
        plbamd_box b;
        b.x0 = (int32_t)x0; b.x1 = (int32_t)x1; b.y0 = (int32_t)y0;
        b.y1 = (int32_t)y1; b.z0 = (int32_t)z0; b.z1 = (int32_t)z1;
        return b;
    }
};

// boundary::BcType (boundaryCondition/boundaryCondition.h:40); the integer
// values are the C ABI's PLBAMD_BC_* constants.
namespace boundary {
typedef enum { dirichlet, neumann, freeslip, density, outflow, normalOutflow } BcType;
}

inline bool intersect(Box3D const& a, Box3D const& b, Box3D& out) {
    out = Box3D(std::max(a.x0, b.x0), std::min(a.x1, b.x1), std::max(a.y0, b.y0), std::min(a.y1, b.y1),
                std::max(a.z0, b.z0), std::min(a.z1, b.z1));
    return out.x0 <= out.x1 && out.y0 <= out.y1 && out.z0 <= out.z1;
}

// core/array.h
template<typename T, int N>
class Array {
public:
    Array() { for (int i = 0; i < N; ++i) d[i] = T(); }
    Array(T a, T b) { static_assert(N == 2, "2 components"); d[0] = a; d[1] = b; }
    Array(T a, T b, T c) { static_assert(N == 3, "3 components"); d[0] = a; d[1] = b; d[2] = c; }
    T& operator[](int i) { return d[i]; }
    T const& operator[](int i) const { return d[i]; }
    void resetToZero() { for (int i = 0; i < N; ++i) d[i] = T(); }
private:
    T d[N];
};

namespace descriptors {
// latticeBoltzmann/nearestNeighborLattices3D.hh:159-211
template<typename T>
struct D3Q19Descriptor {
    enum { d = 3, q = 19, numPop = 19 };
    static constexpr int vicinity = 1;
    static int c(int i, int a) {
        static const int v[19][3] = {{0, 0, 0},  {-1, 0, 0}, {0, -1, 0}, {0, 0, -1}, {-1, -1, 0},
                                     {-1, 1, 0}, {-1, 0, -1}, {-1, 0, 1}, {0, -1, -1}, {0, -1, 1},
                                     {1, 0, 0},  {0, 1, 0},  {0, 0, 1},  {1, 1, 0},  {1, -1, 0},
                                     {1, 0, 1},  {1, 0, -1}, {0, 1, 1},  {0, 1, -1}};
        return v[i][a];
    }
    static T t(int i) {
        if (i == 0) return (T)1 / (T)3;
        int n2 = c(i, 0) * c(i, 0) + c(i, 1) * c(i, 1) + c(i, 2) * c(i, 2);
        return n2 == 1 ? (T)1 / (T)18 : (T)1 / (T)36;
    }
    static T cs2() { return (T)1 / (T)3; }
    static T rhoBar(T rho) { return rho - (T)1; }
    static T fullRho(T rhoBar) { return rhoBar + (T)1; }
    static T invRho(T rhoBar) { return (T)1 / (rhoBar + (T)1); }
    static constexpr bool forced = false;
};
template<typename T>
struct ForcedD3Q19Descriptor : public D3Q19Descriptor<T> {
    struct ExternalField { enum { forceBeginsAt = 0, sizeOfForce = 3 }; };
    static constexpr bool forced = true;
};
}  // namespace descriptors

template<typename T> struct StorageOf;
template<> struct StorageOf<double> { enum { code = PLBAMD_F64 }; };
template<> struct StorageOf<float> { enum { code = PLBAMD_F32 }; };

// ---------------------------------------------------------------------------
// Dynamics (core/dynamics.h:64-270). The engine runs the arithmetic; these
// objects carry the parameters and identify the cell type.

enum class DynKind { BGK, RBGK, GuoBGK, BounceBack, NoDynamics, MomentumExchange, WallLink };

template<typename T, template<typename U> class Descriptor>
class Dynamics {
public:
    virtual ~Dynamics() {}
    virtual Dynamics* clone() const = 0;
    virtual DynKind kind() const = 0;
    virtual T getOmega() const { return T(); }
    virtual void setOmega(T) {}
    virtual int getId() const { return (int)kind(); }
    // density parameter of the solid dynamics (reported by the analysis functions)
    virtual T getDensityParameter() const { return (T)1; }
    // bgk_ma2_equilibrium (dynamicsTemplates3D.h:1994-1997)
    virtual T computeEquilibrium(plint iPop, T rhoBar, Array<T, 3> const& j, T jSqr, T = T()) const {
        typedef Descriptor<T> D;
        T invRho = D::invRho(rhoBar);
        T c_j = D::c(iPop, 0) * j[0] + D::c(iPop, 1) * j[1] + D::c(iPop, 2) * j[2];
        return D::t(iPop) * (rhoBar + (T)3. * c_j + invRho * ((T)4.5 * c_j * c_j - (T)1.5 * jSqr));
    }
};

template<typename T, template<typename U> class Descriptor>
class BasicBulkDynamics : public Dynamics<T, Descriptor> {
public:
    explicit BasicBulkDynamics(T omega_) : omega(omega_) {}
    T getOmega() const override { return omega; }
    void setOmega(T o) override { omega = o; }
protected:
    T omega;
};

template<typename T, template<typename U> class Descriptor>
class BGKdynamics : public BasicBulkDynamics<T, Descriptor> {
public:
    explicit BGKdynamics(T omega_) : BasicBulkDynamics<T, Descriptor>(omega_) {}
    BGKdynamics* clone() const override { return new BGKdynamics(*this); }
    DynKind kind() const override { return DynKind::BGK; }
};

template<typename T, template<typename U> class Descriptor>
class RegularizedBGKdynamics : public BasicBulkDynamics<T, Descriptor> {
public:
    explicit RegularizedBGKdynamics(T omega_) : BasicBulkDynamics<T, Descriptor>(omega_) {}
    RegularizedBGKdynamics* clone() const override { return new RegularizedBGKdynamics(*this); }
    DynKind kind() const override { return DynKind::RBGK; }
};

template<typename T, template<typename U> class Descriptor>
class GuoExternalForceBGKdynamics : public BasicBulkDynamics<T, Descriptor> {
public:
    explicit GuoExternalForceBGKdynamics(T omega_) : BasicBulkDynamics<T, Descriptor>(omega_) {}
    GuoExternalForceBGKdynamics* clone() const override { return new GuoExternalForceBGKdynamics(*this); }
    DynKind kind() const override { return DynKind::GuoBGK; }
};

template<typename T, template<typename U> class Descriptor>
class BounceBack : public Dynamics<T, Descriptor> {
public:
    // core/dynamics.h:574: rho defaults to T()
    explicit BounceBack(T rho_ = T()) : rho(rho_) {}
    BounceBack* clone() const override { return new BounceBack(*this); }
    DynKind kind() const override { return DynKind::BounceBack; }
    T getDensityParameter() const override { return rho; }
private:
    T rho;
};

template<typename T, template<typename U> class Descriptor>
class NoDynamics : public Dynamics<T, Descriptor> {
public:
    // core/dynamics.h:687
    explicit NoDynamics(T rho_ = T()) : rho(rho_) {}
    NoDynamics* clone() const override { return new NoDynamics(*this); }
    DynKind kind() const override { return DynKind::NoDynamics; }
    T getDensityParameter() const override { return rho; }
private:
    T rho;
};

template<typename T, template<typename U> class Descriptor>
class MomentumExchangeBounceBack : public Dynamics<T, Descriptor> {
public:
    // boundaryCondition/bounceBackModels.h:62
    explicit MomentumExchangeBounceBack(Array<plint, 3> forceIds_, T rho_ = T()) : forceIds(forceIds_), rho(rho_) {}
    MomentumExchangeBounceBack* clone() const override { return new MomentumExchangeBounceBack(*this); }
    DynKind kind() const override { return DynKind::MomentumExchange; }
    T getDensityParameter() const override { return rho; }
private:
    Array<plint, 3> forceIds;
    T rho;
};

// Link-wise half-way bounce-back solid (same-step reflection, no Palabos class).
template<typename T, template<typename U> class Descriptor>
class HalfwayBounceBackWall : public Dynamics<T, Descriptor> {
public:
    HalfwayBounceBackWall* clone() const override { return new HalfwayBounceBackWall(*this); }
    DynKind kind() const override { return DynKind::WallLink; }
};

// ---------------------------------------------------------------------------
// Statistics (core/blockStatistics.h:70-120, core/latticeStatistics.h:38-43)

namespace LatticeStatistics {
enum { avRhoBar = 0, avUSqr = 1, maxUSqr = 0 };
}

class BlockStatistics {
public:
    double getAverage(plint which) const { return which == LatticeStatistics::avRhoBar ? s[PLBAMD_STAT_AVG_RHOBAR] : s[PLBAMD_STAT_AVG_USQR]; }
    double getMax(plint) const { return s[PLBAMD_STAT_MAX_USQR]; }
    double getSum(plint which) const { return (which >= 0 && which < 3) ? s[PLBAMD_STAT_FORCE_X + which] : 0.0; }
    double getPorousDrag(int a) const { return s[PLBAMD_STAT_POROUS_X + a]; }
    double s[PLBAMD_NUM_STATS] = {};
};

class StatSubscriber {
public:
    plint subscribeSum() { return next++; }
    plint subscribeAverage() { return next++; }
private:
    plint next = 0;
};

struct PeriodicitySwitch3D {
    bool flags[3] = {false, false, false};
    void toggle(plint dir, bool p) { flags[dir] = p; }
    void toggleAll(bool p) { flags[0] = flags[1] = flags[2] = p; }
    bool get(plint dir) const { return flags[dir]; }
};

template<typename T, template<typename U> class Descriptor> class MultiBlockLattice3D;

// A cell of the host mirror (core/cell.h:91-377, population access only).
template<typename T, template<typename U> class Descriptor>
class Cell {
public:
    Cell(T* f_, MultiBlockLattice3D<T, Descriptor>* lat_, plint x_, plint y_, plint z_)
        : f(f_), lat(lat_), x(x_), y(y_), z(z_) {}
    T& operator[](plint i) { return f[i]; }
    T const& operator[](plint i) const { return f[i]; }
    T* getRawPopulations() { return f; }
    // Density and velocity as the cell's dynamics reports them, computed on the
    // device (Dynamics::computeDensity / computeVelocity).
    T computeDensity() const {
        double rho, u[3];
        lat->cellMacroscopic(x, y, z, rho, u);
        return (T)rho;
    }
    void computeVelocity(Array<T, 3>& u) const {
        double rho, v[3];
        lat->cellMacroscopic(x, y, z, rho, v);
        for (int a = 0; a < 3; ++a) u[a] = (T)v[a];
    }
    void defineVelocity(Array<T, 3> const& u);
    void defineDensity(T rho);
private:
    T* f;
    MultiBlockLattice3D<T, Descriptor>* lat;
    plint x, y, z;
};

// ---------------------------------------------------------------------------
// The lattice (multiBlock/multiBlockLattice3D.h:72-146). One device block.

template<typename T, template<typename U> class Descriptor>
class MultiBlockLattice3D {
public:
    MultiBlockLattice3D(plint nx_, plint ny_, plint nz_, Dynamics<T, Descriptor>* background_)
        : nx(nx_), ny(ny_), nz(nz_), background(background_) {
        if (!background) throw PlbException("MultiBlockLattice3D: null background dynamics");
        DynKind k = background->kind();
        if (k != DynKind::BGK && k != DynKind::RBGK && k != DynKind::GuoBGK)
            throw PlbException("MultiBlockLattice3D: background dynamics must be BGK, regularized BGK or Guo BGK");
        tags.assign((size_t)(nx * ny * nz), (uint8_t)PLBAMD_TAG_BULK);
    }
    virtual ~MultiBlockLattice3D() {
        if (handle > 0) plbamd_lattice_destroy(handle);
        delete background;
    }
    MultiBlockLattice3D(MultiBlockLattice3D const&) = delete;
    MultiBlockLattice3D& operator=(MultiBlockLattice3D const&) = delete;

    plint getNx() const { return nx; }
    plint getNy() const { return ny; }
    plint getNz() const { return nz; }
    Box3D getBoundingBox() const { return Box3D(0, nx - 1, 0, ny - 1, 0, nz - 1); }
    Dynamics<T, Descriptor> const& getBackgroundDynamics() const { return *background; }

    PeriodicitySwitch3D& periodicity() {
        if (handle > 0) throw PlbException("periodicity must be set before the lattice is first used");
        return periodic;
    }
    void toggleInternalStatistics(bool on) {
        statsOn = on;
        if (handle > 0) detail::check(plbamd_toggle_statistics(handle, on ? 1 : 0), "toggle_statistics");
    }
    StatSubscriber& internalStatSubscription() { return subscriber; }
    BlockStatistics const& getInternalStatistics() const { return stats; }

    Cell<T, Descriptor> get(plint x, plint y, plint z) {
        syncHost();
        hostDirty = true;
        return Cell<T, Descriptor>(&mirror[index(x, y, z) * 19], this, x, y, z);
    }

    // MultiBlock3D::initialize (multiBlock3D.cpp:330-332): one pass of the data
    // processors (Neumann-type copies, edge and corner closures); no envelope to
    // fill on one block.
    void initialize() {
        materialize();
        pushHost();
        detail::check(plbamd_execute_processors(handle), "execute_processors");
        hostValid = false;
    }

    // MultiBlockLattice3D::collideAndStream (multiBlockLattice3D.hh:318-360).
    void collideAndStream() { collideAndStream(1); }
    void collideAndStream(plint steps) {
        materialize();
        pushHost();
        detail::check(plbamd_collide_and_stream(handle, (int)steps), "collide_and_stream");
        hostValid = false;
        if (statsOn) detail::check(plbamd_get_statistics(handle, stats.s), "get_statistics");
    }
    double collideAndStreamTimed(plint steps) {
        materialize();
        pushHost();
        double ms = 0;
        detail::check(plbamd_collide_and_stream_timed(handle, (int)steps, &ms), "collide_and_stream_timed");
        hostValid = false;
        if (statsOn) detail::check(plbamd_get_statistics(handle, stats.s), "get_statistics");
        return ms;
    }
    plint getTime() {
        if (handle <= 0) return 0;
        int64_t t = 0;
        detail::check(plbamd_get_time(handle, &t), "get_time");
        return (plint)t;
    }

    // -- used by the free functions below ------------------------------------
    void setTag(Box3D const& b, uint8_t tag) {
        for (plint x = b.x0; x <= b.x1; ++x)
            for (plint y = b.y0; y <= b.y1; ++y)
                for (plint z = b.z0; z <= b.z1; ++z) tags[index(x, y, z)] = tag;
        tagsDirty = true;
    }
    int addPlane(int family, int kind, int dir, int ori, Box3D const& box, bool wrapped,
                 boundary::BcType bcType = boundary::dirichlet) {
        Plane p;
        p.family = family; p.kind = kind; p.dir = dir; p.ori = ori; p.box = box; p.wrapped = wrapped;
        p.values.assign((size_t)box.nCells() * (kind == PLBAMD_VELOCITY ? 3 : 1), kind == PLBAMD_VELOCITY ? 0.0 : 1.0);
        p.id = -1;
        p.bcType = (int)bcType;
        planes.push_back(p);
        planesDirty = true;
        return (int)planes.size() - 1;
    }
    // An external velocity edge along `plane` (normals along (plane+1)%3 and
    // (plane+2)%3) or corner (normals along x, y, z): StoreVelocityDynamics and the
    // finite-difference processor (boundaryInstantiator3D.h:1113-1232).
    int addEdge(int plane, int normal1, int normal2, Box3D const& box, boundary::BcType bcType) {
        Plane p;
        p.family = PLBAMD_REGULARIZED; p.kind = PLBAMD_VELOCITY; p.dir = plane; p.ori = 0; p.box = box;
        p.wrapped = false;
        p.shape = 1;
        p.normal[0] = normal1; p.normal[1] = normal2;
        p.values.assign((size_t)box.nCells() * 3, 0.0);
        p.id = -1;
        p.bcType = (int)bcType;
        planes.push_back(p);
        planesDirty = true;
        return (int)planes.size() - 1;
    }
    int addCorner(int xNormal, int yNormal, int zNormal, plint x, plint y, plint z, boundary::BcType bcType) {
        Plane p;
        p.family = PLBAMD_REGULARIZED; p.kind = PLBAMD_VELOCITY; p.dir = 0; p.ori = 0; p.box = Box3D(x, x, y, y, z, z);
        p.wrapped = false;
        p.shape = 2;
        p.normal[0] = xNormal; p.normal[1] = yNormal; p.normal[2] = zNormal;
        p.values.assign(3, 0.0);
        p.id = -1;
        p.bcType = (int)bcType;
        planes.push_back(p);
        planesDirty = true;
        return (int)planes.size() - 1;
    }
    void setPlaneValue(plint x, plint y, plint z, int kind, const double* v) {
        for (auto& p : planes) {
            if (p.kind != kind || !p.box.contains(x, y, z)) continue;
            plint k = ((x - p.box.x0) * p.box.getNy() + (y - p.box.y0)) * p.box.getNz() + (z - p.box.z0);
            int nv = kind == PLBAMD_VELOCITY ? 3 : 1;
            for (int a = 0; a < nv; ++a) p.values[(size_t)(k * nv + a)] = v[a];
            p.valuesDirty = true;
            planesDirty = true;
        }
    }
    // One density parameter per solid kind (the engine reports it for every
    // cell of that kind); two different ones for one kind are refused.
    void setSolidDensity(int tag, double rho) {
        const int k = tag - PLBAMD_TAG_BOUNCE_BACK;
        if (solidRhoSet[k] && solidRho[k] != rho)
            throw PlbException("defineDynamics: one density parameter per solid dynamics kind");
        solidRho[k] = rho;
        solidRhoSet[k] = true;
        tagsDirty = true;
    }
    void setForce(Array<T, 3> const& F) {
        if (handle > 0) throw PlbException("the body force must be set before the lattice is first used");
        force[0] = F[0]; force[1] = F[1]; force[2] = F[2];
    }
    void equilibrium(Box3D const& b, std::vector<double> const& rho, std::vector<double> const& u) {
        materialize();
        pushHost();
        plbamd_box cb = b.toC();
        detail::check(plbamd_initialize_at_equilibrium(handle, &cb, rho.data(), u.data()), "initialize_at_equilibrium");
        hostValid = false;
    }
    // Checkpoint files of this lattice's type (io/serializerIO_3D.cpp:37-66).
    void saveBinary(std::string const& fName, bool enforceUint) {
        materialize();
        pushHost();
        detail::check(plbamd_save_binary_block(handle, fName.c_str(), StorageOf<T>::code, Descriptor<T>::forced ? 1 : 0,
                                               enforceUint ? 1 : 0),
                      "save_binary_block");
    }
    void loadBinary(std::string const& fName, bool enforceUint) {
        materialize();
        pushHost();
        detail::check(plbamd_load_binary_block(handle, fName.c_str(), StorageOf<T>::code, Descriptor<T>::forced ? 1 : 0,
                                               enforceUint ? 1 : 0),
                      "load_binary_block");
        hostValid = false;
    }
    // Analysis fields of a box: [cell][7] (density, rhoBar, u, energy, |u|).
    void analysisFields(Box3D const& b, std::vector<double>& out) {
        materialize();
        pushHost();
        out.resize((size_t)b.nCells() * 7);
        plbamd_box cb = b.toC();
        detail::check(plbamd_compute_fields(handle, &cb, out.data()), "compute_fields");
    }
    void moments(Box3D const& b, double* sums) {
        materialize();
        pushHost();
        plbamd_box cb = b.toC();
        detail::check(plbamd_reduce_moments(handle, &cb, sums), "reduce_moments");
    }
    void cellMacroscopic(plint x, plint y, plint z, double& rho, double u[3]) {
        materialize();
        pushHost();
        plbamd_box cb = Box3D(x, x, y, y, z, z).toC();
        detail::check(plbamd_compute_macroscopic(handle, &cb, &rho, u), "compute_macroscopic");
    }
    int nativeHandle() { materialize(); pushHost(); return handle; }

private:
    struct Plane {
        int family, kind, dir, ori, id;
        bool wrapped;
        Box3D box;
        std::vector<double> values;
        bool valuesDirty = true;
        int shape = 0;                 // 0 plane, 1 edge, 2 corner
        int normal[3] = {0, 0, 0};     // edge: normal1, normal2; corner: x, y, z normals
        int bcType = 0;
    };
    friend class Cell<T, Descriptor>;

    size_t index(plint x, plint y, plint z) const { return (size_t)((x * ny + y) * nz + z); }

    void materialize() {
        if (handle <= 0) {
            detail::ensureInit();
            plbamd_lattice_desc d{};
            d.nx = (int32_t)nx; d.ny = (int32_t)ny; d.nz = (int32_t)nz;
            d.storage = StorageOf<T>::code;
            d.dynamics = background->kind() == DynKind::RBGK ? PLBAMD_RBGK : PLBAMD_BGK;
            for (int a = 0; a < 3; ++a) d.periodic[a] = periodic.flags[a] ? 1 : 0;
            d.omega = (double)background->getOmega();
            for (int a = 0; a < 3; ++a) d.force[a] = force[a];
            int h = 0;
            detail::check(plbamd_lattice_create(&d, &h), "lattice_create");
            handle = h;
            detail::check(plbamd_toggle_statistics(handle, statsOn ? 1 : 0), "toggle_statistics");
        }
        if (tagsDirty) {
            detail::check(plbamd_set_tags(handle, tags.data()), "set_tags");
            for (int k = 0; k < 3; ++k)
                if (solidRhoSet[k])
                    detail::check(plbamd_set_solid_density(handle, PLBAMD_TAG_BOUNCE_BACK + k, solidRho[k]),
                                  "set_solid_density");
            tagsDirty = false;
        }
        if (planesDirty) {
            for (auto& p : planes) {
                if (p.id < 0) {
                    plbamd_box cb = p.box.toC();
                    if (p.shape == 1) {
                        detail::check(plbamd_add_velocity_edge(handle, p.dir, p.normal[0], p.normal[1], &cb, &p.id),
                                      "add_velocity_edge");
                    } else if (p.shape == 2) {
                        detail::check(plbamd_add_velocity_corner(handle, p.normal[0], p.normal[1], p.normal[2],
                                                                 cb.x0, cb.y0, cb.z0, &p.id),
                                      "add_velocity_corner");
                    } else {
                        const int fam = p.family | (p.wrapped ? PLBAMD_WRAPPED : 0);
                        detail::check(plbamd_add_plane(handle, fam, p.kind, p.dir, p.ori, &cb, &p.id), "add_plane");
                    }
                    if (p.bcType != 0) detail::check(plbamd_set_boundary_type(handle, p.id, p.bcType), "set_boundary_type");
                }
                if (p.valuesDirty) {
                    detail::check(plbamd_set_plane_values(handle, p.id, p.values.data(), p.values.size()),
                                  "set_plane_values");
                    p.valuesDirty = false;
                }
            }
            planesDirty = false;
        }
    }
    void syncHost() {
        materialize();
        if (hostValid) return;
        mirror.resize((size_t)(nx * ny * nz * 19));
        plbamd_box cb = getBoundingBox().toC();
        detail::check(plbamd_download(handle, &cb, StorageOf<T>::code, mirror.data(), mirror.size() * sizeof(T)),
                      "download");
        hostValid = true;
        hostDirty = false;
    }
    void pushHost() {
        if (!hostDirty || !hostValid) return;
        plbamd_box cb = getBoundingBox().toC();
        detail::check(plbamd_upload(handle, &cb, StorageOf<T>::code, mirror.data(), mirror.size() * sizeof(T)), "upload");
        hostDirty = false;
    }
    plint nx, ny, nz;
    Dynamics<T, Descriptor>* background;
    PeriodicitySwitch3D periodic;
    StatSubscriber subscriber;
    BlockStatistics stats;
    std::vector<uint8_t> tags;
    double solidRho[3] = {1, 1, 1};
    bool solidRhoSet[3] = {false, false, false};
    std::vector<Plane> planes;
    double force[3] = {0, 0, 0};
    int handle = 0;
    bool statsOn = true;
    bool tagsDirty = false, planesDirty = false;
    std::vector<T> mirror;
    bool hostValid = false, hostDirty = false;
};

// BlockLattice3D (atomicBlock/blockLattice3D.h:106-200): a single block whose
// collideAndStream() is always periodic (blockLattice3D.hh:318-326).
template<typename T, template<typename U> class Descriptor>
class BlockLattice3D : public MultiBlockLattice3D<T, Descriptor> {
public:
    BlockLattice3D(plint nx_, plint ny_, plint nz_, Dynamics<T, Descriptor>* background_)
        : MultiBlockLattice3D<T, Descriptor>(nx_, ny_, nz_, background_) {
        this->periodicity().toggleAll(true);
    }
};

template<typename T, template<typename U> class Descriptor>
void Cell<T, Descriptor>::defineVelocity(Array<T, 3> const& u) {
    double v[3] = {(double)u[0], (double)u[1], (double)u[2]};
    lat->setPlaneValue(x, y, z, PLBAMD_VELOCITY, v);
}

template<typename T, template<typename U> class Descriptor>
void Cell<T, Descriptor>::defineDensity(T rho) {
    double v[1] = {(double)rho};
    lat->setPlaneValue(x, y, z, PLBAMD_PRESSURE, v);
}

// ---------------------------------------------------------------------------
// Boundary conditions (boundaryCondition/boundaryCondition3D.h:40-476)

template<typename T, template<typename U> class Descriptor>
class OnLatticeBoundaryCondition3D {
public:
    OnLatticeBoundaryCondition3D(int family_, bool wrapped_) : family(family_), wrapped(wrapped_) {}
    virtual ~OnLatticeBoundaryCondition3D() {}
    typedef MultiBlockLattice3D<T, Descriptor> L;
    typedef boundary::BcType B;
    static constexpr B D = boundary::dirichlet;
    void addVelocityBoundary0N(Box3D b, L& l, B t = D) { l.addPlane(family, PLBAMD_VELOCITY, 0, -1, b, wrapped, t); }
    void addVelocityBoundary0P(Box3D b, L& l, B t = D) { l.addPlane(family, PLBAMD_VELOCITY, 0, 1, b, wrapped, t); }
    void addVelocityBoundary1N(Box3D b, L& l, B t = D) { l.addPlane(family, PLBAMD_VELOCITY, 1, -1, b, wrapped, t); }
    void addVelocityBoundary1P(Box3D b, L& l, B t = D) { l.addPlane(family, PLBAMD_VELOCITY, 1, 1, b, wrapped, t); }
    void addVelocityBoundary2N(Box3D b, L& l, B t = D) { l.addPlane(family, PLBAMD_VELOCITY, 2, -1, b, wrapped, t); }
    void addVelocityBoundary2P(Box3D b, L& l, B t = D) { l.addPlane(family, PLBAMD_VELOCITY, 2, 1, b, wrapped, t); }
    void addPressureBoundary0N(Box3D b, L& l, B t = D) { l.addPlane(family, PLBAMD_PRESSURE, 0, -1, b, wrapped, t); }
    void addPressureBoundary0P(Box3D b, L& l, B t = D) { l.addPlane(family, PLBAMD_PRESSURE, 0, 1, b, wrapped, t); }
    void addPressureBoundary1N(Box3D b, L& l, B t = D) { l.addPlane(family, PLBAMD_PRESSURE, 1, -1, b, wrapped, t); }
    void addPressureBoundary1P(Box3D b, L& l, B t = D) { l.addPlane(family, PLBAMD_PRESSURE, 1, 1, b, wrapped, t); }
    void addPressureBoundary2N(Box3D b, L& l, B t = D) { l.addPlane(family, PLBAMD_PRESSURE, 2, -1, b, wrapped, t); }
    void addPressureBoundary2P(Box3D b, L& l, B t = D) { l.addPlane(family, PLBAMD_PRESSURE, 2, 1, b, wrapped, t); }

    // External velocity edges (boundaryCondition3D.h): the Zou-He and regularized
    // factories install the same StoreVelocityDynamics + OuterVelocityEdgeFunctional3D.
    void addExternalVelocityEdge0NN(Box3D b, L& l, B t = D) { l.addEdge(0, -1, -1, b, t); }
    void addExternalVelocityEdge0NP(Box3D b, L& l, B t = D) { l.addEdge(0, -1, 1, b, t); }
    void addExternalVelocityEdge0PN(Box3D b, L& l, B t = D) { l.addEdge(0, 1, -1, b, t); }
    void addExternalVelocityEdge0PP(Box3D b, L& l, B t = D) { l.addEdge(0, 1, 1, b, t); }
    void addExternalVelocityEdge1NN(Box3D b, L& l, B t = D) { l.addEdge(1, -1, -1, b, t); }
    void addExternalVelocityEdge1NP(Box3D b, L& l, B t = D) { l.addEdge(1, -1, 1, b, t); }
    void addExternalVelocityEdge1PN(Box3D b, L& l, B t = D) { l.addEdge(1, 1, -1, b, t); }
    void addExternalVelocityEdge1PP(Box3D b, L& l, B t = D) { l.addEdge(1, 1, 1, b, t); }
    void addExternalVelocityEdge2NN(Box3D b, L& l, B t = D) { l.addEdge(2, -1, -1, b, t); }
    void addExternalVelocityEdge2NP(Box3D b, L& l, B t = D) { l.addEdge(2, -1, 1, b, t); }
    void addExternalVelocityEdge2PN(Box3D b, L& l, B t = D) { l.addEdge(2, 1, -1, b, t); }
    void addExternalVelocityEdge2PP(Box3D b, L& l, B t = D) { l.addEdge(2, 1, 1, b, t); }
    void addExternalVelocityCornerNNN(plint x, plint y, plint z, L& l, B t = D) { l.addCorner(-1, -1, -1, x, y, z, t); }
    void addExternalVelocityCornerNNP(plint x, plint y, plint z, L& l, B t = D) { l.addCorner(-1, -1, 1, x, y, z, t); }
    void addExternalVelocityCornerNPN(plint x, plint y, plint z, L& l, B t = D) { l.addCorner(-1, 1, -1, x, y, z, t); }
    void addExternalVelocityCornerNPP(plint x, plint y, plint z, L& l, B t = D) { l.addCorner(-1, 1, 1, x, y, z, t); }
    void addExternalVelocityCornerPNN(plint x, plint y, plint z, L& l, B t = D) { l.addCorner(1, -1, -1, x, y, z, t); }
    void addExternalVelocityCornerPNP(plint x, plint y, plint z, L& l, B t = D) { l.addCorner(1, -1, 1, x, y, z, t); }
    void addExternalVelocityCornerPPN(plint x, plint y, plint z, L& l, B t = D) { l.addCorner(1, 1, -1, x, y, z, t); }
    void addExternalVelocityCornerPPP(plint x, plint y, plint z, L& l, B t = D) { l.addCorner(1, 1, 1, x, y, z, t); }

    // Velocity conditions on the six face interiors, the twelve edges and the
    // eight corners of `block` that intersect `applicationDomain`
    // (boundaryCondition3D.hh:300-400, BlockSurface3D boxes of
    // core/blockSurface3D.cpp:44-198, boundary width 1).
    void setVelocityConditionOnBlockBoundaries(L& l, Box3D block, Box3D app, B t = D) {
        const plint x0 = block.x0, x1 = block.x1, y0 = block.y0, y1 = block.y1, z0 = block.z0, z1 = block.z1;
        Box3D i;
        if (intersect(Box3D(x0, x0, y0 + 1, y1 - 1, z0 + 1, z1 - 1), app, i)) addVelocityBoundary0N(i, l, t);
        if (intersect(Box3D(x1, x1, y0 + 1, y1 - 1, z0 + 1, z1 - 1), app, i)) addVelocityBoundary0P(i, l, t);
        if (intersect(Box3D(x0 + 1, x1 - 1, y0, y0, z0 + 1, z1 - 1), app, i)) addVelocityBoundary1N(i, l, t);
        if (intersect(Box3D(x0 + 1, x1 - 1, y1, y1, z0 + 1, z1 - 1), app, i)) addVelocityBoundary1P(i, l, t);
        if (intersect(Box3D(x0 + 1, x1 - 1, y0 + 1, y1 - 1, z0, z0), app, i)) addVelocityBoundary2N(i, l, t);
        if (intersect(Box3D(x0 + 1, x1 - 1, y0 + 1, y1 - 1, z1, z1), app, i)) addVelocityBoundary2P(i, l, t);
        // edges: plane x (y, z normals), plane y (z, x normals), plane z (x, y normals)
        if (intersect(Box3D(x0 + 1, x1 - 1, y0, y0, z0, z0), app, i)) addExternalVelocityEdge0NN(i, l, t);
        if (intersect(Box3D(x0 + 1, x1 - 1, y0, y0, z1, z1), app, i)) addExternalVelocityEdge0NP(i, l, t);
        if (intersect(Box3D(x0 + 1, x1 - 1, y1, y1, z0, z0), app, i)) addExternalVelocityEdge0PN(i, l, t);
        if (intersect(Box3D(x0 + 1, x1 - 1, y1, y1, z1, z1), app, i)) addExternalVelocityEdge0PP(i, l, t);
        if (intersect(Box3D(x0, x0, y0 + 1, y1 - 1, z0, z0), app, i)) addExternalVelocityEdge1NN(i, l, t);
        if (intersect(Box3D(x1, x1, y0 + 1, y1 - 1, z0, z0), app, i)) addExternalVelocityEdge1NP(i, l, t);
        if (intersect(Box3D(x0, x0, y0 + 1, y1 - 1, z1, z1), app, i)) addExternalVelocityEdge1PN(i, l, t);
        if (intersect(Box3D(x1, x1, y0 + 1, y1 - 1, z1, z1), app, i)) addExternalVelocityEdge1PP(i, l, t);
        if (intersect(Box3D(x0, x0, y0, y0, z0 + 1, z1 - 1), app, i)) addExternalVelocityEdge2NN(i, l, t);
        if (intersect(Box3D(x0, x0, y1, y1, z0 + 1, z1 - 1), app, i)) addExternalVelocityEdge2NP(i, l, t);
        if (intersect(Box3D(x1, x1, y0, y0, z0 + 1, z1 - 1), app, i)) addExternalVelocityEdge2PN(i, l, t);
        if (intersect(Box3D(x1, x1, y1, y1, z0 + 1, z1 - 1), app, i)) addExternalVelocityEdge2PP(i, l, t);
        if (intersect(Box3D(x0, x0, y0, y0, z0, z0), app, i)) addExternalVelocityCornerNNN(i.x0, i.y0, i.z0, l, t);
        if (intersect(Box3D(x0, x0, y0, y0, z1, z1), app, i)) addExternalVelocityCornerNNP(i.x0, i.y0, i.z0, l, t);
        if (intersect(Box3D(x0, x0, y1, y1, z0, z0), app, i)) addExternalVelocityCornerNPN(i.x0, i.y0, i.z0, l, t);
        if (intersect(Box3D(x0, x0, y1, y1, z1, z1), app, i)) addExternalVelocityCornerNPP(i.x0, i.y0, i.z0, l, t);
        if (intersect(Box3D(x1, x1, y0, y0, z0, z0), app, i)) addExternalVelocityCornerPNN(i.x0, i.y0, i.z0, l, t);
        if (intersect(Box3D(x1, x1, y0, y0, z1, z1), app, i)) addExternalVelocityCornerPNP(i.x0, i.y0, i.z0, l, t);
        if (intersect(Box3D(x1, x1, y1, y1, z0, z0), app, i)) addExternalVelocityCornerPPN(i.x0, i.y0, i.z0, l, t);
        if (intersect(Box3D(x1, x1, y1, y1, z1, z1), app, i)) addExternalVelocityCornerPPP(i.x0, i.y0, i.z0, l, t);
    }
    void setVelocityConditionOnBlockBoundaries(L& l, Box3D app, B t = D) {
        setVelocityConditionOnBlockBoundaries(l, l.getBoundingBox(), app, t);
    }
    void setVelocityConditionOnBlockBoundaries(L& l, B t = D) {
        setVelocityConditionOnBlockBoundaries(l, l.getBoundingBox(), l.getBoundingBox(), t);
    }
    // Pressure conditions on the six face interiors (boundaryCondition3D.hh:420-446;
    // the reference installs nothing on edges and corners).
    void setPressureConditionOnBlockBoundaries(L& l, Box3D block, Box3D app, B t = D) {
        const plint x0 = block.x0, x1 = block.x1, y0 = block.y0, y1 = block.y1, z0 = block.z0, z1 = block.z1;
        Box3D i;
        if (intersect(Box3D(x0, x0, y0 + 1, y1 - 1, z0 + 1, z1 - 1), app, i)) addPressureBoundary0N(i, l, t);
        if (intersect(Box3D(x1, x1, y0 + 1, y1 - 1, z0 + 1, z1 - 1), app, i)) addPressureBoundary0P(i, l, t);
        if (intersect(Box3D(x0 + 1, x1 - 1, y0, y0, z0 + 1, z1 - 1), app, i)) addPressureBoundary1N(i, l, t);
        if (intersect(Box3D(x0 + 1, x1 - 1, y1, y1, z0 + 1, z1 - 1), app, i)) addPressureBoundary1P(i, l, t);
        if (intersect(Box3D(x0 + 1, x1 - 1, y0 + 1, y1 - 1, z0, z0), app, i)) addPressureBoundary2N(i, l, t);
        if (intersect(Box3D(x0 + 1, x1 - 1, y0 + 1, y1 - 1, z1, z1), app, i)) addPressureBoundary2P(i, l, t);
    }
    void setPressureConditionOnBlockBoundaries(L& l, Box3D app, B t = D) {
        setPressureConditionOnBlockBoundaries(l, l.getBoundingBox(), app, t);
    }
    void setPressureConditionOnBlockBoundaries(L& l, B t = D) {
        setPressureConditionOnBlockBoundaries(l, l.getBoundingBox(), l.getBoundingBox(), t);
    }
private:
    int family;
    bool wrapped;
};

// zouHeBoundary3D.hh:338-350, boundaryCondition3D.hh:1163-1175
template<typename T, template<typename U> class Descriptor>
OnLatticeBoundaryCondition3D<T, Descriptor>* createZouHeBoundaryCondition3D() {
    return new OnLatticeBoundaryCondition3D<T, Descriptor>(PLBAMD_ZOUHE, true);
}
template<typename T, template<typename U> class Descriptor>
OnLatticeBoundaryCondition3D<T, Descriptor>* createDynamicsBasedZouHeBoundaryCondition3D() {
    return new OnLatticeBoundaryCondition3D<T, Descriptor>(PLBAMD_ZOUHE, false);
}
template<typename T, template<typename U> class Descriptor>
OnLatticeBoundaryCondition3D<T, Descriptor>* createLocalBoundaryCondition3D() {
    return new OnLatticeBoundaryCondition3D<T, Descriptor>(PLBAMD_REGULARIZED, true);
}
template<typename T, template<typename U> class Descriptor>
OnLatticeBoundaryCondition3D<T, Descriptor>* createDynamicsBasedLocalBoundaryCondition3D() {
    return new OnLatticeBoundaryCondition3D<T, Descriptor>(PLBAMD_REGULARIZED, false);
}

// ---------------------------------------------------------------------------
// Data initialisers and analysis (dataProcessors/dataInitializerWrapper3D.h,
// dataAnalysisWrapper3D.h)

template<typename T, template<typename U> class Descriptor>
void defineDynamics(MultiBlockLattice3D<T, Descriptor>& lattice, Box3D domain, Dynamics<T, Descriptor>* dynamics) {
    uint8_t tag;
    switch (dynamics->kind()) {
    case DynKind::BounceBack: tag = PLBAMD_TAG_BOUNCE_BACK; break;
    case DynKind::NoDynamics: tag = PLBAMD_TAG_NO_DYNAMICS; break;
    case DynKind::MomentumExchange: tag = PLBAMD_TAG_MOMENTUM_EXCHANGE; break;
    case DynKind::WallLink: tag = PLBAMD_TAG_WALL_LINK; break;
    default:
        if (dynamics->getOmega() != lattice.getBackgroundDynamics().getOmega() ||
            dynamics->kind() != lattice.getBackgroundDynamics().kind()) {
            delete dynamics;
            throw PlbException("defineDynamics: bulk cells must use the background dynamics");
        }
        tag = PLBAMD_TAG_BULK;
    }
    if (tag >= PLBAMD_TAG_BOUNCE_BACK && tag <= PLBAMD_TAG_MOMENTUM_EXCHANGE) {
        try {
            lattice.setSolidDensity(tag, (double)dynamics->getDensityParameter());
        } catch (...) {
            delete dynamics;
            throw;
        }
    }
    lattice.setTag(domain, tag);
    delete dynamics;  // defineDynamics takes ownership of the prototype
}

template<typename T, template<typename U> class Descriptor>
void defineDynamics(MultiBlockLattice3D<T, Descriptor>& lattice, plint x, plint y, plint z,
                    Dynamics<T, Descriptor>* dynamics) {
    defineDynamics(lattice, Box3D(x, x, y, y, z, z), dynamics);
}

// Cells where `inside(x, y, z)` is true (a DomainFunctional3D of Palabos).
template<typename T, template<typename U> class Descriptor, class Functional>
void defineDynamics(MultiBlockLattice3D<T, Descriptor>& lattice, Box3D domain, Functional inside,
                    Dynamics<T, Descriptor>* dynamics) {
    for (plint x = domain.x0; x <= domain.x1; ++x)
        for (plint y = domain.y0; y <= domain.y1; ++y)
            for (plint z = domain.z0; z <= domain.z1; ++z)
                if (inside(x, y, z)) defineDynamics(lattice, Box3D(x, x, y, y, z, z), dynamics->clone());
    delete dynamics;
}

template<typename T, template<typename U> class Descriptor>
void initializeMomentumExchange(MultiBlockLattice3D<T, Descriptor>&, Box3D) {
    // fluid directions are derived on the device from the cell tags
}

template<typename T, template<typename U> class Descriptor>
void setBoundaryVelocity(MultiBlockLattice3D<T, Descriptor>& lattice, Box3D domain, Array<T, 3> const& u) {
    double v[3] = {(double)u[0], (double)u[1], (double)u[2]};
    for (plint x = domain.x0; x <= domain.x1; ++x)
        for (plint y = domain.y0; y <= domain.y1; ++y)
            for (plint z = domain.z0; z <= domain.z1; ++z) lattice.setPlaneValue(x, y, z, PLBAMD_VELOCITY, v);
}

template<typename T, template<typename U> class Descriptor>
void setBoundaryDensity(MultiBlockLattice3D<T, Descriptor>& lattice, Box3D domain, T rho) {
    double v[1] = {(double)rho};
    for (plint x = domain.x0; x <= domain.x1; ++x)
        for (plint y = domain.y0; y <= domain.y1; ++y)
            for (plint z = domain.z0; z <= domain.z1; ++z) lattice.setPlaneValue(x, y, z, PLBAMD_PRESSURE, v);
}

template<typename T, template<typename U> class Descriptor>
void setExternalVector(MultiBlockLattice3D<T, Descriptor>& lattice, Box3D domain, int, Array<T, 3> const& F) {
    Box3D all = lattice.getBoundingBox();
    if (domain.x0 != all.x0 || domain.x1 != all.x1 || domain.y0 != all.y0 || domain.y1 != all.y1 ||
        domain.z0 != all.z0 || domain.z1 != all.z1)
        throw PlbException("setExternalVector: the body force must be uniform over the lattice");
    lattice.setForce(F);
}

template<typename T, template<typename U> class Descriptor>
void initializeAtEquilibrium(MultiBlockLattice3D<T, Descriptor>& lattice, Box3D domain, T rho, Array<T, 3> const& u) {
    size_t n = (size_t)domain.nCells();
    std::vector<double> r(n, (double)rho), v(3 * n);
    for (size_t k = 0; k < n; ++k)
        for (int a = 0; a < 3; ++a) v[3 * k + a] = (double)u[a];
    lattice.equilibrium(domain, r, v);
}

// Per-cell density and velocity from f(x, y, z, rho, u).
template<typename T, template<typename U> class Descriptor, class RhoUFunction>
void initializeAtEquilibrium(MultiBlockLattice3D<T, Descriptor>& lattice, Box3D domain, RhoUFunction f) {
    size_t n = (size_t)domain.nCells();
    std::vector<double> r(n), v(3 * n);
    size_t k = 0;
    for (plint x = domain.x0; x <= domain.x1; ++x)
        for (plint y = domain.y0; y <= domain.y1; ++y)
            for (plint z = domain.z0; z <= domain.z1; ++z, ++k) {
                T rho;
                Array<T, 3> u;
                f(x, y, z, rho, u);
                r[k] = (double)rho;
                for (int a = 0; a < 3; ++a) v[3 * k + a] = (double)u[a];
            }
    lattice.equilibrium(domain, r, v);
}

// dataAnalysisWrapper3D.hh:1226-1257
template<typename T, template<typename U> class Descriptor>
T computeAverageDensity(MultiBlockLattice3D<T, Descriptor>& lattice, Box3D domain) {
    double s[5];
    lattice.moments(domain, s);
    return Descriptor<T>::fullRho((T)(s[0] / (double)domain.nCells()));
}
template<typename T, template<typename U> class Descriptor>
T computeAverageDensity(MultiBlockLattice3D<T, Descriptor>& lattice) {
    return computeAverageDensity(lattice, lattice.getBoundingBox());
}
template<typename T, template<typename U> class Descriptor>
T computeAverageEnergy(MultiBlockLattice3D<T, Descriptor>& lattice, Box3D domain) {
    double s[5];
    lattice.moments(domain, s);
    return (T)(s[1] / 2.0) / (T)domain.nCells();
}
template<typename T, template<typename U> class Descriptor>
T computeAverageEnergy(MultiBlockLattice3D<T, Descriptor>& lattice) {
    return computeAverageEnergy(lattice, lattice.getBoundingBox());
}

// Scalar and vector fields on a box (the multiScalarField3D.h /
// multiTensorField3D.h role for the analysis functions below): values in C
// order over the box, get() in the box's own coordinates from 0.
template<typename T>
class MultiScalarField3D {
public:
    explicit MultiScalarField3D(Box3D const& b) : box(b), data((size_t)b.nCells()) {}
    plint getNx() const { return box.getNx(); }
    plint getNy() const { return box.getNy(); }
    plint getNz() const { return box.getNz(); }
    Box3D getBoundingBox() const { return Box3D(0, getNx() - 1, 0, getNy() - 1, 0, getNz() - 1); }
    T& get(plint x, plint y, plint z) { return data[(size_t)((x * getNy() + y) * getNz() + z)]; }
    T const& get(plint x, plint y, plint z) const { return data[(size_t)((x * getNy() + y) * getNz() + z)]; }
private:
    Box3D box;
    std::vector<T> data;
};
template<typename T, int nDim>
class MultiTensorField3D {
public:
    explicit MultiTensorField3D(Box3D const& b) : box(b), data((size_t)b.nCells()) {}
    plint getNx() const { return box.getNx(); }
    plint getNy() const { return box.getNy(); }
    plint getNz() const { return box.getNz(); }
    Array<T, nDim>& get(plint x, plint y, plint z) { return data[(size_t)((x * getNy() + y) * getNz() + z)]; }
    Array<T, nDim> const& get(plint x, plint y, plint z) const { return data[(size_t)((x * getNy() + y) * getNz() + z)]; }
private:
    Box3D box;
    std::vector<Array<T, nDim>> data;
};

namespace detail {
template<typename T, template<typename U> class Descriptor>
std::unique_ptr<MultiScalarField3D<T>> analysisScalar(MultiBlockLattice3D<T, Descriptor>& lattice, Box3D domain,
                                                      int column) {
    std::vector<double> f;
    lattice.analysisFields(domain, f);
    std::unique_ptr<MultiScalarField3D<T>> out(new MultiScalarField3D<T>(domain));
    size_t k = 0;
    for (plint x = 0; x < domain.getNx(); ++x)
        for (plint y = 0; y < domain.getNy(); ++y)
            for (plint z = 0; z < domain.getNz(); ++z, ++k) out->get(x, y, z) = (T)f[7 * k + (size_t)column];
    return out;
}
}  // namespace detail

// dataProcessors/dataAnalysisWrapper3D.h:695-777 (per cell the cell's dynamics
// decides; the values are computed on the device in T).
template<typename T, template<typename U> class Descriptor>
std::unique_ptr<MultiScalarField3D<T>> computeDensity(MultiBlockLattice3D<T, Descriptor>& lattice, Box3D domain) {
    return detail::analysisScalar(lattice, domain, 0);
}
template<typename T, template<typename U> class Descriptor>
std::unique_ptr<MultiScalarField3D<T>> computeDensity(MultiBlockLattice3D<T, Descriptor>& lattice) {
    return computeDensity(lattice, lattice.getBoundingBox());
}
template<typename T, template<typename U> class Descriptor>
std::unique_ptr<MultiScalarField3D<T>> computeRhoBar(MultiBlockLattice3D<T, Descriptor>& lattice, Box3D domain) {
    return detail::analysisScalar(lattice, domain, 1);
}
template<typename T, template<typename U> class Descriptor>
std::unique_ptr<MultiScalarField3D<T>> computeRhoBar(MultiBlockLattice3D<T, Descriptor>& lattice) {
    return computeRhoBar(lattice, lattice.getBoundingBox());
}
template<typename T, template<typename U> class Descriptor>
std::unique_ptr<MultiScalarField3D<T>> computeKineticEnergy(MultiBlockLattice3D<T, Descriptor>& lattice, Box3D domain) {
    return detail::analysisScalar(lattice, domain, 5);
}
template<typename T, template<typename U> class Descriptor>
std::unique_ptr<MultiScalarField3D<T>> computeKineticEnergy(MultiBlockLattice3D<T, Descriptor>& lattice) {
    return computeKineticEnergy(lattice, lattice.getBoundingBox());
}
template<typename T, template<typename U> class Descriptor>
std::unique_ptr<MultiScalarField3D<T>> computeVelocityNorm(MultiBlockLattice3D<T, Descriptor>& lattice, Box3D domain) {
    return detail::analysisScalar(lattice, domain, 6);
}
template<typename T, template<typename U> class Descriptor>
std::unique_ptr<MultiScalarField3D<T>> computeVelocityNorm(MultiBlockLattice3D<T, Descriptor>& lattice) {
    return computeVelocityNorm(lattice, lattice.getBoundingBox());
}
template<typename T, template<typename U> class Descriptor>
std::unique_ptr<MultiScalarField3D<T>> computeVelocityComponent(MultiBlockLattice3D<T, Descriptor>& lattice,
                                                                Box3D domain, plint iComponent) {
    return detail::analysisScalar(lattice, domain, 2 + (int)iComponent);
}
template<typename T, template<typename U> class Descriptor>
std::unique_ptr<MultiTensorField3D<T, 3>> computeVelocity(MultiBlockLattice3D<T, Descriptor>& lattice, Box3D domain) {
    std::vector<double> f;
    lattice.analysisFields(domain, f);
    std::unique_ptr<MultiTensorField3D<T, 3>> out(new MultiTensorField3D<T, 3>(domain));
    size_t k = 0;
    for (plint x = 0; x < domain.getNx(); ++x)
        for (plint y = 0; y < domain.getNy(); ++y)
            for (plint z = 0; z < domain.getNz(); ++z, ++k)
                for (int a = 0; a < 3; ++a) out->get(x, y, z)[a] = (T)f[7 * k + 2 + (size_t)a];
    return out;
}
template<typename T, template<typename U> class Descriptor>
std::unique_ptr<MultiTensorField3D<T, 3>> computeVelocity(MultiBlockLattice3D<T, Descriptor>& lattice) {
    return computeVelocity(lattice, lattice.getBoundingBox());
}

// io/serializerIO_3D.h:48,58: the file holds the raw f-bar (and, for
// ForcedD3Q19Descriptor, the force externals) of every cell, base64-coded.
template<typename T, template<typename U> class Descriptor>
void saveBinaryBlock(MultiBlockLattice3D<T, Descriptor>& lattice, std::string fName, bool enforceUint = false) {
    lattice.saveBinary(fName, enforceUint);
}
template<typename T, template<typename U> class Descriptor>
void loadBinaryBlock(MultiBlockLattice3D<T, Descriptor>& lattice, std::string fName, bool enforceUint = false) {
    lattice.loadBinary(fName, enforceUint);
}

// multiBlockLattice3D.hh:531-546
template<typename T, template<typename U> class Descriptor>
double getStoredAverageDensity(MultiBlockLattice3D<T, Descriptor> const& lattice) {
    return Descriptor<T>::fullRho((T)lattice.getInternalStatistics().getAverage(LatticeStatistics::avRhoBar));
}
template<typename T, template<typename U> class Descriptor>
double getStoredAverageEnergy(MultiBlockLattice3D<T, Descriptor> const& lattice) {
    return 0.5 * lattice.getInternalStatistics().getAverage(LatticeStatistics::avUSqr);
}
template<typename T, template<typename U> class Descriptor>
double getStoredMaxVelocity(MultiBlockLattice3D<T, Descriptor> const& lattice) {
    return std::sqrt(lattice.getInternalStatistics().getMax(LatticeStatistics::maxUSqr));
}

// core/units.h:39-95
template<typename T>
class IncomprFlowParam {
public:
    /// Physical velocity and length given (core/units.h:49-52).
    IncomprFlowParam(T physicalU_, T latticeU_, T Re_, T physicalLength_, plint resolution_, T lx_, T ly_, T lz_ = T())
        : physicalU(physicalU_), latticeU(latticeU_), Re(Re_), physicalLength(physicalLength_),
          resolution(resolution_), lx(lx_), ly(ly_), lz(lz_) {}
    /// Dimensionless form: physical velocity and length 1 (core/units.h:54-59).
    IncomprFlowParam(T latticeU_, T Re_, plint resolution_, T lx_, T ly_, T lz_ = T())
        : physicalU((T)1), latticeU(latticeU_), Re(Re_), physicalLength((T)1),
          resolution(resolution_), lx(lx_), ly(ly_), lz(lz_) {}
    T getLatticeU() const { return latticeU; }
    T getPhysicalU() const { return physicalU; }
    T getRe() const { return Re; }
    T getPhysicalLength() const { return physicalLength; }
    plint getResolution() const { return resolution; }
    /// extents in dimensionless units: the physical length times lx, ly, lz
    T getLx() const { return physicalLength * lx; }
    T getLy() const { return physicalLength * ly; }
    T getLz() const { return physicalLength * lz; }
    T getDeltaX() const { return physicalLength / (T)resolution; }
    T getDeltaT() const { return getDeltaX() * latticeU / physicalU; }
    /// dimensionless length or time to cells or steps, rounded to nearest (int, as the reference)
    plint nCell(T l) const { return (int)(l / getDeltaX() + (T)0.5); }
    plint nStep(T t) const { return (int)(t / getDeltaT() + (T)0.5); }
    /// cells along x, y, z; one more with offLattice
    plint getNx(bool offLattice = false) const { return nCell(getLx()) + 1 + (int)offLattice; }
    plint getNy(bool offLattice = false) const { return nCell(getLy()) + 1 + (int)offLattice; }
    plint getNz(bool offLattice = false) const { return nCell(getLz()) + 1 + (int)offLattice; }
    T getLatticeNu() const { return latticeU * (T)resolution / Re; }
    T getTau() const { return (T)3 * getLatticeNu() + (T)0.5; }
    T getOmega() const { return (T)1 / getTau(); }
private:
    T physicalU, latticeU, Re, physicalLength;
    plint resolution;
    T lx, ly, lz;
};

// ---------------------------------------------------------------------------
// Co-processor plug-in (coProcessors/coProcessor3D.h:59-186)

template<typename T>
class CoProcessor3D {
public:
    virtual ~CoProcessor3D() {}
    /// Returns 1 on success, 0 on failure (coProcessor3D.h:73-75).
    virtual int addDomain(plint nx, plint ny, plint nz, T omega, int& domainHandle) = 0;
    virtual int send(int domainHandle, Box3D const& subDomain, std::vector<char> const& data) = 0;
    virtual int receive(int domainHandle, Box3D const& subDomain, std::vector<char>& data) const = 0;
    virtual int collideAndStream(int domainHandle) = 0;
};

// D3Q19 BGK domains on the MI355X engine. A domain is a Palabos block with its
// envelope; coordinates 0..n-1; data is the static cell layout of T f-bar
// ([nx][ny][nz][19], blockLattice3D.hh:650-690). The outer layer uses the
// same-step edge reflection (the co-processor contract leaves it undefined).
template<typename T>
class HipCoProcessor3D : public CoProcessor3D<T> {
public:
    ~HipCoProcessor3D() override {
        for (int h : handles) plbamd_lattice_destroy(h);
    }
    int addDomain(plint nx, plint ny, plint nz, T omega, int& domainHandle) override {
        try {
            detail::ensureInit();
        } catch (...) {
            return 0;
        }
        plbamd_lattice_desc d{};
        d.nx = (int32_t)nx; d.ny = (int32_t)ny; d.nz = (int32_t)nz;
        d.storage = StorageOf<T>::code;
        d.dynamics = PLBAMD_BGK;
        d.omega = (double)omega;
        int h = 0;
        if (plbamd_lattice_create(&d, &h) != PLBAMD_OK) return 0;
        plbamd_toggle_statistics(h, 0);
        handles.push_back(h);
        domainHandle = h;
        return 1;
    }
    int send(int domainHandle, Box3D const& subDomain, std::vector<char> const& data) override {
        plbamd_box b = subDomain.toC();
        return plbamd_upload(domainHandle, &b, StorageOf<T>::code, data.data(), data.size()) == PLBAMD_OK ? 1 : 0;
    }
    int receive(int domainHandle, Box3D const& subDomain, std::vector<char>& data) const override {
        data.resize((size_t)subDomain.nCells() * 19 * sizeof(T));
        plbamd_box b = subDomain.toC();
        return plbamd_download(domainHandle, &b, StorageOf<T>::code, data.data(), data.size()) == PLBAMD_OK ? 1 : 0;
    }
    int collideAndStream(int domainHandle) override {
        return plbamd_collide_and_stream(domainHandle, 1) == PLBAMD_OK ? 1 : 0;
    }
private:
    std::vector<int> handles;
};

namespace global {
// The process-wide co-processor (coProcessor3D.h:180-186), here the HIP one.
template<typename T>
CoProcessor3D<T>& defaultCoProcessor3D() {
    static HipCoProcessor3D<T> instance;
    return instance;
}
}  // namespace global

}  // namespace plb
