// palabos_amd engine: host-side state of a lattice and the helpers shared by
// the translation units of the library (engine.hip: set-up and C ABI;
// step.hip: the step launches of one storage type and dynamics;
// storage_ops.hip: transfers, checkpoints, equilibrium and moments of one
// storage type; adjoint.hip: the adjoint gradient). The kernels and the data
// layout are described in kernels.h.
#ifndef PALABOS_AMD_ENGINE_H
#define PALABOS_AMD_ENGINE_H

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/palabos_amd.h"
#include "adjoint.h"
#include "checkpoint.h"
#include "fd_boundary.h"
#include "kernels.h"

namespace plbamd {

extern thread_local std::string g_lastError;

int fail(int code, const std::string& msg);

#define PLBAMD_HIP(expr)                                                                  \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess)                                                             \
            return fail(PLBAMD_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define PLBAMD_TRY(expr)          \
    do {                          \
        int rc_ = (expr);         \
        if (rc_) return rc_;      \
    } while (0)

// Envelope width of multi-block lattices: the AA pattern needs two layers
// refreshed every second step (kernels.h).
constexpr int kEnvelope = 2;
constexpr int kEvenThreads = 256;   // threads per workgroup of the linear even step

// A boundary object: a plane (velocity or pressure), an external velocity
// edge or an external velocity corner, with its boundary type.
enum BoundaryShape { SHAPE_PLANE = 0, SHAPE_EDGE = 1, SHAPE_CORNER = 2 };

struct Plane {
    int family, kind, dir, ori;
    bool wrapped;
    plbamd_box box;
    std::vector<double> values;
    int shape = SHAPE_PLANE;
    int normal[3] = {0, 0, 0};   // outward normal (planes: ori along dir)
    int bcType = PLBAMD_BC_DIRICHLET;
};

struct Lattice {
    plbamd_lattice_desc desc;
    int device = -1;                  // HIP device the lattice lives on
    bool multi = false;
    std::vector<plbamd_box> bulk;     // global bulk box of each block
    std::vector<BlockGeom> geom;      // device geometry of each block (host copy)
    BlockGeom* dGeom = nullptr;
    uint16_t* dWgBlock = nullptr;   // block of every bulk work item (multi-block)
    uint32_t* dWalls = nullptr;     // link-wall mask per allocated cell (cells next to link walls)
    bool linearEven = false;        // even steps through bulk_even_linear
    uint64_t nEven = 0;             // its workgroups
    uint16_t* dEvenBlock = nullptr; // its block of every workgroup (multi-block)
    CopyTask* dTasks = nullptr;     // envelope exchange: one workgroup per task
    uint32_t nTasks = 0;
    // the adjoint state's exchange: the first envelope layer only (the reverse
    // step reads lambda one cell beyond the bulk)
    CopySegment* dSegs1 = nullptr;
    CopyTask* dTasks1 = nullptr;
    uint32_t nTasks1 = 0;
    uint64_t allocCells = 0;          // cells of all blocks (per-cell arrays)
    uint64_t cells = 0;               // cells of the domain
    size_t elemSize = 0;
    void* arena = nullptr;
    int64_t time = 0;
    bool statsOn = true;
    // adjoint: device bytes for moments recorded during the forward sweep (the
    // last segments then need no recompute); -1: the free memory beyond a reserve
    int64_t adjointRecordBytes = -1;
    int adjointKept = 0;               // segments the last adjoint call recorded in its forward sweep
    int64_t adjointKeptBytes = 0;      // and their device bytes
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;

    std::vector<uint8_t> hostTags;    // PLBAMD_TAG_* per domain cell, or empty
    std::vector<Plane> planes;
    std::vector<double> hostAlpha;    // per domain cell, or empty
    double solidRho[3] = {1.0, 1.0, 1.0};  // density parameter of BounceBack, NoDynamics, momentum exchange

    bool dirty = true;                // device tags / specials / porosity need a rebuild
    bool anyWrapped = false;
    uint8_t* dTags = nullptr;
    SpecialCell* dSpecial = nullptr;
    uint32_t nSpecial = 0;
    double* dValues = nullptr;
    void* dAlpha = nullptr;        // porosity per cell in the arithmetic type (double for FP64 storage, else float)
    double* dBulkPartials = nullptr;
    double* dBulkLevel1 = nullptr;  // kStatsLevel1 records: first level of the statistics reduction
    uint64_t nBulkPartials = 0;       // = workgroups of the bulk launch
    int xcdRemap = 0;
    std::vector<uint32_t> planeValueBase;   // first value of each plane in dValues
    bool staged = false;           // odd steps through bulk_odd_staged
    bool pairs = false;            // 16-bit storage through bulk_pairs
    bool packed = false;           // bulk_pairs with whole rows per workgroup
    bool zWrap = false;            // every block periodic along z
    int ldsPitch = 0;              // packed: LDS pair words per staged direction
    bool cachedZ = false;          // staged odd steps with ordinary stores for the z-moving outputs (CACHEDZ)
    size_t pairsLds = 0;           // its dynamic LDS bytes (odd steps)
    size_t stagedLds = 0;          // its dynamic LDS bytes
    double* dSpecialPartials = nullptr;
    uint64_t nSpecialPartials = 0;
    double* dStats = nullptr;         // 16 doubles
    bool statsValid = false;

    CopySegment* dSegs = nullptr;
    int nSegs = 0;
    uint64_t segCells = 0;

    // data processors after the stream (fd_boundary.h): records of the four
    // phases [procPhase[p], procPhase[p + 1])
    ProcCell* dProc = nullptr;
    void* dExtField = nullptr;        // external rhoBar / j field, [4][allocCells] of the arithmetic type
    uint32_t procPhase[5] = {0, 0, 0, 0, 0};
    bool anyCopy = false;             // some boundary copies its value (the device values change)

    dim3 bulkBlock;
    int zChunk = 0;

    ~Lattice() {
        hipFree(arena); hipFree(dGeom); hipFree(dTags); hipFree(dSpecial); hipFree(dValues); hipFree(dAlpha);
        hipFree(dBulkPartials); hipFree(dBulkLevel1); hipFree(dSpecialPartials); hipFree(dStats); hipFree(dSegs); hipFree(dSegs1); hipFree(dTasks1);
        hipFree(dWgBlock); hipFree(dTasks); hipFree(dWalls); hipFree(dEvenBlock); hipFree(dProc); hipFree(dExtField);
        if (ev0) hipEventDestroy(ev0);
        if (ev1) hipEventDestroy(ev1);
        if (stream) hipStreamDestroy(stream);
    }
};

// One collide-and-stream step of every block. With `record` (FP64 BGK only,
// statistics off), the kernels also write the pre-collision moments of every
// cell and the pre-completion populations of every special record.
struct Recording {
    void* moments = nullptr;       // MomentType<S> [4][momStride]
    uint64_t momStride = 0;
    double* specialF = nullptr;
};

// Device buffers of one call, freed on every return path.
struct DeviceBuffers {
    std::vector<void*> p;
    void* get(size_t bytes) {
        void* q = nullptr;
        if (hipMalloc(&q, bytes) != hipSuccess) { (void)hipGetLastError(); return nullptr; }
        p.push_back(q);
        return q;
    }
    ~DeviceBuffers() { for (void* q : p) (void)hipFree(q); }
};

inline dim3 itemsGrid(uint64_t n) { return dim3((unsigned)std::max<uint64_t>(1, (n + 255) / 256)); }

Lattice* find(int handle);
bool boxValid(const Lattice& L, const plbamd_box& b);
uint64_t boxCells(const plbamd_box& b);
bool intersect(const plbamd_box& a, const plbamd_box& b, plbamd_box& out);
size_t storageSize(int storage);
dim3 gridFor(uint64_t n);
int rebuild(Lattice& L);
int step(Lattice& L, int nsteps);
int recordStep(Lattice& L, bool odd, const Recording& r);
bool boxArgs(const Lattice& L, size_t bi, const plbamd_box& g, BoxArgs& a);
std::string base64Host(const uint8_t* p, size_t n);
bool unbase64Host(const char* s, size_t n, uint8_t* out);
int adjointGradient(Lattice& L, int T, int K, const plbamd_objective& obj, double* Jout, double* gradAlpha,
                    double* gradScale, double* timing);

// Typed operations, explicitly instantiated per storage type S (and dynamics
// DYN) in their own translation units.
template<typename S, int DYN>
int launchStepTyped(Lattice& L, bool odd, bool stats, const Recording* record = nullptr);
template<typename S>
int exchange(Lattice& L);
template<typename S>
hipError_t transferDevice(Lattice& L, const plbamd_box& b, int elem, void* staging, bool up);
template<typename S>
int transfer(Lattice& L, const plbamd_box& b, int elem, void* data, size_t bytes, bool up);
template<typename S>
int checkpoint(Lattice& L, const char* path, int elem, bool force, bool enforceUint, bool save);
template<typename S>
int equilibrium(Lattice& L, const plbamd_box& b, const double* rho, const double* u);
// Caller-provided room for the per-workgroup records of moments(): `base`
// holds maxRecords records, then kMomentsLevel1 first-level records, then the
// final record (8 doubles each). With it, moments() allocates nothing and
// refuses lattices whose device set-up is out of date (rebuild allocates).
constexpr uint32_t kMomentsLevel1 = 256;
struct MomentsScratch {
    double* base;
    uint64_t maxRecords;
};
template<typename S>
int moments(Lattice& L, const plbamd_box& b, double* rho, double* u, double* sums, double* fields = nullptr,
            const MomentsScratch* scratch = nullptr);
// One pass of the data processors on the state of layout parity `parity`
// (launches only).
template<typename S>
int launchProcessors(Lattice& L, int parity);
// The external-moment field and step (external.h): op 0 compute on b, 1 get, 2 set.
template<typename S>
int externalField(Lattice& L, const plbamd_box& b, double* rhoBar, double* j, int op);
template<typename S>
int externalSteps(Lattice& L, int nsteps);

// Finds the device address of the fixed scratch array of
// plbamd_reduce_moments_noalloc on `device`, which must be current (noalloc.hip).
int scratchSetup(int device);

}  // namespace plbamd

#endif
