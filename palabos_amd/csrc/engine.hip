// palabos_amd engine: D3Q19 collide-and-stream on gfx950 (MI355X) — host
// runtime and C ABI (include/palabos_amd.h). The kernels and the data layout
// are described in kernels.h.
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
#include "kernels.h"

namespace plbamd {

thread_local std::string g_lastError;

int fail(int code, const std::string& msg) {
    g_lastError = msg;
    return code;
}

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

struct Plane {
    int family, kind, dir, ori;
    bool wrapped;
    plbamd_box box;
    std::vector<double> values;
};

struct Lattice {
    plbamd_lattice_desc desc;
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
    double* dAlpha = nullptr;
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
    size_t pairsLds = 0;           // its dynamic LDS bytes (odd steps)
    size_t stagedLds = 0;          // its dynamic LDS bytes
    double* dSpecialPartials = nullptr;
    uint64_t nSpecialPartials = 0;
    double* dStats = nullptr;         // 16 doubles
    bool statsValid = false;

    CopySegment* dSegs = nullptr;
    int nSegs = 0;
    uint64_t segCells = 0;

    dim3 bulkBlock;
    int zChunk = 0;

    ~Lattice() {
        hipFree(arena); hipFree(dGeom); hipFree(dTags); hipFree(dSpecial); hipFree(dValues); hipFree(dAlpha);
        hipFree(dBulkPartials); hipFree(dBulkLevel1); hipFree(dSpecialPartials); hipFree(dStats); hipFree(dSegs); hipFree(dSegs1); hipFree(dTasks1);
        hipFree(dWgBlock); hipFree(dTasks); hipFree(dWalls); hipFree(dEvenBlock);
        if (ev0) hipEventDestroy(ev0);
        if (ev1) hipEventDestroy(ev1);
        if (stream) hipStreamDestroy(stream);
    }
};

std::mutex g_mutex;
std::map<int, std::unique_ptr<Lattice>> g_lattices;
int g_nextHandle = 1;

Lattice* find(int handle) {
    std::lock_guard<std::mutex> lock(g_mutex);
    auto it = g_lattices.find(handle);
    return it == g_lattices.end() ? nullptr : it->second.get();
}

bool boxValid(const Lattice& L, const plbamd_box& b) {
    return b.x0 >= 0 && b.y0 >= 0 && b.z0 >= 0 && b.x0 <= b.x1 && b.y0 <= b.y1 && b.z0 <= b.z1 &&
           b.x1 < L.desc.nx && b.y1 < L.desc.ny && b.z1 < L.desc.nz;
}

uint64_t boxCells(const plbamd_box& b) {
    return (uint64_t)(b.x1 - b.x0 + 1) * (b.y1 - b.y0 + 1) * (b.z1 - b.z0 + 1);
}

bool intersect(const plbamd_box& a, const plbamd_box& b, plbamd_box& out) {
    out.x0 = std::max(a.x0, b.x0); out.x1 = std::min(a.x1, b.x1);
    out.y0 = std::max(a.y0, b.y0); out.y1 = std::min(a.y1, b.y1);
    out.z0 = std::max(a.z0, b.z0); out.z1 = std::min(a.z1, b.z1);
    return out.x0 <= out.x1 && out.y0 <= out.y1 && out.z0 <= out.z1;
}

size_t storageSize(int storage) {
    return storage == PLBAMD_F64 ? 8 : (storage == PLBAMD_F32 ? 4 : 2);
}

dim3 gridFor(uint64_t n) {
    uint64_t b = (n + 255) / 256;
    if (b > 65536) b = 65536;
    if (b == 0) b = 1;
    return dim3((unsigned)b);
}

int validateDesc(const plbamd_lattice_desc* desc) {
    if (desc->nx <= 0 || desc->ny <= 0 || desc->nz <= 0 || desc->nx > 65535 || desc->ny > 65535 || desc->nz > 65535)
        return fail(PLBAMD_EINVAL, "lattice_create: extents must be in [1, 65535]");
    if (desc->storage < PLBAMD_F64 || desc->storage > PLBAMD_BF16)
        return fail(PLBAMD_EINVAL, "lattice_create: unknown storage type");
    if (desc->dynamics != PLBAMD_BGK && desc->dynamics != PLBAMD_RBGK)
        return fail(PLBAMD_EINVAL, "lattice_create: unknown dynamics");
    if (!(desc->omega > 0.0 && desc->omega < 2.0)) return fail(PLBAMD_EINVAL, "lattice_create: omega must lie in (0, 2)");
    return PLBAMD_OK;
}

// Lays out the blocks in the arena, sets the launch decomposition and the
// envelope exchange segments, and allocates the device memory.
int setup(Lattice& L) {
    const plbamd_lattice_desc& d = L.desc;
    const int n[3] = {d.nx, d.ny, d.nz};
    const int per[3] = {d.periodic[0] != 0, d.periodic[1] != 0, d.periodic[2] != 0};
    // Launch shape: TX threads along z (a multiple of the 64-lane wavefront),
    // CPT cells each per row segment; TY rows per workgroup when a row is short.
    // Measured on MI355X (512^3 periodic): 256 threads for FP64, 512 for the
    // 4- and 2-byte storages (fewer, larger workgroups for the cheaper cells),
    // workgroups remapped so that neighbouring row segments share an XCD.
    int txMax = storageSize(d.storage) >= 8 ? 256 : 512, cpt = 1;
    L.xcdRemap = 1;
    if (const char* e = getenv("PLBAMD_XCD_REMAP")) L.xcdRemap = atoi(e) != 0;
    // at most 512 threads: the step kernels are compiled for 512-thread workgroups
    // (a 1024-thread bound caps them at 128 registers and the heavier variants spill)
    if (const char* e = getenv("PLBAMD_BULK_TX")) txMax = std::max(64, std::min(512, atoi(e) / 64 * 64));
    if (const char* e = getenv("PLBAMD_BULK_CPT")) cpt = std::max(1, std::min(16, atoi(e)));
    int maxNz = 0;
    L.geom.clear();
    for (const plbamd_box& b : L.bulk) {
        BlockGeom g{};
        const int lo[3] = {b.x0, b.y0, b.z0}, hi[3] = {b.x1, b.y1, b.z1};
        int env0[3], env1[3];
        for (int a = 0; a < 3; ++a) {
            if (!L.multi) { env0[a] = env1[a] = 0; continue; }
            env0[a] = (lo[a] > 0 || per[a]) ? kEnvelope : 0;
            env1[a] = (hi[a] < n[a] - 1 || per[a]) ? kEnvelope : 0;
        }
        // 16-bit storage runs two cells per thread when every block is even along
        // z: widen one z envelope to three cells where that makes the block even.
        if (L.multi && storageSize(d.storage) == 2 && (hi[2] - lo[2] + 1 + env0[2] + env1[2]) % 2 != 0) {
            if (env1[2] > 0) env1[2] += 1;
            else if (env0[2] > 0) env0[2] += 1;
        }
        g.nx = hi[0] - lo[0] + 1 + env0[0] + env1[0];
        g.ny = hi[1] - lo[1] + 1 + env0[1] + env1[1];
        g.nz = hi[2] - lo[2] + 1 + env0[2] + env1[2];
        g.px = L.multi ? 0 : per[0];
        g.py = L.multi ? 0 : per[1];
        g.pz = L.multi ? 0 : per[2];
        g.bx0 = env0[0]; g.by0 = env0[1]; g.bz0 = env0[2];
        g.bx1 = env0[0] + hi[0] - lo[0]; g.by1 = env0[1] + hi[1] - lo[1]; g.bz1 = env0[2] + hi[2] - lo[2];
        g.gx0 = lo[0] - env0[0]; g.gy0 = lo[1] - env0[1]; g.gz0 = lo[2] - env0[2];
        maxNz = std::max(maxNz, g.nz);
        L.geom.push_back(g);
    }
    if (L.multi && (int64_t)L.geom.size() > 65535) return fail(PLBAMD_EINVAL, "multiblock: at most 65535 blocks");
    // Odd steps of FP64 and FP32 lattices go through the LDS-staged kernel (one
    // row segment per workgroup; measured 9 % and 13 % faster at 512^3); 16-bit
    // storage keeps the unstaged kernel, which is faster there (its loads are
    // narrow enough that the extra LDS traffic costs more than the misaligned
    // lines). PLBAMD_STAGED=0/1 overrides. The unstaged kernel packs several
    // short rows into one workgroup.
    bool staged = cpt == 1 && storageSize(d.storage) >= 4;
    if (const char* e = getenv("PLBAMD_STAGED")) staged = cpt == 1 && atoi(e) != 0;
    // 16-bit storage with every block's nz even: two cells per thread
    // (bulk_pairs, 32-bit pair accesses, its own LDS staging of odd steps).
    bool pairs = cpt == 1 && storageSize(d.storage) == 2;
    for (const BlockGeom& g : L.geom) pairs = pairs && (g.nz % 2 == 0);
    if (const char* e = getenv("PLBAMD_PAIRS")) pairs = pairs && atoi(e) != 0;
    if (pairs) {
        staged = false;
        if (!getenv("PLBAMD_BULK_TX")) txMax = 256;
    }
    const int cellsPerThread = pairs ? 2 : 1;
    int tx = (int)std::min<int64_t>(txMax, ((int64_t)(maxNz + cellsPerThread - 1) / cellsPerThread + 63) / 64 * 64);
    if ((staged || pairs) && !getenv("PLBAMD_BULK_TX")) {
        // one row segment per workgroup: take the width (a multiple of 64 up to the
        // maximum) that leaves the fewest idle threads over every block's rows,
        // the widest among equals
        uint64_t best = UINT64_MAX;
        for (int w = 64; w <= tx; w += 64) {
            uint64_t threads = 0;
            for (const BlockGeom& g : L.geom) {
                const int units = (g.nz + cellsPerThread - 1) / cellsPerThread;
                threads += (uint64_t)g.nx * g.ny * ((units + w - 1) / w) * w;
            }
            if (threads <= best) { best = threads; L.bulkBlock.x = w; }
        }
        tx = (int)L.bulkBlock.x;
    }
    // Pair kernel with whole rows per workgroup: when every block's row fits one
    // workgroup of up to 256 threads, take the width that leaves the fewest idle
    // threads with each workgroup filled by as many whole rows as fit; used when
    // it idles fewer threads than the row segments (short rows of multi-block
    // lattices). PLBAMD_PACKED=0/1 overrides.
    bool packed = false;
    int packedW = 0;
    if (pairs && !getenv("PLBAMD_BULK_TX")) {
        uint64_t segThreads = 0;
        for (const BlockGeom& g : L.geom) segThreads += (uint64_t)g.nx * g.ny * ((g.nz / 2 + tx - 1) / tx) * tx;
        uint64_t best = UINT64_MAX;
        for (int w = 64; w <= 256; w += 64) {
            uint64_t threads = 0;
            bool fits = true;
            for (const BlockGeom& g : L.geom) {
                const int units = g.nz / 2;
                // at least two pairs per row: the low and the high halo come from two lanes
                if (units > w || units < 2) { fits = false; break; }
                const uint64_t rows = (uint64_t)g.nx * g.ny, r = (uint64_t)(w / units);
                threads += (rows + r - 1) / r * w;
            }
            if (fits && threads <= best) { best = threads; packedW = w; }
        }
        packed = packedW > 0 && best < segThreads;
        if (const char* e = getenv("PLBAMD_PACKED")) packed = packedW > 0 && atoi(e) != 0;
        if (packed) tx = packedW;
    }
    const int ty = (staged || pairs) ? 1 : std::max(1, txMax / tx);
    L.bulkBlock = dim3(tx, ty);
    L.zChunk = tx * (pairs ? 2 : cpt);
    L.staged = staged;
    L.pairs = pairs;
    L.packed = packed;
    L.zWrap = true;
    for (const BlockGeom& g : L.geom) L.zWrap = L.zWrap && g.pz != 0;
    L.ldsPitch = 0;
    if (packed) {
        for (BlockGeom& g : L.geom) {
            const int units = g.nz / 2;
            g.rowsPerWG = tx / units;
            L.ldsPitch = std::max(L.ldsPitch, g.rowsPerWG * (units + 2));
        }
    }
    L.pairsLds = (size_t)kZMoving * (packed ? L.ldsPitch : tx + 2) * 4 + 2 * (size_t)tx * sizeof(uint32_t);
    L.stagedLds = (size_t)kZMoving * (tx + 2) * std::max<size_t>(4, storageSize(d.storage)) + (size_t)tx * sizeof(uint32_t);
    if (L.stagedLds > 65536) L.staged = false;   // stay within the default LDS limit per workgroup
    uint64_t offset = 0, cellOffset = 0;
    uint64_t wg = 0;
    for (BlockGeom& g : L.geom) {
        const uint64_t cells = (uint64_t)g.nx * g.ny * g.nz;
        if (cells >= (1ull << 32)) return fail(PLBAMD_EINVAL, "lattice_create: at most 2^32 - 1 cells per block");
        g.slot = (cells + 63) / 64 * 64;
        g.offset = offset;
        g.cellOffset = cellOffset;
        g.nzChunks = (g.nz + L.zChunk - 1) / L.zChunk;
        g.nyGroups = (g.ny + ty - 1) / ty;
        g.wgBegin = (uint32_t)wg;
        if (L.packed)
            wg += ((uint64_t)g.nx * g.ny + g.rowsPerWG - 1) / g.rowsPerWG;
        else
            wg += (uint64_t)g.nx * g.nyGroups * g.nzChunks;
        if (wg >= (1ull << 31)) return fail(PLBAMD_EINVAL, "lattice_create: lattice too large for one launch");
        offset += g.slot * 19;
        cellOffset += cells;
    }
    L.allocCells = cellOffset;
    L.nBulkPartials = wg;
    // linear even-step decomposition: kEvenThreads threads, 1 or 2 cells each.
    // Used for 16-bit storage (pair accesses) and for multi-block lattices (short
    // rows); one-block FP64/FP32 lattices keep the row kernel, which measured
    // faster there (TG 512^3 FP64: 17.9 vs 17.4 GLUPS).
    L.linearEven = storageSize(d.storage) == 2 ? pairs : L.multi;
    if (const char* e = getenv("PLBAMD_LINEAR_EVEN")) L.linearEven = L.linearEven && atoi(e) != 0;
    {
        const uint64_t perWG = (uint64_t)kEvenThreads * (storageSize(d.storage) == 2 ? 2 : 1);
        uint64_t ew = 0;
        for (BlockGeom& g : L.geom) {
            g.evenBegin = (uint32_t)ew;
            ew += ((uint64_t)g.nx * g.ny * g.nz + perWG - 1) / perWG;
        }
        if (ew >= (1ull << 31)) L.linearEven = false;
        L.nEven = ew;
    }
    L.elemSize = storageSize(d.storage);
    const size_t bytes = offset * L.elemSize;
    if (hipMalloc(&L.arena, bytes) != hipSuccess) {
        (void)hipGetLastError();
        return fail(PLBAMD_ENOMEM, "lattice_create: cannot allocate " + std::to_string(bytes) + " bytes");
    }
    PLBAMD_HIP(hipMemset(L.arena, 0, bytes));
    PLBAMD_HIP(hipMalloc(&L.dGeom, L.geom.size() * sizeof(BlockGeom)));
    PLBAMD_HIP(hipMemcpy(L.dGeom, L.geom.data(), L.geom.size() * sizeof(BlockGeom), hipMemcpyHostToDevice));
    if (L.multi) {
        std::vector<uint16_t> wgBlock(L.nBulkPartials);
        for (size_t b = 0; b < L.geom.size(); ++b) {
            const uint64_t end = b + 1 < L.geom.size() ? L.geom[b + 1].wgBegin : L.nBulkPartials;
            std::fill(wgBlock.begin() + L.geom[b].wgBegin, wgBlock.begin() + end, (uint16_t)b);
        }
        PLBAMD_HIP(hipMalloc(&L.dWgBlock, wgBlock.size() * sizeof(uint16_t)));
        PLBAMD_HIP(hipMemcpy(L.dWgBlock, wgBlock.data(), wgBlock.size() * sizeof(uint16_t), hipMemcpyHostToDevice));
    }
    PLBAMD_HIP(hipMalloc(&L.dBulkPartials, std::max(L.nBulkPartials, L.nEven) * 16 * sizeof(double)));
    PLBAMD_HIP(hipMalloc(&L.dBulkLevel1, kStatsLevel1 * 16 * sizeof(double)));
    if (L.multi && L.linearEven) {
        std::vector<uint16_t> eb(L.nEven);
        for (size_t b = 0; b < L.geom.size(); ++b) {
            const uint64_t end = b + 1 < L.geom.size() ? L.geom[b + 1].evenBegin : L.nEven;
            std::fill(eb.begin() + L.geom[b].evenBegin, eb.begin() + end, (uint16_t)b);
        }
        PLBAMD_HIP(hipMalloc(&L.dEvenBlock, eb.size() * sizeof(uint16_t)));
        PLBAMD_HIP(hipMemcpy(L.dEvenBlock, eb.data(), eb.size() * sizeof(uint16_t), hipMemcpyHostToDevice));
    }
    PLBAMD_HIP(hipMalloc(&L.dStats, 16 * sizeof(double)));
    PLBAMD_HIP(hipMemset(L.dStats, 0, 16 * sizeof(double)));

    // Envelope segments: the allocated box of block b (or its bulk widened by
    // `width` cells, clipped to the allocation) intersected with the bulk of
    // every block c shifted by every periodic image (not b itself unshifted).
    auto buildSegments = [&](int width, std::vector<CopySegment>& segs, uint64_t& total) {
        segs.clear();
        total = 0;
        for (size_t bi = 0; bi < L.geom.size(); ++bi) {
            const BlockGeom& gb = L.geom[bi];
            const int blo[3] = {gb.bx0, gb.by0, gb.bz0}, bhi[3] = {gb.bx1, gb.by1, gb.bz1};
            const int ext[3] = {gb.nx, gb.ny, gb.nz}, org[3] = {gb.gx0, gb.gy0, gb.gz0};
            int alo[3], ahi[3];
            for (int a = 0; a < 3; ++a) {
                alo[a] = org[a] + std::max(0, blo[a] - width);
                ahi[a] = org[a] + std::min(ext[a] - 1, bhi[a] + width);
            }
            for (size_t ci = 0; ci < L.geom.size(); ++ci) {
                const plbamd_box& cb = L.bulk[ci];
                const BlockGeom& gc = L.geom[ci];
                for (int sx = -1; sx <= 1; ++sx)
                    for (int sy = -1; sy <= 1; ++sy)
                        for (int sz = -1; sz <= 1; ++sz) {
                            const int s[3] = {sx, sy, sz};
                            bool okShift = true;
                            for (int a = 0; a < 3; ++a) if (s[a] != 0 && !per[a]) okShift = false;
                            if (!okShift || (ci == bi && sx == 0 && sy == 0 && sz == 0)) continue;
                            const int clo[3] = {cb.x0 + sx * n[0], cb.y0 + sy * n[1], cb.z0 + sz * n[2]};
                            const int chi[3] = {cb.x1 + sx * n[0], cb.y1 + sy * n[1], cb.z1 + sz * n[2]};
                            int olo[3], ohi[3];
                            bool any = true;
                            for (int a = 0; a < 3; ++a) {
                                olo[a] = std::max(alo[a], clo[a]);
                                ohi[a] = std::min(ahi[a], chi[a]);
                                if (olo[a] > ohi[a]) any = false;
                            }
                            if (!any) continue;
                            CopySegment seg{};
                            seg.src = (uint16_t)ci;
                            seg.dst = (uint16_t)bi;
                            // source: bulk of c; local = global - shift * n - origin of c
                            seg.sx = olo[0] - sx * n[0] - gc.gx0;
                            seg.sy = olo[1] - sy * n[1] - gc.gy0;
                            seg.sz = olo[2] - sz * n[2] - gc.gz0;
                            seg.dx = olo[0] - gb.gx0;
                            seg.dy = olo[1] - gb.gy0;
                            seg.dz = olo[2] - gb.gz0;
                            seg.lx = ohi[0] - olo[0] + 1;
                            seg.ly = ohi[1] - olo[1] + 1;
                            seg.lz = ohi[2] - olo[2] + 1;
                            seg.cellBegin = total;
                            total += (uint64_t)seg.lx * seg.ly * seg.lz;
                            segs.push_back(seg);
                        }
            }
        }
    };
    // device copy of a segment list and its tasks (256 cells each)
    auto uploadSegments = [&](const std::vector<CopySegment>& segs, CopySegment** dSegs, CopyTask** dTasks,
                              uint32_t* nTasks) -> int {
        if (segs.empty()) return PLBAMD_OK;
        PLBAMD_HIP(hipMalloc(dSegs, segs.size() * sizeof(CopySegment)));
        PLBAMD_HIP(hipMemcpy(*dSegs, segs.data(), segs.size() * sizeof(CopySegment), hipMemcpyHostToDevice));
        std::vector<CopyTask> tasks;
        for (size_t k = 0; k < segs.size(); ++k) {
            const uint64_t cnt = (uint64_t)segs[k].lx * segs[k].ly * segs[k].lz;
            for (uint64_t f = 0; f < cnt; f += 256) tasks.push_back(CopyTask{(uint32_t)k, (uint32_t)f});
        }
        *nTasks = (uint32_t)tasks.size();
        PLBAMD_HIP(hipMalloc(dTasks, tasks.size() * sizeof(CopyTask)));
        PLBAMD_HIP(hipMemcpy(*dTasks, tasks.data(), tasks.size() * sizeof(CopyTask), hipMemcpyHostToDevice));
        return PLBAMD_OK;
    };
    if (L.multi) {
        std::vector<CopySegment> segs;
        uint64_t total = 0;
        buildSegments(1 << 30, segs, total);   // the whole allocation: both envelope layers
        L.nSegs = (int)segs.size();
        L.segCells = total;
        PLBAMD_TRY(uploadSegments(segs, &L.dSegs, &L.dTasks, &L.nTasks));
        buildSegments(1, segs, total);         // the first layer (the adjoint state)
        PLBAMD_TRY(uploadSegments(segs, &L.dSegs1, &L.dTasks1, &L.nTasks1));
    }
    PLBAMD_HIP(hipStreamCreateWithFlags(&L.stream, hipStreamNonBlocking));
    PLBAMD_HIP(hipEventCreate(&L.ev0));
    PLBAMD_HIP(hipEventCreate(&L.ev1));
    return PLBAMD_OK;
}

// Builds the device tags, the special-cell list and the porosity array of every
// block from the domain-wide description (tags, planes, porosity).
int rebuild(Lattice& L) {
    if (!L.dirty) return PLBAMD_OK;
    const int nx = L.desc.nx, ny = L.desc.ny, nz = L.desc.nz;
    const int per[3] = {L.desc.periodic[0] != 0, L.desc.periodic[1] != 0, L.desc.periodic[2] != 0};
    auto at = [&](int x, int y, int z) { return ((uint64_t)x * ny + y) * nz + z; };
    auto wrap = [&](int& v, int n, int p) {
        if (v < 0 || v >= n) { if (!p) return false; v = ((v % n) + n) % n; }
        return true;
    };
    const bool anyPlane = !L.planes.empty();
    std::vector<int> planeOf;
    std::vector<uint32_t> boxIndex;
    if (anyPlane) {
        planeOf.assign(L.cells, -1);
        boxIndex.assign(L.cells, 0);
        for (size_t p = 0; p < L.planes.size(); ++p) {
            const plbamd_box& b = L.planes[p].box;
            uint32_t k = 0;
            for (int x = b.x0; x <= b.x1; ++x)
                for (int y = b.y0; y <= b.y1; ++y)
                    for (int z = b.z0; z <= b.z1; ++z, ++k) {
                        planeOf[at(x, y, z)] = (int)p;
                        boxIndex[at(x, y, z)] = k;
                    }
        }
    }
    const bool haveTags = !L.hostTags.empty();
    auto tagOf = [&](uint64_t c) -> uint8_t { return haveTags ? L.hostTags[c] : (uint8_t)PLBAMD_TAG_BULK; };
    std::vector<uint32_t>& planeValueBase = L.planeValueBase;
    planeValueBase.assign(L.planes.size(), 0);
    std::vector<double> values;
    for (size_t p = 0; p < L.planes.size(); ++p) {
        planeValueBase[p] = (uint32_t)values.size();
        values.insert(values.end(), L.planes[p].values.begin(), L.planes[p].values.end());
    }
    std::vector<uint8_t> dev(L.allocCells, BT_BULK);
    std::vector<uint32_t> walls;   // allocated on the first cell next to a link wall
    std::vector<double> alpha;
    if (!L.hostAlpha.empty()) alpha.assign(L.allocCells, 0.0);
    std::vector<SpecialCell> special;
    bool anyNonBulk = anyPlane;
    L.anyWrapped = false;
    for (size_t bi = 0; bi < L.geom.size(); ++bi) {
        const BlockGeom& g = L.geom[bi];
        for (int lx = 0; lx < g.nx; ++lx)
            for (int ly = 0; ly < g.ny; ++ly)
                for (int lz = 0; lz < g.nz; ++lz) {
                    int x = g.gx0 + lx, y = g.gy0 + ly, z = g.gz0 + lz;
                    wrap(x, nx, per[0]); wrap(y, ny, per[1]); wrap(z, nz, per[2]);
                    const uint64_t c = at(x, y, z);
                    const uint64_t local = ((uint64_t)lx * g.ny + ly) * g.nz + lz;
                    const uint64_t slot = g.cellOffset + local;
                    if (!alpha.empty()) alpha[slot] = L.hostAlpha[c];
                    const uint8_t t = tagOf(c);
                    const int p = anyPlane ? planeOf[c] : -1;
                    if (t == PLBAMD_TAG_WALL_LINK) { dev[slot] = BT_SKIP; anyNonBulk = true; continue; }
                    uint32_t wall = 0, fluid = 0;
                    if (haveTags) {
                        for (int i = 1; i < 19; ++i) {
                            int ax = x + cx(i), ay = y + cy(i), az = z + cz(i);
                            const bool in = wrap(ax, nx, per[0]) && wrap(ay, ny, per[1]) && wrap(az, nz, per[2]);
                            const uint8_t nt = in ? tagOf(at(ax, ay, az)) : (uint8_t)PLBAMD_TAG_BULK;
                            if (in && nt == PLBAMD_TAG_WALL_LINK) wall |= 1u << i;
                            // bounceBackModels3D.hh:48-65: fluid unless BB / MEBB / NoDynamics;
                            // beyond a non-periodic edge the envelope holds the background dynamics.
                            if (nt != PLBAMD_TAG_MOMENTUM_EXCHANGE && nt != PLBAMD_TAG_BOUNCE_BACK &&
                                nt != PLBAMD_TAG_NO_DYNAMICS && nt != PLBAMD_TAG_WALL_LINK)
                                fluid |= 1u << i;
                        }
                    }
                    const bool bulk = lx >= g.bx0 && lx <= g.bx1 && ly >= g.by0 && ly <= g.by1 &&
                                      lz >= g.bz0 && lz <= g.bz1;
                    SpecialCell s{};
                    s.cell = (uint32_t)local;
                    s.x = (uint16_t)lx; s.y = (uint16_t)ly; s.z = (uint16_t)lz;
                    s.block = (uint16_t)bi;
                    s.wall = wall;
                    const uint16_t bulkBit = (uint16_t)((bulk ? 1 : 0) << 8);
                    if (p >= 0) {
                        const Plane& P = L.planes[p];
                        s.info = (uint16_t)(SK_PLANE | (P.family << 2) | (P.kind << 3) | (P.dir << 4) |
                                            ((P.ori > 0 ? 1 : 0) << 6) | ((P.wrapped ? 1 : 0) << 7)) | bulkBit;
                        s.value = planeValueBase[p] + boxIndex[c] * (P.kind == PLBAMD_VELOCITY ? 3u : 1u);
                        if (P.wrapped) L.anyWrapped = true;
                        dev[slot] = BT_SKIP;
                        special.push_back(s);
                    } else if (t == PLBAMD_TAG_MOMENTUM_EXCHANGE) {
                        s.info = (uint16_t)SK_MOMENTUM_EXCHANGE | bulkBit;
                        s.fluid = fluid;
                        dev[slot] = BT_SKIP;
                        special.push_back(s);
                        anyNonBulk = true;
                    } else if (t == PLBAMD_TAG_BOUNCE_BACK || t == PLBAMD_TAG_NO_DYNAMICS) {
                        // node dynamics keep their full-way semantics next to a link wall
                        dev[slot] = (t == PLBAMD_TAG_BOUNCE_BACK) ? BT_BOUNCE_BACK : BT_NO_DYNAMICS;
                        anyNonBulk = true;
                    } else if (wall) {
                        // background dynamics with the link mask, inside the bulk kernels
                        if (walls.empty()) walls.assign(L.allocCells, 0u);
                        walls[slot] = wall;
                        dev[slot] = BT_WALL_NEAR;
                        anyNonBulk = true;
                    }
                }
    }
    hipFree(L.dTags); L.dTags = nullptr;
    hipFree(L.dSpecial); L.dSpecial = nullptr;
    hipFree(L.dValues); L.dValues = nullptr;
    hipFree(L.dSpecialPartials); L.dSpecialPartials = nullptr;
    hipFree(L.dAlpha); L.dAlpha = nullptr;
    hipFree(L.dWalls); L.dWalls = nullptr;
    L.nSpecial = (uint32_t)special.size();
    if (anyNonBulk) {
        PLBAMD_HIP(hipMalloc(&L.dTags, L.allocCells));
        PLBAMD_HIP(hipMemcpy(L.dTags, dev.data(), L.allocCells, hipMemcpyHostToDevice));
    }
    if (!walls.empty()) {
        PLBAMD_HIP(hipMalloc(&L.dWalls, walls.size() * sizeof(uint32_t)));
        PLBAMD_HIP(hipMemcpy(L.dWalls, walls.data(), walls.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    }
    if (!special.empty()) {
        PLBAMD_HIP(hipMalloc(&L.dSpecial, special.size() * sizeof(SpecialCell)));
        PLBAMD_HIP(hipMemcpy(L.dSpecial, special.data(), special.size() * sizeof(SpecialCell), hipMemcpyHostToDevice));
        L.nSpecialPartials = (special.size() + 255) / 256;
        PLBAMD_HIP(hipMalloc(&L.dSpecialPartials, L.nSpecialPartials * 16 * sizeof(double)));
    }
    if (!values.empty()) {
        PLBAMD_HIP(hipMalloc(&L.dValues, values.size() * sizeof(double)));
        PLBAMD_HIP(hipMemcpy(L.dValues, values.data(), values.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    if (!alpha.empty()) {
        PLBAMD_HIP(hipMalloc(&L.dAlpha, alpha.size() * sizeof(double)));
        PLBAMD_HIP(hipMemcpy(L.dAlpha, alpha.data(), alpha.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    L.dirty = false;
    return PLBAMD_OK;
}

template<typename S>
int exchange(Lattice& L) {
    if (!L.multi || L.nSegs == 0) return PLBAMD_OK;
    hipLaunchKernelGGL((exchange_envelopes<S>), dim3(L.nTasks), dim3(256), 0, L.stream, L.arena, L.dGeom, L.dSegs,
                       L.dTasks);
    PLBAMD_HIP(hipGetLastError());
    return PLBAMD_OK;
}

// One collide-and-stream step of every block. With `record` (FP64 BGK only,
// statistics off), the kernels also write the pre-collision moments of every
// cell and the pre-completion populations of every special record.
struct Recording {
    void* moments = nullptr;       // MomentType<S> [4][momStride]
    uint64_t momStride = 0;
    double* specialF = nullptr;
};

template<typename S, int DYN>
int launchStepTyped(Lattice& L, bool odd, bool stats, const Recording* record = nullptr) {
    StepArgs a{};
    a.arena = L.arena;
    a.blocks = L.dGeom;
    a.nBlocks = (int)L.geom.size();
    a.zChunk = L.zChunk;
    a.omega = L.desc.omega;
    for (int c = 0; c < 3; ++c) a.force[c] = L.desc.force[c];
    a.tags = L.dTags;
    a.alpha = L.dAlpha;
    a.partials = L.dBulkPartials;
    a.xcdRemap = L.xcdRemap;
    a.wgBlock = L.dWgBlock;
    a.walls = L.dWalls;
    a.ldsPitch = L.ldsPitch;
    const bool forced = L.dAlpha || L.desc.force[0] != 0 || L.desc.force[1] != 0 || L.desc.force[2] != 0;
    const bool tags = L.dTags != nullptr;
    const dim3 grid((unsigned)L.nBulkPartials);
    if (record && stats) return fail(PLBAMD_EINVAL, "recording step: statistics must be off");
    if (!odd && L.linearEven) {
        // even steps in memory order (bulk_even_linear)
        a.evenBlock = L.dEvenBlock;
        const dim3 eg((unsigned)L.nEven), eb(kEvenThreads);
        if (record) {
            a.moments = record->moments;
            a.momStride = record->momStride;
            if constexpr (std::is_same<S, double>::value || sizeof(typename Store<S>::R) == 2) {
#define PLBAMD_EVR(TG, FO) hipLaunchKernelGGL((bulk_even_linear<S, DYN, TG, false, FO, true>), eg, eb, 0, L.stream, a)
                if (tags) { if (forced) PLBAMD_EVR(true, true); else PLBAMD_EVR(true, false); }
                else { if (forced) PLBAMD_EVR(false, true); else PLBAMD_EVR(false, false); }
#undef PLBAMD_EVR
            } else {
                return fail(PLBAMD_EINVAL, "recording step: FP64 or 16-bit storage");
            }
        } else {
#define PLBAMD_EV(TG, ST, FO) hipLaunchKernelGGL((bulk_even_linear<S, DYN, TG, ST, FO>), eg, eb, 0, L.stream, a)
#define PLBAMD_EV_F(TG, ST) if (forced) PLBAMD_EV(TG, ST, true); else PLBAMD_EV(TG, ST, false)
#define PLBAMD_EV_S(TG) if (stats) { PLBAMD_EV_F(TG, true); } else { PLBAMD_EV_F(TG, false); }
            if (tags) { PLBAMD_EV_S(true) } else { PLBAMD_EV_S(false) }
#undef PLBAMD_EV_S
#undef PLBAMD_EV_F
#undef PLBAMD_EV
        }
    } else if (record) {
        a.moments = record->moments;
        a.momStride = record->momStride;
        if constexpr (std::is_same<S, double>::value) {
#define PLBAMD_REC(O, TG, FO)                                                                                   \
    do {                                                                                                        \
        if (O && L.staged)                                                                                      \
            hipLaunchKernelGGL((bulk_odd_staged<double, DYN, TG, false, FO, true>), grid, L.bulkBlock,           \
                               L.stagedLds, L.stream, a);                                                       \
        else                                                                                                    \
            hipLaunchKernelGGL((bulk_step<double, DYN, O, TG, false, FO, true>), grid, L.bulkBlock, 0, L.stream, a); \
    } while (0)
#define PLBAMD_REC_F(O, TG) if (forced) PLBAMD_REC(O, TG, true); else PLBAMD_REC(O, TG, false)
#define PLBAMD_REC_T(O) if (tags) { PLBAMD_REC_F(O, true); } else { PLBAMD_REC_F(O, false); }
            if (odd) { PLBAMD_REC_T(true) } else { PLBAMD_REC_T(false) }
#undef PLBAMD_REC_T
#undef PLBAMD_REC_F
#undef PLBAMD_REC
        } else if constexpr (sizeof(typename Store<S>::R) == 2) {
            if (!L.pairs) return fail(PLBAMD_EINVAL, "recording step: 16-bit storage needs every block even along z");
#define PLBAMD_RECP(O, TG, FO)                                                                                  \
    do {                                                                                                        \
        if (L.packed)                                                                                           \
            hipLaunchKernelGGL((bulk_pairs<S, DYN, O, TG, false, FO, true, true>), grid, L.bulkBlock,           \
                               O ? L.pairsLds : 0, L.stream, a);                                                \
        else                                                                                                    \
            hipLaunchKernelGGL((bulk_pairs<S, DYN, O, TG, false, FO, true>), grid, L.bulkBlock,                 \
                               O ? L.pairsLds : 0, L.stream, a);                                                \
    } while (0)
#define PLBAMD_RECP_F(O, TG) if (forced) PLBAMD_RECP(O, TG, true); else PLBAMD_RECP(O, TG, false)
#define PLBAMD_RECP_T(O) if (tags) { PLBAMD_RECP_F(O, true); } else { PLBAMD_RECP_F(O, false); }
            if (odd) { PLBAMD_RECP_T(true) } else { PLBAMD_RECP_T(false) }
#undef PLBAMD_RECP_T
#undef PLBAMD_RECP_F
#undef PLBAMD_RECP
        } else {
            return fail(PLBAMD_EINVAL, "recording step: FP64 or 16-bit storage");
        }
    } else if (L.pairs) {
        if constexpr (sizeof(typename Store<S>::R) == 2) {
#define PLBAMD_PAIRS(O, TG, ST, FO)                                                                             \
    do {                                                                                                        \
        if (L.packed)                                                                                           \
            hipLaunchKernelGGL((bulk_pairs<S, DYN, O, TG, ST, FO, false, true>), grid, L.bulkBlock,             \
                               O ? L.pairsLds : 0, L.stream, a);                                                \
        else if (O && L.zWrap)                                                                                  \
            hipLaunchKernelGGL((bulk_pairs<S, DYN, O, TG, ST, FO, false, false, true>), grid, L.bulkBlock,      \
                               L.pairsLds, L.stream, a);                                                        \
        else                                                                                                    \
            hipLaunchKernelGGL((bulk_pairs<S, DYN, O, TG, ST, FO>), grid, L.bulkBlock, O ? L.pairsLds : 0,      \
                               L.stream, a);                                                                    \
    } while (0)
#define PLBAMD_PAIRS_F(O, TG, ST) if (forced) PLBAMD_PAIRS(O, TG, ST, true); else PLBAMD_PAIRS(O, TG, ST, false)
#define PLBAMD_PAIRS_S(O, TG) if (stats) { PLBAMD_PAIRS_F(O, TG, true); } else { PLBAMD_PAIRS_F(O, TG, false); }
#define PLBAMD_PAIRS_T(O) if (tags) { PLBAMD_PAIRS_S(O, true) } else { PLBAMD_PAIRS_S(O, false) }
            if (odd) { PLBAMD_PAIRS_T(true) } else { PLBAMD_PAIRS_T(false) }
#undef PLBAMD_PAIRS_T
#undef PLBAMD_PAIRS_S
#undef PLBAMD_PAIRS_F
#undef PLBAMD_PAIRS
        } else {
            return fail(PLBAMD_EINVAL, "pair kernel: 16-bit storage only");
        }
    } else {
#define PLBAMD_BULK(O, TG, ST, FO)                                                                              \
    do {                                                                                                        \
        if (O && L.staged)                                                                                      \
            hipLaunchKernelGGL((bulk_odd_staged<S, DYN, TG, ST, FO>), grid, L.bulkBlock, L.stagedLds, L.stream,  \
                               a);                                                                              \
        else                                                                                                    \
            hipLaunchKernelGGL((bulk_step<S, DYN, O, TG, ST, FO>), grid, L.bulkBlock, 0, L.stream, a);         \
    } while (0)
#define PLBAMD_BULK_F(O, TG, ST) if (forced) PLBAMD_BULK(O, TG, ST, true); else PLBAMD_BULK(O, TG, ST, false)
#define PLBAMD_BULK_S(O, TG) if (stats) { PLBAMD_BULK_F(O, TG, true); } else { PLBAMD_BULK_F(O, TG, false); }
#define PLBAMD_BULK_T(O) if (tags) { PLBAMD_BULK_S(O, true) } else { PLBAMD_BULK_S(O, false) }
        if (odd) { PLBAMD_BULK_T(true) } else { PLBAMD_BULK_T(false) }
#undef PLBAMD_BULK_T
#undef PLBAMD_BULK_S
#undef PLBAMD_BULK_F
#undef PLBAMD_BULK
    }
    PLBAMD_HIP(hipGetLastError());
    if (L.nSpecial) {
        SpecialArgs s{};
        s.arena = L.arena; s.blocks = L.dGeom; s.omega = L.desc.omega;
        for (int c = 0; c < 3; ++c) s.force[c] = L.desc.force[c];
        s.alpha = L.dAlpha;
        s.cells = L.dSpecial; s.count = L.nSpecial; s.values = L.dValues;
        s.partials = L.dSpecialPartials;
        if (record) {
            s.moments = record->moments;
            s.momStride = record->momStride;
            s.dumpF = record->specialF;
        }
        dim3 sg((L.nSpecial + 255) / 256), sb(256);
#define PLBAMD_SPECIAL(O, ST, FO) hipLaunchKernelGGL((special_step<S, DYN, O, ST, FO>), sg, sb, 0, L.stream, s)
#define PLBAMD_SPECIAL_F(O, ST) if (forced) PLBAMD_SPECIAL(O, ST, true); else PLBAMD_SPECIAL(O, ST, false)
        if (odd) {
            if (stats) { PLBAMD_SPECIAL_F(true, true); } else { PLBAMD_SPECIAL_F(true, false); }
        } else {
            if (stats) { PLBAMD_SPECIAL_F(false, true); } else { PLBAMD_SPECIAL_F(false, false); }
        }
#undef PLBAMD_SPECIAL_F
#undef PLBAMD_SPECIAL
        PLBAMD_HIP(hipGetLastError());
    }
    if (stats) {
        // two levels when there are many records (a single workgroup reading
        // 65536 records took 40 % of a 256^3 step): kStatsLevel1 workgroups in a
        // fixed order, then one
        const uint64_t nb = (!odd && L.linearEven) ? L.nEven : L.nBulkPartials;
        const double* bulk = L.dBulkPartials;
        uint64_t nr = nb;
        if (nb > 4 * kStatsLevel1) {
            hipLaunchKernelGGL(reduce_bulk_records, dim3(kStatsLevel1), dim3(256), 0, L.stream, L.dBulkPartials, nb,
                               L.dBulkLevel1);
            bulk = L.dBulkLevel1;
            nr = kStatsLevel1;
        }
        hipLaunchKernelGGL(reduce_partials, dim3(1), dim3(256), 0, L.stream, bulk, nr,
                           L.dSpecialPartials, L.nSpecial ? L.nSpecialPartials : 0, L.dStats);
        PLBAMD_HIP(hipGetLastError());
    }
    if (odd) PLBAMD_TRY(exchange<S>(L));
    return PLBAMD_OK;
}

template<typename S>
int launchStep(Lattice& L, bool odd, bool stats) {
    return L.desc.dynamics == PLBAMD_BGK ? launchStepTyped<S, 0>(L, odd, stats)
                                         : launchStepTyped<S, 1>(L, odd, stats);
}

int step(Lattice& L, int nsteps) {
    PLBAMD_TRY(rebuild(L));
    for (int s = 0; s < nsteps; ++s) {
        const bool odd = (L.time & 1) != 0;
        const bool stats = L.statsOn && s == nsteps - 1;
        int rc;
        switch (L.desc.storage) {
        case PLBAMD_F64: rc = launchStep<double>(L, odd, stats); break;
        case PLBAMD_F32: rc = launchStep<float>(L, odd, stats); break;
        case PLBAMD_F16: rc = launchStep<f16s>(L, odd, stats); break;
        default: rc = launchStep<bf16s>(L, odd, stats); break;
        }
        if (rc) return rc;
        L.time += 1;
        if (stats) L.statsValid = true;
    }
    return PLBAMD_OK;
}

// Transfer arguments for block bi and the part of the global box `g` it owns.
bool boxArgs(const Lattice& L, size_t bi, const plbamd_box& g, BoxArgs& a) {
    plbamd_box I;
    if (!intersect(g, L.bulk[bi], I)) return false;
    const BlockGeom& bg = L.geom[bi];
    a = BoxArgs{};
    a.g = bg;
    a.block = (int)bi;
    a.parity = (int)(L.time & 1);
    a.ix0 = I.x0 - bg.gx0; a.iy0 = I.y0 - bg.gy0; a.iz0 = I.z0 - bg.gz0;
    a.ilx = I.x1 - I.x0 + 1; a.ily = I.y1 - I.y0 + 1; a.ilz = I.z1 - I.z0 + 1;
    a.ox = I.x0 - g.x0; a.oy = I.y0 - g.y0; a.oz = I.z0 - g.z0;
    a.sly = g.y1 - g.y0 + 1; a.slz = g.z1 - g.z0 + 1;
    a.special = L.dSpecial;
    a.nSpecial = L.nSpecial;
    a.walls = L.dWalls;
    return true;
}

// Populations of box b between the lattice and a device buffer of boxCells(b)
// * 19 elements of type elem, C order; launches only (no host copy, no wait).
template<typename S>
hipError_t transferDevice(Lattice& L, const plbamd_box& b, int elem, void* staging, bool up) {
    hipError_t err = hipSuccess;
    for (size_t bi = 0; bi < L.geom.size() && err == hipSuccess; ++bi) {
        BoxArgs a;
        if (!boxArgs(L, bi, b, a)) continue;
        const dim3 grid = gridFor((uint64_t)a.ilx * a.ily * a.ilz), block(256);
        if (up) {
            if (elem == PLBAMD_F64) hipLaunchKernelGGL((upload_box<S, double>), grid, block, 0, L.stream, L.arena, a, (const double*)staging);
            else if (elem == PLBAMD_F32) hipLaunchKernelGGL((upload_box<S, float>), grid, block, 0, L.stream, L.arena, a, (const float*)staging);
            else hipLaunchKernelGGL((upload_box_raw<S>), grid, block, 0, L.stream, L.arena, a, (const uint16_t*)staging);
        } else {
            if (elem == PLBAMD_F64) hipLaunchKernelGGL((download_box<S, double>), grid, block, 0, L.stream, L.arena, a, (double*)staging);
            else if (elem == PLBAMD_F32) hipLaunchKernelGGL((download_box<S, float>), grid, block, 0, L.stream, L.arena, a, (float*)staging);
            else hipLaunchKernelGGL((download_box_raw<S>), grid, block, 0, L.stream, L.arena, a, (uint16_t*)staging);
            err = hipGetLastError();
            if (err == hipSuccess && L.anyWrapped && L.nSpecial) {
                const dim3 g2 = gridFor(L.nSpecial);
                if (elem == PLBAMD_F64)
                    hipLaunchKernelGGL((complete_download<S, double>), g2, block, 0, L.stream, L.arena, a, L.dValues, (double*)staging, 0);
                else if (elem == PLBAMD_F32)
                    hipLaunchKernelGGL((complete_download<S, float>), g2, block, 0, L.stream, L.arena, a, L.dValues, (float*)staging, 0);
                else
                    hipLaunchKernelGGL((complete_download<S, float>), g2, block, 0, L.stream, L.arena, a, L.dValues, (float*)staging, 1);
            }
        }
        if (err == hipSuccess) err = hipGetLastError();
    }
    return err;
}

template<typename S>
int transfer(Lattice& L, const plbamd_box& b, int elem, void* data, size_t bytes, bool up) {
    const uint64_t n = boxCells(b);
    const size_t hostElem = (elem == PLBAMD_F64) ? 8 : (elem == PLBAMD_F32 ? 4 : 2);
    if (bytes != n * 19 * hostElem) return fail(PLBAMD_EINVAL, "transfer: byte count does not match the box");
    if ((elem == PLBAMD_F16 || elem == PLBAMD_BF16) && elem != L.desc.storage)
        return fail(PLBAMD_EINVAL, "transfer: 16-bit host data must match the storage type");
    if (up && L.multi && (L.time & 1))
        return fail(PLBAMD_ESTATE, "upload: a multi-block lattice accepts populations after an even number of steps");
    PLBAMD_TRY(rebuild(L));
    void* staging = nullptr;
    PLBAMD_HIP(hipMalloc(&staging, bytes));
    hipError_t err = hipSuccess;
    if (up) err = hipMemcpyAsync(staging, data, bytes, hipMemcpyHostToDevice, L.stream);
    if (err == hipSuccess) err = transferDevice<S>(L, b, elem, staging, up);
    if (err == hipSuccess && up && exchange<S>(L) != PLBAMD_OK) err = hipErrorUnknown;
    if (err == hipSuccess && !up) err = hipMemcpyAsync(data, staging, bytes, hipMemcpyDeviceToHost, L.stream);
    if (err == hipSuccess) err = hipStreamSynchronize(L.stream);
    hipFree(staging);
    if (err != hipSuccess) return fail(PLBAMD_EHIP, std::string("transfer: ") + hipGetErrorString(err));
    return PLBAMD_OK;
}

// ---------------------------------------------------------------------------
// Palabos checkpoint files (saveBinaryBlock / loadBinaryBlock,
// io/serializerIO_3D.cpp:37-66): the cells are serialized and base64-coded on
// the device; the host writes or reads the text.

std::string base64Host(const uint8_t* p, size_t n) {
    std::string out;
    for (size_t k = 0; k < n; k += 3) {
        const uint32_t b0 = p[k], b1 = k + 1 < n ? p[k + 1] : 0u, b2 = k + 2 < n ? p[k + 2] : 0u;
        const uint32_t g = (b0 << 16) | (b1 << 8) | b2;
        out += b64_char(g >> 18);
        out += b64_char((g >> 12) & 63);
        out += k + 1 < n ? b64_char((g >> 6) & 63) : '=';
        out += k + 2 < n ? b64_char(g & 63) : '=';
    }
    return out;
}

// Decodes the n bytes of 4 * ceil(n / 3) characters; false on a character
// outside the alphabet where the bytes need one.
bool unbase64Host(const char* s, size_t n, uint8_t* out) {
    for (size_t k = 0; k < n; k += 3, s += 4) {
        const size_t take = n - k < 3 ? n - k : 3;
        uint32_t v[4];
        for (int c = 0; c < 4; ++c) v[c] = b64_value((uint8_t)s[c]);
        for (size_t c = 0; c <= take; ++c) if (v[c] > 63) return false;
        const uint32_t x = ((v[0] & 63) << 18) | ((v[1] & 63) << 12) | ((v[2] & 63) << 6) | (v[3] & 63);
        out[k] = (uint8_t)(x >> 16);
        if (take > 1) out[k + 1] = (uint8_t)((x >> 8) & 255u);
        if (take > 2) out[k + 2] = (uint8_t)(x & 255u);
    }
    return true;
}

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

static dim3 itemsGrid(uint64_t n) { return dim3((unsigned)std::max<uint64_t>(1, (n + 255) / 256)); }

template<typename S, typename T>
int checkpointTyped(Lattice& L, const char* path, int elem, bool force, bool enforceUint, bool save) {
    const plbamd_box whole{0, L.desc.nx - 1, 0, L.desc.ny - 1, 0, L.desc.nz - 1};
    const uint64_t cells = boxCells(whole);
    const int recLen = force ? 22 : 19;
    const uint64_t dataBytes = cells * recLen * sizeof(T);
    const uint64_t nChars = (dataBytes + 2) / 3 * 4;
    if (enforceUint && dataBytes > 0xffffffffull)
        return fail(PLBAMD_EINVAL, "binary block: the data does not fit a 32-bit size header (enforce_uint)");
    if (dataBytes / kB64Bytes + 1 > (uint64_t)INT32_MAX * 256)
        return fail(PLBAMD_EINVAL, "binary block: lattice too large for one conversion launch");
    PLBAMD_TRY(rebuild(L));
    DeviceBuffers buf;
    T* pops = static_cast<T*>(buf.get(cells * 19 * sizeof(T)));
    T* rec = force ? static_cast<T*>(buf.get(dataBytes)) : pops;
    char* text = static_cast<char*>(buf.get((nChars + 63) / 64 * 64));
    int* bad = static_cast<int*>(buf.get(2 * sizeof(int)));
    if (!pops || !rec || !text || !bad) return fail(PLBAMD_ENOMEM, "binary block: cannot allocate device buffers");
    const T fx = (T)L.desc.force[0], fy = (T)L.desc.force[1], fz = (T)L.desc.force[2];
    if (save) {
        // cells in no block are zeros, as the reference's forward ordering writes them
        PLBAMD_HIP(hipMemsetAsync(pops, 0, cells * 19 * sizeof(T), L.stream));
        PLBAMD_HIP(transferDevice<S>(L, whole, elem, pops, false));
        if (force) hipLaunchKernelGGL((append_force<T>), itemsGrid(cells), dim3(256), 0, L.stream, pops, rec, cells, fx, fy, fz);
        hipLaunchKernelGGL(base64_encode, itemsGrid(dataBytes / kB64Bytes + 1), dim3(256), 0, L.stream,
                           reinterpret_cast<const uint8_t*>(rec), dataBytes, text);
        PLBAMD_HIP(hipGetLastError());
        std::vector<char> host(nChars);
        PLBAMD_HIP(hipMemcpyAsync(host.data(), text, nChars, hipMemcpyDeviceToHost, L.stream));
        PLBAMD_HIP(hipStreamSynchronize(L.stream));
        uint8_t header[8];
        const uint64_t size64 = dataBytes;
        std::memcpy(header, &size64, 8);   // little-endian host
        const std::string head = base64Host(header, enforceUint ? 4 : 8);
        FILE* fp = std::fopen(path, "wb");
        if (!fp) return fail(PLBAMD_EINVAL, std::string("save_binary_block: cannot open ") + path);
        const bool ok = std::fwrite(head.data(), 1, head.size(), fp) == head.size() &&
                        std::fwrite(host.data(), 1, host.size(), fp) == host.size();
        if (std::fclose(fp) != 0 || !ok) return fail(PLBAMD_EINVAL, std::string("save_binary_block: cannot write ") + path);
        return PLBAMD_OK;
    }
    if (L.multi && (L.time & 1))
        return fail(PLBAMD_ESTATE, "load_binary_block: a multi-block lattice accepts populations after an even number of steps");
    std::string file;
    {
        FILE* fp = std::fopen(path, "rb");
        if (!fp) return fail(PLBAMD_EINVAL, std::string("load_binary_block: cannot open ") + path);
        if (std::fseek(fp, 0, SEEK_END) == 0) {
            const long size = std::ftell(fp);
            if (size > 0) file.reserve((size_t)size);
            std::rewind(fp);
        }
        char chunk[1 << 16];
        size_t got;
        while ((got = std::fread(chunk, 1, sizeof(chunk), fp)) > 0) file.append(chunk, got);
        std::fclose(fp);
    }
    // the reference reads with operator>>, which skips white space (every
    // white-space character is at most ' '; every base64 character above it)
    if (std::any_of(file.begin(), file.end(), [](char c) { return (unsigned char)c <= ' '; }))
        file.erase(std::remove_if(file.begin(), file.end(), [](char c) { return std::isspace((unsigned char)c) != 0; }),
                   file.end());
    const size_t headChars = enforceUint ? 8 : 12;
    if (file.size() < headChars) return fail(PLBAMD_EINVAL, "load_binary_block: file shorter than its size header");
    uint8_t header[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (!unbase64Host(file.data(), enforceUint ? 4 : 8, header))
        return fail(PLBAMD_EINVAL, "load_binary_block: the size header is not base64");
    uint64_t size64 = 0;
    std::memcpy(&size64, header, 8);
    if (size64 != dataBytes)
        return fail(PLBAMD_EINVAL, "load_binary_block: the file holds " + std::to_string(size64) +
                                       " bytes of cells, this lattice takes " + std::to_string(dataBytes));
    if (file.size() - headChars < nChars) return fail(PLBAMD_EINVAL, "load_binary_block: the file is truncated");
    PLBAMD_HIP(hipMemsetAsync(bad, 0, 2 * sizeof(int), L.stream));
    PLBAMD_HIP(hipMemcpyAsync(text, file.data() + headChars, nChars, hipMemcpyHostToDevice, L.stream));
    hipLaunchKernelGGL(base64_decode, itemsGrid(dataBytes / kB64Bytes + 1), dim3(256), 0, L.stream, text, dataBytes,
                       reinterpret_cast<uint8_t*>(rec), bad);
    if (force) hipLaunchKernelGGL((split_force<T>), itemsGrid(cells), dim3(256), 0, L.stream, rec, pops, cells, fx, fy, fz, bad + 1);
    PLBAMD_HIP(hipGetLastError());
    int flags[2] = {0, 0};
    PLBAMD_HIP(hipMemcpyAsync(flags, bad, sizeof(flags), hipMemcpyDeviceToHost, L.stream));
    PLBAMD_HIP(hipStreamSynchronize(L.stream));
    if (flags[0]) return fail(PLBAMD_EINVAL, "load_binary_block: the cell data is not base64");
    if (flags[1]) return fail(PLBAMD_EINVAL, "load_binary_block: the force externals differ from the lattice's force");
    PLBAMD_HIP(transferDevice<S>(L, whole, elem, pops, true));
    PLBAMD_TRY(exchange<S>(L));
    PLBAMD_HIP(hipStreamSynchronize(L.stream));
    return PLBAMD_OK;
}

template<typename S>
int checkpoint(Lattice& L, const char* path, int elem, bool force, bool enforceUint, bool save) {
    if (elem == PLBAMD_F64) return checkpointTyped<S, double>(L, path, elem, force, enforceUint, save);
    return checkpointTyped<S, float>(L, path, elem, force, enforceUint, save);
}

template<typename S>
int equilibrium(Lattice& L, const plbamd_box& b, const double* rho, const double* u) {
    const uint64_t n = boxCells(b);
    if (L.multi && (L.time & 1))
        return fail(PLBAMD_ESTATE, "initialize_at_equilibrium: a multi-block lattice accepts populations after an even number of steps");
    PLBAMD_TRY(rebuild(L));
    double* d = nullptr;
    PLBAMD_HIP(hipMalloc(&d, n * 4 * sizeof(double)));
    hipError_t err = hipMemcpyAsync(d, rho, n * sizeof(double), hipMemcpyHostToDevice, L.stream);
    if (err == hipSuccess) err = hipMemcpyAsync(d + n, u, 3 * n * sizeof(double), hipMemcpyHostToDevice, L.stream);
    for (size_t bi = 0; bi < L.geom.size() && err == hipSuccess; ++bi) {
        BoxArgs a;
        if (!boxArgs(L, bi, b, a)) continue;
        hipLaunchKernelGGL((equilibrium_box<S>), gridFor((uint64_t)a.ilx * a.ily * a.ilz), dim3(256), 0, L.stream,
                           L.arena, a, d, d + n);
        err = hipGetLastError();
    }
    if (err == hipSuccess && exchange<S>(L) != PLBAMD_OK) err = hipErrorUnknown;
    if (err == hipSuccess) err = hipStreamSynchronize(L.stream);
    hipFree(d);
    if (err != hipSuccess) return fail(PLBAMD_EHIP, std::string("equilibrium: ") + hipGetErrorString(err));
    return PLBAMD_OK;
}

template<typename S>
int moments(Lattice& L, const plbamd_box& b, double* rho, double* u, double* sums, double* fields = nullptr) {
    const uint64_t n = boxCells(b);
    PLBAMD_TRY(rebuild(L));
    std::vector<BoxArgs> parts;
    uint64_t nRecords = 0;
    const bool planes = L.nSpecial > 0 && L.dValues != nullptr;
    for (size_t bi = 0; bi < L.geom.size(); ++bi) {
        BoxArgs a;
        if (boxArgs(L, bi, b, a)) {
            parts.push_back(a);
            nRecords += gridFor((uint64_t)a.ilx * a.ily * a.ilz).x;
            if (planes) nRecords += gridFor(L.nSpecial).x;
        }
    }
    const size_t fieldDoubles = (rho || u) ? n * 4 : 0;
    const size_t analysisDoubles = fields ? n * 7 : 0;
    // records reduced in two levels when there are many: 256 workgroups, then one
    const uint32_t level1 = nRecords > 4096 ? 256u : 0u;
    double* d = nullptr;
    PLBAMD_HIP(hipMalloc(&d, (fieldDoubles + analysisDoubles + nRecords * 8 + 8 * level1 + 8) * sizeof(double)));
    double* dRho = rho ? d : nullptr;
    double* dU = u ? d + n : nullptr;
    double* dFields = fields ? d + fieldDoubles : nullptr;
    double* dPart = d + fieldDoubles + analysisDoubles;
    double* dLevel1 = dPart + nRecords * 8;
    double* dSum = dLevel1 + 8 * level1;
    MomentsArgs m{};
    m.tags = L.dTags;
    m.values = L.dValues;
    for (int c = 0; c < 3; ++c) m.force[c] = L.desc.force[c];
    m.forced = L.desc.force[0] != 0 || L.desc.force[1] != 0 || L.desc.force[2] != 0;
    m.fields = dFields;
    for (int k = 0; k < 3; ++k) m.solidRho[k] = L.solidRho[k];
    hipError_t err = hipSuccess;
    uint64_t rec = 0;
    for (const BoxArgs& a : parts) {
        const dim3 grid = gridFor((uint64_t)a.ilx * a.ily * a.ilz);
        if (a.parity)
            hipLaunchKernelGGL((moments_box<S, 1>), grid, dim3(256), 0, L.stream, L.arena, a, m, dRho, dU, dPart + rec * 8);
        else
            hipLaunchKernelGGL((moments_box<S, 0>), grid, dim3(256), 0, L.stream, L.arena, a, m, dRho, dU, dPart + rec * 8);
        rec += grid.x;
        err = hipGetLastError();
        if (err != hipSuccess) break;
        if (planes) {
            const dim3 pg = gridFor(L.nSpecial);
            hipLaunchKernelGGL((moments_planes<S>), pg, dim3(256), 0, L.stream, L.arena, a, m, dRho, dU, dPart + rec * 8);
            rec += pg.x;
            err = hipGetLastError();
            if (err != hipSuccess) break;
        }
    }
    if (err == hipSuccess) {
        if (level1) {
            hipLaunchKernelGGL(sum_records, dim3(level1), dim3(256), 0, L.stream, dPart, (uint32_t)nRecords, dLevel1);
            hipLaunchKernelGGL(sum_records, dim3(1), dim3(256), 0, L.stream, dLevel1, level1, dSum);
        } else {
            hipLaunchKernelGGL(sum_records, dim3(1), dim3(256), 0, L.stream, dPart, (uint32_t)nRecords, dSum);
        }
        err = hipGetLastError();
    }
    double host[8];
    if (err == hipSuccess && rho) err = hipMemcpyAsync(rho, dRho, n * sizeof(double), hipMemcpyDeviceToHost, L.stream);
    if (err == hipSuccess && u) err = hipMemcpyAsync(u, dU, 3 * n * sizeof(double), hipMemcpyDeviceToHost, L.stream);
    if (err == hipSuccess && fields)
        err = hipMemcpyAsync(fields, dFields, 7 * n * sizeof(double), hipMemcpyDeviceToHost, L.stream);
    if (err == hipSuccess) err = hipMemcpyAsync(host, dSum, sizeof(host), hipMemcpyDeviceToHost, L.stream);
    if (err == hipSuccess) err = hipStreamSynchronize(L.stream);
    hipFree(d);
    if (err != hipSuccess) return fail(PLBAMD_EHIP, std::string("moments: ") + hipGetErrorString(err));
    if (sums) {
        sums[0] = host[0];  // sum of rhoBar
        sums[1] = host[1];  // sum of |u|^2
        sums[2] = host[3];  // sum of u
        sums[3] = host[4];
        sums[4] = host[5];
    }
    return PLBAMD_OK;
}


// ---------------------------------------------------------------------------
// Adjoint gradient (adjoint.h): forward T steps with a checkpoint every K
// steps, then segments backwards: recompute the segment from its checkpoint
// with the forward kernels, recording per step the pre-collision moments
// (rhoBar, j) of every cell and the populations of the plane cells, then K
// reverse steps.

int recordStep(Lattice& L, bool odd, const Recording& r) {
    const bool bgk = L.desc.dynamics == PLBAMD_BGK;
    switch (L.desc.storage) {
    case PLBAMD_F64: return bgk ? launchStepTyped<double, 0>(L, odd, false, &r) : launchStepTyped<double, 1>(L, odd, false, &r);
    case PLBAMD_F16: return bgk ? launchStepTyped<f16s, 0>(L, odd, false, &r) : launchStepTyped<f16s, 1>(L, odd, false, &r);
    case PLBAMD_BF16: return bgk ? launchStepTyped<bf16s, 0>(L, odd, false, &r) : launchStepTyped<bf16s, 1>(L, odd, false, &r);
    default: return fail(PLBAMD_EINVAL, "adjoint: FP64 or 16-bit storage");
    }
}

int adjointGradient(Lattice& L, int T, int K, const plbamd_objective& obj, double* Jout, double* gradAlpha,
                    double* gradScale, double* timing) {
    const int st = L.desc.storage;
    if (st != PLBAMD_F64 && st != PLBAMD_F16 && st != PLBAMD_BF16)
        return fail(PLBAMD_EINVAL, "adjoint: FP64 or 16-bit storage");
    const bool forced = L.desc.force[0] != 0 || L.desc.force[1] != 0 || L.desc.force[2] != 0;
    if (!L.hostTags.empty()) return fail(PLBAMD_EINVAL, "adjoint: no per-cell tags (walls are the domain edges)");
    if (T <= 0 || K <= 0 || (K & 1)) return fail(PLBAMD_EINVAL, "adjoint: steps > 0 and an even checkpoint interval");
    if (obj.first_step < 0 || obj.first_step >= T) return fail(PLBAMD_EINVAL, "adjoint: first_step in [0, nsteps)");
    if (L.time & 1) return fail(PLBAMD_ESTATE, "adjoint: start after an even number of steps");
    if (obj.inlet_x < 0 || obj.inlet_x >= L.desc.nx || obj.outlet_x < 0 || obj.outlet_x >= L.desc.nx)
        return fail(PLBAMD_EINVAL, "adjoint: measuring planes outside the lattice");
    if (!L.planes.empty()) {
        if (L.multi) return fail(PLBAMD_EINVAL, "adjoint: boundary planes on single-block lattices only");
        if (L.desc.dynamics != PLBAMD_BGK) return fail(PLBAMD_EINVAL, "adjoint: boundary planes with BGK only");
        if (!L.hostAlpha.empty())
            for (const Plane& p : L.planes) {
                const plbamd_box& b = p.box;
                for (int x = b.x0; x <= b.x1; ++x)
                    for (int y = b.y0; y <= b.y1; ++y)
                        for (int z = b.z0; z <= b.z1; ++z)
                            if (L.hostAlpha[((uint64_t)x * L.desc.ny + y) * L.desc.nz + z] != 0.0)
                                return fail(PLBAMD_EINVAL, "adjoint: porosity must vanish on boundary planes");
            }
    }
    PLBAMD_TRY(rebuild(L));
    if (st != PLBAMD_F64 && !L.pairs) return fail(PLBAMD_EINVAL, "adjoint: 16-bit storage needs every block even along z");
    const BlockGeom& last = L.geom.back();
    const uint64_t arenaElems = last.offset + last.slot * 19;
    const size_t stateBytes = arenaElems * L.elemSize;
    const int nCk = (T + K - 1) / K;
    const uint64_t cells = L.allocCells;
    const uint64_t momElems = 4 * cells;                       // per recorded step
    const uint64_t pfElems = 19 * (uint64_t)std::max<uint32_t>(1, L.nSpecial);
    // reverse kernels: one launch over every block, 64 (z) x 4 (y) cells per workgroup
    const dim3 ablock(64, 4);
    std::vector<uint32_t> adjBegin(L.geom.size() + 1, 0);
    for (size_t b = 0; b < L.geom.size(); ++b) {
        const BlockGeom& g = L.geom[b];
        adjBegin[b + 1] = adjBegin[b] + (uint32_t)(((g.nz + 63) / 64) * ((g.ny + 3) / 4) * g.nx);
    }
    const uint32_t adjWGs = adjBegin.back();
    std::vector<uint16_t> adjBlock(adjWGs);
    for (size_t b = 0; b < L.geom.size(); ++b)
        std::fill(adjBlock.begin() + adjBegin[b], adjBlock.begin() + adjBegin[b + 1], (uint16_t)b);
    const dim3 agrid(adjWGs);
    const uint32_t planeWGs = (uint32_t)((20 * (uint64_t)L.nSpecial + 255) / 256);
    // device memory: checkpoints, a working state, the recorded moments and plane
    // populations of one segment, two adjoint states, gradients, partial sums
    // the adjoint state is FP64 for an FP64 forward, FP32 for a 16-bit one
    const bool lam64 = st == PLBAMD_F64;
    const size_t lamSize = lam64 ? sizeof(double) : sizeof(float);
    // the recorded moments have the adjoint state's type (MomentType of the storage)
    char *ck = nullptr, *work = nullptr, *lamA = nullptr, *lamB = nullptr, *mom = nullptr;
    double *pf = nullptr, *gA = nullptr, *gS = nullptr, *acc = nullptr, *sum = nullptr;
    uint16_t* dAdjBlock = nullptr;
    uint32_t* dAdjBegin = nullptr;
    auto release = [&]() {
        hipFree(ck); hipFree(work); hipFree(mom); hipFree(pf); hipFree(lamA); hipFree(lamB); hipFree(gA);
        hipFree(gS); hipFree(acc); hipFree(sum); hipFree(dAdjBlock); hipFree(dAdjBegin);
    };
    const uint64_t nAcc = (uint64_t)adjWGs + planeWGs;   // J accumulators: bulk workgroups, then plane ones
    hipError_t err = hipMalloc(&ck, stateBytes * nCk);
    if (err == hipSuccess) err = hipMalloc(&work, stateBytes);
    if (err == hipSuccess) err = hipMalloc(&mom, momElems * (K + 1) * lamSize);
    if (err == hipSuccess) err = hipMalloc(&pf, pfElems * (K + 1) * sizeof(double));
    if (err == hipSuccess) err = hipMalloc(&lamA, arenaElems * lamSize);
    if (err == hipSuccess) err = hipMalloc(&lamB, arenaElems * lamSize);
    if (err == hipSuccess) err = hipMalloc(&gA, cells * sizeof(double));
    if (err == hipSuccess) err = hipMalloc(&gS, std::max<uint32_t>(1, L.nSpecial) * sizeof(double));
    if (err == hipSuccess) err = hipMalloc(&acc, nAcc * sizeof(double));
    if (err == hipSuccess) err = hipMalloc(&sum, 8 * sizeof(double));
    if (err == hipSuccess) err = hipMalloc(&dAdjBlock, adjBlock.size() * sizeof(uint16_t));
    if (err == hipSuccess) err = hipMalloc(&dAdjBegin, adjBegin.size() * sizeof(uint32_t));
    if (err == hipSuccess)
        err = hipMemcpy(dAdjBlock, adjBlock.data(), adjBlock.size() * sizeof(uint16_t), hipMemcpyHostToDevice);
    if (err == hipSuccess)
        err = hipMemcpy(dAdjBegin, adjBegin.data(), adjBegin.size() * sizeof(uint32_t), hipMemcpyHostToDevice);
    if (err != hipSuccess) {
        (void)hipGetLastError();
        release();
        return fail(PLBAMD_ENOMEM, "adjoint: cannot allocate the checkpoint and adjoint buffers");
    }
    if (err == hipSuccess) err = hipMemsetAsync(gA, 0, cells * sizeof(double), L.stream);
    if (err == hipSuccess) err = hipMemsetAsync(gS, 0, std::max<uint32_t>(1, L.nSpecial) * sizeof(double), L.stream);
    if (err == hipSuccess) err = hipMemsetAsync(acc, 0, nAcc * sizeof(double), L.stream);
    // cells the odd steps skip (outermost envelope rows) keep zero moments
    if (err == hipSuccess) err = hipMemsetAsync(mom, 0, momElems * (K + 1) * lamSize, L.stream);
    const bool saveStats = L.statsOn;
    L.statsOn = false;
    float msF = 0, msB = 0;
    // forward, with checkpoints (even time: the AA array is in natural layout)
    if (err == hipSuccess) err = hipEventRecord(L.ev0, L.stream);
    for (int c = 0; c < nCk && err == hipSuccess; ++c) {
        err = hipMemcpyAsync(ck + (uint64_t)c * stateBytes, L.arena, stateBytes, hipMemcpyDeviceToDevice, L.stream);
        if (err != hipSuccess) break;
        if (step(L, std::min(K, T - c * K)) != PLBAMD_OK) err = hipErrorUnknown;
    }
    if (err == hipSuccess) err = hipEventRecord(L.ev1, L.stream);
    if (err == hipSuccess) err = hipEventSynchronize(L.ev1);
    if (err == hipSuccess) err = hipEventElapsedTime(&msF, L.ev0, L.ev1);
    // backward
    AdjointArgs a{};
    a.momStride = cells;
    a.alpha = L.dAlpha;
    a.gAlpha = gA;
    a.tags = L.dTags;
    a.omega = L.desc.omega;
    for (int c = 0; c < 3; ++c) a.force[c] = L.desc.force[c];
    a.invT = 1.0 / (T - obj.first_step);
    a.wdp = obj.weight_pressure_drop;
    a.wdrag = obj.weight_drag;
    a.wflux = obj.weight_flux / ((double)L.desc.nx * L.desc.ny * L.desc.nz);
    a.inW = (1.0 / 3.0) / ((double)L.desc.ny * L.desc.nz);
    a.outW = a.inW;
    a.xin = obj.inlet_x;
    a.xout = obj.outlet_x;
    a.blocks = L.dGeom;
    a.wgBlock = dAdjBlock;
    a.wgBegin = dAdjBegin;
    a.acc = acc;
    const bool rbgk = L.desc.dynamics == PLBAMD_RBGK;
#define PLBAMD_ADJ(LT) \
    do { \
        if (rbgk) { \
            if (forced) hipLaunchKernelGGL((adjoint_bulk<LT, 1, true>), gr, ablock, 0, L.stream, args); \
            else hipLaunchKernelGGL((adjoint_bulk<LT, 1, false>), gr, ablock, 0, L.stream, args); \
        } else { \
            if (forced) hipLaunchKernelGGL((adjoint_bulk<LT, 0, true>), gr, ablock, 0, L.stream, args); \
            else hipLaunchKernelGGL((adjoint_bulk<LT, 0, false>), gr, ablock, 0, L.stream, args); \
        } \
    } while (0)
    auto reverseKernel = [&](const AdjointArgs& args, const dim3& gr) {
        if (lam64) PLBAMD_ADJ(double);
        else PLBAMD_ADJ(float);
    };
#undef PLBAMD_ADJ
    auto exchangeLam = [&](char* lam) {
        if (!L.multi || !L.nSegs) return;
        // the first envelope layer only: the reverse step gathers lambda one
        // cell beyond the bulk, and every envelope cell it computes itself is
        // replaced by this exchange
        if (!L.nTasks1) return;
        if (lam64)
            hipLaunchKernelGGL((exchange_envelopes<double>), dim3(L.nTasks1), dim3(256), 0, L.stream, (void*)lam,
                               L.dGeom, L.dSegs1, L.dTasks1);
        else
            hipLaunchKernelGGL((exchange_envelopes<float>), dim3(L.nTasks1), dim3(256), 0, L.stream, (void*)lam,
                               L.dGeom, L.dSegs1, L.dTasks1);
    };
    // the plane cells' reverse step runs on a second stream next to the bulk one
    hipStream_t side = nullptr;
    hipEvent_t evIn = nullptr, evOut = nullptr;
    if (err == hipSuccess && L.nSpecial) {
        err = hipStreamCreateWithFlags(&side, hipStreamNonBlocking);
        if (err == hipSuccess) err = hipEventCreateWithFlags(&evIn, hipEventDisableTiming);
        if (err == hipSuccess) err = hipEventCreateWithFlags(&evOut, hipEventDisableTiming);
    }
    char* lamNext = lamA;
    char* lamCur = lamB;
    void* const arena = L.arena;
    const int64_t time0 = L.time;
    L.arena = work;   // the recompute runs the step kernels on the working state
    if (err == hipSuccess) err = hipEventRecord(L.ev0, L.stream);
    for (int c = nCk - 1; c >= 0 && err == hipSuccess; --c) {
        const int t0 = c * K, len = std::min(K, T - t0);
        err = hipMemcpyAsync(work, ck + (uint64_t)c * stateBytes, stateBytes, hipMemcpyDeviceToDevice, L.stream);
        // steps t0 .. t0 + len - 1 (and, in the last segment, one more for f^T)
        const int nRec = (c == nCk - 1) ? len + 1 : len;
        for (int m = 0; m < nRec && err == hipSuccess; ++m) {
            Recording r;
            r.moments = mom + (uint64_t)m * momElems * lamSize;
            r.momStride = cells;
            r.specialF = L.nSpecial ? pf + (uint64_t)m * pfElems : nullptr;
            if (recordStep(L, (m & 1) != 0, r) != PLBAMD_OK) err = hipErrorUnknown;
        }
        if (err == hipSuccess && c == nCk - 1) {
            a.mom = mom + (uint64_t)len * momElems * lamSize;
            a.explicitTerm = 1;
            a.lam = lamNext;
            if (lam64) hipLaunchKernelGGL(adjoint_final<double>, agrid, ablock, 0, L.stream, a);
            else hipLaunchKernelGGL(adjoint_final<float>, agrid, ablock, 0, L.stream, a);
            err = hipGetLastError();
            exchangeLam(lamNext);
        }
        for (int m = len - 1; m >= 0 && err == hipSuccess; --m) {
            const int t = t0 + m;
            a.mom = mom + (uint64_t)m * momElems * lamSize;
            a.planeF = pf + (uint64_t)m * pfElems;
            a.explicitTerm = t > obj.first_step ? 1 : 0;
            a.lamNext = lamNext;
            a.lam = lamCur;
            if (err == hipSuccess && L.nSpecial) {
                // plane cells (single block) on the side stream, after the previous step
                AdjointSpecialArgs sa{};
                AdjointArgs pa = a;
                pa.g = L.geom[0];
                sa.base = pa;
                sa.cells = L.dSpecial;
                sa.count = L.nSpecial;
                sa.values = L.dValues;
                sa.partials = acc + adjWGs;
                sa.gScale = gS;
                err = hipEventRecord(evIn, L.stream);
                if (err == hipSuccess) err = hipStreamWaitEvent(side, evIn, 0);
                if (err == hipSuccess) {
                    if (lam64) hipLaunchKernelGGL(adjoint_planes<double>, dim3(planeWGs), dim3(256), 0, side, sa);
                    else hipLaunchKernelGGL(adjoint_planes<float>, dim3(planeWGs), dim3(256), 0, side, sa);
                    err = hipGetLastError();
                }
                if (err == hipSuccess) err = hipEventRecord(evOut, side);
            }
            if (err == hipSuccess) {
                reverseKernel(a, agrid);
                err = hipGetLastError();
            }
            if (err == hipSuccess && L.nSpecial) err = hipStreamWaitEvent(L.stream, evOut, 0);
            exchangeLam(lamCur);
            std::swap(lamNext, lamCur);
        }
    }
    L.arena = arena;
    L.time = time0;
    if (err == hipSuccess) err = hipEventRecord(L.ev1, L.stream);
    if (side) {
        (void)hipStreamSynchronize(side);
        (void)hipStreamDestroy(side);
    }
    if (evIn) (void)hipEventDestroy(evIn);
    if (evOut) (void)hipEventDestroy(evOut);
    if (err == hipSuccess) err = hipEventSynchronize(L.ev1);
    if (err == hipSuccess) err = hipEventElapsedTime(&msB, L.ev0, L.ev1);
    L.statsOn = saveStats;
    if (err == hipSuccess) {
        hipLaunchKernelGGL(sum_values, dim3(1), dim3(256), 0, L.stream, acc, (uint64_t)adjWGs, acc + adjWGs,
                           (uint64_t)planeWGs, sum);
        err = hipGetLastError();
    }
    double hsum[8] = {0};
    if (err == hipSuccess) err = hipMemcpyAsync(hsum, sum, sizeof(hsum), hipMemcpyDeviceToHost, L.stream);
    std::vector<double> hgA(gradAlpha ? cells : 0), hgS(std::max<uint32_t>(1, L.nSpecial));
    if (err == hipSuccess && gradAlpha)
        err = hipMemcpyAsync(hgA.data(), gA, cells * sizeof(double), hipMemcpyDeviceToHost, L.stream);
    if (err == hipSuccess) err = hipMemcpyAsync(hgS.data(), gS, hgS.size() * sizeof(double), hipMemcpyDeviceToHost, L.stream);
    std::vector<SpecialCell> hs(L.nSpecial);
    if (err == hipSuccess && L.nSpecial)
        err = hipMemcpyAsync(hs.data(), L.dSpecial, L.nSpecial * sizeof(SpecialCell), hipMemcpyDeviceToHost, L.stream);
    if (err == hipSuccess) err = hipStreamSynchronize(L.stream);
    release();
    if (err != hipSuccess) return fail(PLBAMD_EHIP, std::string("adjoint: ") + hipGetErrorString(err));
    if (Jout) *Jout = hsum[0];
    if (gradAlpha) {
        // bulk cells of every block to the global [nx][ny][nz] array
        const int ny = L.desc.ny, nz = L.desc.nz;
        for (const BlockGeom& g : L.geom)
            for (int x = g.bx0; x <= g.bx1; ++x)
                for (int y = g.by0; y <= g.by1; ++y)
                    for (int z = g.bz0; z <= g.bz1; ++z) {
                        const uint64_t local = ((uint64_t)x * g.ny + y) * g.nz + z;
                        const uint64_t glob = ((uint64_t)(g.gx0 + x) * ny + (g.gy0 + y)) * nz + (g.gz0 + z);
                        gradAlpha[glob] = hgA[g.cellOffset + local];
                    }
    }
    if (gradScale) {
        // plane of each special record from its value index
        std::vector<uint32_t> base(L.planes.size() + 1, 0);
        for (size_t p = 0; p < L.planes.size(); ++p) base[p + 1] = base[p] + (uint32_t)L.planes[p].values.size();
        for (size_t p = 0; p < L.planes.size(); ++p) gradScale[p] = 0.0;
        for (uint32_t k = 0; k < L.nSpecial; ++k) {
            if ((hs[k].info & 3u) != SK_PLANE) continue;
            const size_t p = std::upper_bound(base.begin(), base.end(), hs[k].value) - base.begin() - 1;
            gradScale[p] += hgS[k];
        }
    }
    if (timing) {
        timing[0] = msF;
        timing[1] = msB;
    }
    return PLBAMD_OK;
}

int registerLattice(std::unique_ptr<Lattice> L, int* handle) {
    std::lock_guard<std::mutex> lock(g_mutex);
    *handle = g_nextHandle++;
    g_lattices[*handle] = std::move(L);
    return PLBAMD_OK;
}

}  // namespace plbamd

using namespace plbamd;

extern "C" {

int plbamd_init(int device) {
    int count = 0;
    PLBAMD_HIP(hipGetDeviceCount(&count));
    if (device < 0 || device >= count) return fail(PLBAMD_EINVAL, "plbamd_init: no such device");
    PLBAMD_HIP(hipSetDevice(device));
    return PLBAMD_OK;
}

const char* plbamd_last_error(void) { return g_lastError.c_str(); }

int plbamd_lattice_create(const plbamd_lattice_desc* desc, int* handle) {
    if (!desc || !handle) return fail(PLBAMD_EINVAL, "lattice_create: null argument");
    PLBAMD_TRY(validateDesc(desc));
    auto L = std::make_unique<Lattice>();
    L->desc = *desc;
    L->multi = false;
    L->cells = (uint64_t)desc->nx * desc->ny * desc->nz;
    plbamd_box all{0, desc->nx - 1, 0, desc->ny - 1, 0, desc->nz - 1};
    L->bulk.push_back(all);
    PLBAMD_TRY(setup(*L));
    return registerLattice(std::move(L), handle);
}

int plbamd_multiblock_create(const plbamd_lattice_desc* desc, const plbamd_box* blocks, int nblocks, int* handle) {
    if (!desc || !handle || !blocks || nblocks <= 0) return fail(PLBAMD_EINVAL, "multiblock_create: bad argument");
    PLBAMD_TRY(validateDesc(desc));
    const uint64_t cells = (uint64_t)desc->nx * desc->ny * desc->nz;
    // the blocks must tile the domain: inside, no overlap, every cell covered
    std::vector<uint8_t> cover(cells, 0);
    for (int b = 0; b < nblocks; ++b) {
        const plbamd_box& B = blocks[b];
        if (B.x0 < 0 || B.y0 < 0 || B.z0 < 0 || B.x0 > B.x1 || B.y0 > B.y1 || B.z0 > B.z1 || B.x1 >= desc->nx ||
            B.y1 >= desc->ny || B.z1 >= desc->nz)
            return fail(PLBAMD_EINVAL, "multiblock_create: block " + std::to_string(b) + " leaves the domain");
        for (int x = B.x0; x <= B.x1; ++x)
            for (int y = B.y0; y <= B.y1; ++y)
                for (int z = B.z0; z <= B.z1; ++z) {
                    uint8_t& c = cover[((uint64_t)x * desc->ny + y) * desc->nz + z];
                    if (c) return fail(PLBAMD_EINVAL, "multiblock_create: blocks overlap");
                    c = 1;
                }
    }
    for (uint64_t k = 0; k < cells; ++k)
        if (!cover[k]) return fail(PLBAMD_EINVAL, "multiblock_create: the blocks do not cover the domain");
    auto L = std::make_unique<Lattice>();
    L->desc = *desc;
    L->multi = true;
    L->cells = cells;
    L->bulk.assign(blocks, blocks + nblocks);
    PLBAMD_TRY(setup(*L));
    return registerLattice(std::move(L), handle);
}

int plbamd_lattice_destroy(int handle) {
    std::lock_guard<std::mutex> lock(g_mutex);
    auto it = g_lattices.find(handle);
    if (it == g_lattices.end()) return fail(PLBAMD_EINVAL, "lattice_destroy: unknown handle");
    (void)hipStreamSynchronize(it->second->stream);
    g_lattices.erase(it);
    return PLBAMD_OK;
}

int plbamd_set_tags(int handle, const uint8_t* tags) {
    Lattice* L = find(handle);
    if (!L) return fail(PLBAMD_EINVAL, "set_tags: unknown handle");
    if (!tags) { L->hostTags.clear(); L->dirty = true; return PLBAMD_OK; }
    for (uint64_t k = 0; k < L->cells; ++k)
        if (tags[k] > PLBAMD_TAG_WALL_LINK) return fail(PLBAMD_EINVAL, "set_tags: unknown tag value");
    L->hostTags.assign(tags, tags + L->cells);
    L->dirty = true;
    return PLBAMD_OK;
}

int plbamd_add_plane(int handle, int family, int kind, int dir, int ori, const plbamd_box* box, int* plane) {
    Lattice* L = find(handle);
    if (!L) return fail(PLBAMD_EINVAL, "add_plane: unknown handle");
    if (!box || !boxValid(*L, *box)) return fail(PLBAMD_EINVAL, "add_plane: box outside the lattice");
    const bool wrapped = (family & PLBAMD_WRAPPED) != 0;
    family &= ~PLBAMD_WRAPPED;
    if (family != PLBAMD_ZOUHE && family != PLBAMD_REGULARIZED) return fail(PLBAMD_EINVAL, "add_plane: unknown family");
    if (kind != PLBAMD_VELOCITY && kind != PLBAMD_PRESSURE) return fail(PLBAMD_EINVAL, "add_plane: unknown kind");
    if (dir < 0 || dir > 2 || (ori != -1 && ori != 1)) return fail(PLBAMD_EINVAL, "add_plane: bad direction/orientation");
    const int lo[3] = {box->x0, box->y0, box->z0}, hi[3] = {box->x1, box->y1, box->z1};
    if (lo[dir] != hi[dir]) return fail(PLBAMD_EINVAL, "add_plane: box must be one cell thick along dir");
    if (L->desc.force[0] != 0 || L->desc.force[1] != 0 || L->desc.force[2] != 0)
        return fail(PLBAMD_EINVAL, "add_plane: planes are not combined with a body force");
    Plane p;
    p.family = family; p.kind = kind; p.dir = dir; p.ori = ori; p.box = *box; p.wrapped = wrapped;
    p.values.assign(boxCells(*box) * (kind == PLBAMD_VELOCITY ? 3 : 1), 0.0);
    if (kind == PLBAMD_PRESSURE) std::fill(p.values.begin(), p.values.end(), 1.0);
    L->planes.push_back(p);
    if (plane) *plane = (int)L->planes.size() - 1;
    L->dirty = true;
    return PLBAMD_OK;
}

int plbamd_set_plane_values(int handle, int plane, const double* values, size_t count) {
    Lattice* L = find(handle);
    if (!L) return fail(PLBAMD_EINVAL, "set_plane_values: unknown handle");
    if (plane < 0 || plane >= (int)L->planes.size()) return fail(PLBAMD_EINVAL, "set_plane_values: unknown plane");
    Plane& p = L->planes[plane];
    if (!values || count != p.values.size()) return fail(PLBAMD_EINVAL, "set_plane_values: wrong value count");
    std::copy(values, values + count, p.values.begin());
    if (!L->dirty && L->dValues && plane < (int)L->planeValueBase.size()) {
        // built already: refresh the plane's values on the device (after the queued steps)
        PLBAMD_HIP(hipStreamSynchronize(L->stream));
        PLBAMD_HIP(hipMemcpy(L->dValues + L->planeValueBase[plane], values, count * sizeof(double),
                             hipMemcpyHostToDevice));
        return PLBAMD_OK;
    }
    L->dirty = true;
    return PLBAMD_OK;
}

int plbamd_set_porosity(int handle, const double* alpha) {
    Lattice* L = find(handle);
    if (!L) return fail(PLBAMD_EINVAL, "set_porosity: unknown handle");
    if (alpha) L->hostAlpha.assign(alpha, alpha + L->cells);
    else L->hostAlpha.clear();
    L->dirty = true;
    return PLBAMD_OK;
}

static int dispatchTransfer(int handle, const plbamd_box* box, int elem, void* data, size_t bytes, bool up) {
    Lattice* L = find(handle);
    if (!L) return fail(PLBAMD_EINVAL, "transfer: unknown handle");
    if (!box || !boxValid(*L, *box)) return fail(PLBAMD_EINVAL, "transfer: box outside the lattice");
    if (!data) return fail(PLBAMD_EINVAL, "transfer: null data");
    if (elem < PLBAMD_F64 || elem > PLBAMD_BF16) return fail(PLBAMD_EINVAL, "transfer: unknown element type");
    switch (L->desc.storage) {
    case PLBAMD_F64: return transfer<double>(*L, *box, elem, data, bytes, up);
    case PLBAMD_F32: return transfer<float>(*L, *box, elem, data, bytes, up);
    case PLBAMD_F16: return transfer<f16s>(*L, *box, elem, data, bytes, up);
    default: return transfer<bf16s>(*L, *box, elem, data, bytes, up);
    }
}

static int dispatchCheckpoint(int handle, const char* path, int elem, int force, int enforceUint, bool save) {
    Lattice* L = find(handle);
    const char* what = save ? "save_binary_block" : "load_binary_block";
    if (!L) return fail(PLBAMD_EINVAL, std::string(what) + ": unknown handle");
    if (!path) return fail(PLBAMD_EINVAL, std::string(what) + ": null file name");
    if (elem != PLBAMD_F64 && elem != PLBAMD_F32)
        return fail(PLBAMD_EINVAL, std::string(what) + ": the file element type is FP64 or FP32");
    switch (L->desc.storage) {
    case PLBAMD_F64: return checkpoint<double>(*L, path, elem, force != 0, enforceUint != 0, save);
    case PLBAMD_F32: return checkpoint<float>(*L, path, elem, force != 0, enforceUint != 0, save);
    case PLBAMD_F16: return checkpoint<f16s>(*L, path, elem, force != 0, enforceUint != 0, save);
    default: return checkpoint<bf16s>(*L, path, elem, force != 0, enforceUint != 0, save);
    }
}

int plbamd_save_binary_block(int handle, const char* file_name, int element, int force_externals, int enforce_uint) {
    return dispatchCheckpoint(handle, file_name, element, force_externals, enforce_uint, true);
}

int plbamd_load_binary_block(int handle, const char* file_name, int element, int force_externals, int enforce_uint) {
    return dispatchCheckpoint(handle, file_name, element, force_externals, enforce_uint, false);
}

int plbamd_upload(int handle, const plbamd_box* box, int elem, const void* data, size_t bytes) {
    return dispatchTransfer(handle, box, elem, const_cast<void*>(data), bytes, true);
}

int plbamd_download(int handle, const plbamd_box* box, int elem, void* data, size_t bytes) {
    return dispatchTransfer(handle, box, elem, data, bytes, false);
}

int plbamd_initialize_at_equilibrium(int handle, const plbamd_box* box, const double* rho, const double* u) {
    Lattice* L = find(handle);
    if (!L) return fail(PLBAMD_EINVAL, "initialize_at_equilibrium: unknown handle");
    if (!box || !boxValid(*L, *box) || !rho || !u) return fail(PLBAMD_EINVAL, "initialize_at_equilibrium: bad argument");
    switch (L->desc.storage) {
    case PLBAMD_F64: return equilibrium<double>(*L, *box, rho, u);
    case PLBAMD_F32: return equilibrium<float>(*L, *box, rho, u);
    case PLBAMD_F16: return equilibrium<f16s>(*L, *box, rho, u);
    default: return equilibrium<bf16s>(*L, *box, rho, u);
    }
}

int plbamd_collide_and_stream(int handle, int nsteps) {
    Lattice* L = find(handle);
    if (!L) return fail(PLBAMD_EINVAL, "collide_and_stream: unknown handle");
    if (nsteps < 0) return fail(PLBAMD_EINVAL, "collide_and_stream: negative step count");
    PLBAMD_TRY(step(*L, nsteps));
    PLBAMD_HIP(hipStreamSynchronize(L->stream));
    return PLBAMD_OK;
}

int plbamd_collide_and_stream_timed(int handle, int nsteps, double* ms) {
    Lattice* L = find(handle);
    if (!L) return fail(PLBAMD_EINVAL, "collide_and_stream_timed: unknown handle");
    if (nsteps < 0 || !ms) return fail(PLBAMD_EINVAL, "collide_and_stream_timed: bad argument");
    PLBAMD_TRY(rebuild(*L));
    PLBAMD_HIP(hipEventRecord(L->ev0, L->stream));
    PLBAMD_TRY(step(*L, nsteps));
    PLBAMD_HIP(hipEventRecord(L->ev1, L->stream));
    PLBAMD_HIP(hipEventSynchronize(L->ev1));
    float elapsed = 0.f;
    PLBAMD_HIP(hipEventElapsedTime(&elapsed, L->ev0, L->ev1));
    *ms = elapsed;
    return PLBAMD_OK;
}

int plbamd_toggle_statistics(int handle, int on) {
    Lattice* L = find(handle);
    if (!L) return fail(PLBAMD_EINVAL, "toggle_statistics: unknown handle");
    L->statsOn = on != 0;
    return PLBAMD_OK;
}

int plbamd_get_statistics(int handle, double* out) {
    Lattice* L = find(handle);
    if (!L) return fail(PLBAMD_EINVAL, "get_statistics: unknown handle");
    if (!out) return fail(PLBAMD_EINVAL, "get_statistics: null output");
    if (!L->statsValid) return fail(PLBAMD_ESTATE, "get_statistics: no step has gathered statistics yet");
    double h[16];
    PLBAMD_HIP(hipMemcpyAsync(h, L->dStats, sizeof(h), hipMemcpyDeviceToHost, L->stream));
    PLBAMD_HIP(hipStreamSynchronize(L->stream));
    const double count = h[3];
    out[PLBAMD_STAT_AVG_RHOBAR] = count > 0 ? h[0] / count : 0.0;
    out[PLBAMD_STAT_AVG_USQR] = count > 0 ? h[1] / count : 0.0;
    out[PLBAMD_STAT_MAX_USQR] = h[2];
    out[PLBAMD_STAT_FORCE_X] = h[4];
    out[PLBAMD_STAT_FORCE_Y] = h[5];
    out[PLBAMD_STAT_FORCE_Z] = h[6];
    out[PLBAMD_STAT_COUNT] = count;
    out[PLBAMD_STAT_POROUS_X] = h[8];
    out[PLBAMD_STAT_POROUS_Y] = h[9];
    out[PLBAMD_STAT_POROUS_Z] = h[11];
    return PLBAMD_OK;
}

static int dispatchMoments(int handle, const plbamd_box* box, double* rho, double* u, double* sums,
                           double* fields = nullptr) {
    Lattice* L = find(handle);
    if (!L) return fail(PLBAMD_EINVAL, "moments: unknown handle");
    if (!box || !boxValid(*L, *box)) return fail(PLBAMD_EINVAL, "moments: box outside the lattice");
    switch (L->desc.storage) {
    case PLBAMD_F64: return moments<double>(*L, *box, rho, u, sums, fields);
    case PLBAMD_F32: return moments<float>(*L, *box, rho, u, sums, fields);
    case PLBAMD_F16: return moments<f16s>(*L, *box, rho, u, sums, fields);
    default: return moments<bf16s>(*L, *box, rho, u, sums, fields);
    }
}

int plbamd_set_solid_density(int handle, int tag, double rho) {
    Lattice* L = find(handle);
    if (!L) return fail(PLBAMD_EINVAL, "set_solid_density: unknown handle");
    if (tag < PLBAMD_TAG_BOUNCE_BACK || tag > PLBAMD_TAG_MOMENTUM_EXCHANGE)
        return fail(PLBAMD_EINVAL, "set_solid_density: tag must be BounceBack, NoDynamics or momentum exchange");
    L->solidRho[tag - PLBAMD_TAG_BOUNCE_BACK] = rho;
    return PLBAMD_OK;
}

int plbamd_compute_fields(int handle, const plbamd_box* box, double* out) {
    if (!out) return fail(PLBAMD_EINVAL, "compute_fields: null output");
    return dispatchMoments(handle, box, nullptr, nullptr, nullptr, out);
}

int plbamd_compute_macroscopic(int handle, const plbamd_box* box, double* rho, double* u) {
    if (!rho && !u) return fail(PLBAMD_EINVAL, "compute_macroscopic: no output");
    return dispatchMoments(handle, box, rho, u, nullptr);
}

int plbamd_reduce_moments(int handle, const plbamd_box* box, double* out) {
    if (!out) return fail(PLBAMD_EINVAL, "reduce_moments: null output");
    return dispatchMoments(handle, box, nullptr, nullptr, out);
}

int plbamd_adjoint_gradient(int handle, int nsteps, int checkpoint_every, const plbamd_objective* objective,
                            double* J, double* grad_alpha, double* grad_plane_scale, double* timing_ms) {
    Lattice* L = find(handle);
    if (!L || !objective) return fail(PLBAMD_EINVAL, "adjoint_gradient: bad argument");
    return adjointGradient(*L, nsteps, checkpoint_every, *objective, J, grad_alpha, grad_plane_scale, timing_ms);
}

int plbamd_get_time(int handle, int64_t* t) {
    Lattice* L = find(handle);
    if (!L || !t) return fail(PLBAMD_EINVAL, "get_time: bad argument");
    *t = L->time;
    return PLBAMD_OK;
}

}  // extern "C"
