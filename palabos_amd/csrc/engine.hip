// palabos_amd engine: D3Q19 collide-and-stream on gfx950 (MI355X) — lattice
// set-up, the step loop and the C ABI (include/palabos_amd.h). The kernels and
// the data layout are described in kernels.h.
#include "engine.h"

#include <unordered_map>

namespace plbamd {

// Cells up to which the odd steps write their z-moving outputs with ordinary
// stores (bulk_odd_staged and bulk_pairs CACHEDZ). Measured on periodic n^3
// lattices (step time, nontemporal -> ordinary): FP64 128^3 -9 %, 224^3 -6 %,
// 384^3 -1.4 %, 512^3 +1.5 %; FP32 192^3 -11 %, 384^3 -0.6 %, 512^3 +2 %;
// FP16 192^3 -7 %, 256^3 -7 %, 384^3 -1.4 %, 512^3 +1.8 %: the turn lies
// between 384^3 and 512^3 cells whatever the storage. The sphere channel
// 256x128x128 -10 % (FP32) and -11 % (FP64); the 384x256x256 porous medium
// (FP16) -3 %.
constexpr uint64_t kCachedZCells = uint64_t(1) << 26;

thread_local std::string g_lastError;

int fail(int code, const std::string& msg) {
    g_lastError = msg;
    return code;
}

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
    PLBAMD_HIP(hipGetDevice(&L.device));
    const int n[3] = {d.nx, d.ny, d.nz};
    const int per[3] = {d.periodic[0] != 0, d.periodic[1] != 0, d.periodic[2] != 0};
    // Launch shape: TX threads along z (a multiple of the 64-lane wavefront),
    // CPT cells each per row segment; TY rows per workgroup when a row is short.
    // Up to 512 threads, so that a 512-cell row is one segment of the staged odd
    // step (a periodic row then wraps through LDS with no halo loads: FP64 at
    // 512^3 18.3k against 17.9k MLUPS with 256-thread segments), workgroups
    // remapped so that neighbouring row segments share an XCD.
    int txMax = 512, cpt = 1;
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
    // Staged odd steps take one row segment per workgroup, also for short rows:
    // two or four 128-cell rows per workgroup measured slower on the 256x128x128
    // channel (FP32 odd step 138.7, 141.2, 147.4 us; FP64 262.4, 264.4, 271.4 us).
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
    // ordinary stores for the odd steps' z-moving outputs on lattices of at most
    // kCachedZCells cells (PLBAMD_CACHED_Z=0/1 overrides)
    L.cachedZ = cellOffset <= kCachedZCells;
    if (const char* e = getenv("PLBAMD_CACHED_Z")) L.cachedZ = atoi(e) != 0;
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

// Copy processor of a boundary object (0 none; ProcOp otherwise): the
// boundary types that install one (boundaryInstantiator3D.h:1042-1063,
// :1097-1102, :1139-1157, :1216-1232).
static int copyOp(const Plane& P) {
    if (P.shape == SHAPE_PLANE && P.kind == PLBAMD_PRESSURE) return P.bcType == PLBAMD_BC_NEUMANN ? PROC_COPY_DENSITY : 0;
    if (P.bcType == PLBAMD_BC_NEUMANN || P.bcType == PLBAMD_BC_OUTFLOW) return PROC_COPY_VELOCITY;
    if (P.bcType == PLBAMD_BC_FREESLIP) return PROC_COPY_TANGENTIAL;
    if (P.bcType == PLBAMD_BC_NORMAL_OUTFLOW) return PROC_COPY_NORMAL;
    return 0;
}

// The data-processor records (fd_boundary.h) of a single-block lattice, in four
// phases: copies of planes and edges, edge closures, copies of corners, corner
// closures. Palabos runs each object's copy, then its closure, object by object
// in the order they were added; the phases give the same results when no cell is
// read in the phase that writes it and every pair of touching objects was added
// in phase order (faces, edges, corners), which this checks.
int buildProcessors(Lattice& L, const std::vector<int>& planeOf, const std::vector<uint32_t>& boxIndex,
                    const std::vector<SpecialCell>& special, std::vector<ProcCell>& out) {
    out.clear();
    L.anyCopy = false;
    bool any = false;
    for (const Plane& P : L.planes) any = any || P.shape != SHAPE_PLANE || copyOp(P) != 0;
    for (int k = 0; k < 5; ++k) L.procPhase[k] = 0;
    if (!any) return PLBAMD_OK;
    if (L.multi) return fail(PLBAMD_EINVAL, "edges, corners and non-Dirichlet boundary types need a single-block lattice");
    const int nx = L.desc.nx, ny = L.desc.ny, nz = L.desc.nz;
    const int per[3] = {L.desc.periodic[0] != 0, L.desc.periodic[1] != 0, L.desc.periodic[2] != 0};
    auto at = [&](int x, int y, int z) { return ((uint64_t)x * ny + y) * nz + z; };
    std::unordered_map<uint64_t, uint32_t> specialAt;
    for (uint32_t k = 0; k < special.size(); ++k) specialAt[at(special[k].x, special[k].y, special[k].z)] = k;
    const bool haveTags = !L.hostTags.empty();
    std::vector<ProcCell> phase[4];
    // write phases of every object: its stored values (copy), its populations
    // (closure of an edge or corner; a wrapped plane's completion, which the
    // device performs when the plane is read)
    auto valuesPhase = [&](int p) { return copyOp(L.planes[p]) ? (L.planes[p].shape == SHAPE_CORNER ? 2 : 0) : -1; };
    auto popsPhase = [&](int p) {
        const Plane& P = L.planes[p];
        if (P.shape == SHAPE_EDGE) return 1;
        if (P.shape == SHAPE_CORNER) return 3;
        return P.wrapped ? 0 : -1;
    };
    std::string err;
    // what a processor reads of a neighbour: its velocity, or its density and velocity
    enum { READ_U = 1, READ_RHO = 2 };
    // a neighbour of object `owner` in phase `ph`; returns its reference, or fails
    auto neighbour = [&](int owner, int ph, int x, int y, int z, int dx, int dy, int dz, int reads,
                         uint32_t& ref) -> bool {
        int v[3] = {x + dx, y + dy, z + dz};
        const int n[3] = {nx, ny, nz};
        for (int a = 0; a < 3; ++a) {
            if (v[a] >= 0 && v[a] < n[a]) continue;
            if (!per[a]) { err = "a boundary processor's stencil leaves the lattice"; return false; }
            v[a] = (v[a] % n[a] + n[a]) % n[a];
        }
        const uint64_t c = at(v[0], v[1], v[2]);
        auto it = specialAt.find(c);
        if (it == specialAt.end()) {
            if (haveTags && L.hostTags[c] != PLBAMD_TAG_BULK) {
                err = "a boundary processor reads a solid cell";
                return false;
            }
            ref = (uint32_t)c;
            return true;
        }
        const uint32_t kind = special[it->second].info & 3u;
        if (kind != SK_PLANE && kind != SK_STORE_VELOCITY) {
            err = "a boundary processor reads a momentum-exchange cell";
            return false;
        }
        const int w = planeOf[c];
        // a velocity read of an edge, corner or velocity plane takes its stored
        // value only; every other read takes the populations (and a plane's value)
        const bool storedU = kind == SK_STORE_VELOCITY || L.planes[w].kind == PLBAMD_VELOCITY;
        const bool readsPops = (reads & READ_RHO) || !storedU;
        // the sequential order: w's writes are seen iff w was added before the reader
        // (or is the reader itself, whose copy precedes its closure)
        for (int wp : {valuesPhase(w), readsPops ? popsPhase(w) : -1}) {
            if (wp < 0) continue;
            if (wp == ph) {
                err = "a boundary processor reads a cell written in the same pass";
                return false;
            }
            const bool seen = wp < ph;
            const bool before = w < owner || (w == owner && wp == valuesPhase(w));
            if (seen != before) {
                err = "boundary objects that touch must be added faces first, then edges, then corners";
                return false;
            }
        }
        ref = it->second | kNbSpecial;
        return true;
    };
    for (int p = 0; p < (int)L.planes.size(); ++p) {
        const Plane& P = L.planes[p];
        const plbamd_box& b = P.box;
        const int op = copyOp(P);
        if (op) L.anyCopy = true;
        int n[3];
        for (int a = 0; a < 3; ++a) n[a] = P.shape == SHAPE_PLANE ? (a == P.dir ? P.ori : 0) : P.normal[a];
        for (int x = b.x0; x <= b.x1; ++x)
            for (int y = b.y0; y <= b.y1; ++y)
                for (int z = b.z0; z <= b.z1; ++z) {
                    const uint64_t c = at(x, y, z);
                    if (planeOf[c] != p) continue;
                    ProcCell r{};
                    r.special = specialAt.at(c);
                    for (int a = 0; a < 3; ++a) r.n[a] = (int8_t)n[a];
                    if (op) {
                        const int ph = P.shape == SHAPE_CORNER ? 2 : 0;
                        r.op = (uint8_t)op;
                        if (!neighbour(p, ph, x, y, z, -n[0], -n[1], -n[2], op == PROC_COPY_DENSITY ? READ_RHO : READ_U,
                                       r.nb[0]))
                            return fail(PLBAMD_EINVAL, err);
                        phase[ph].push_back(r);
                    }
                    if (P.shape == SHAPE_EDGE) {
                        const int pl = P.dir, d1 = (pl + 1) % 3, d2 = (pl + 2) % 3;
                        int s1[3] = {0, 0, 0}, s2[3] = {0, 0, 0}, sp[3] = {0, 0, 0};
                        s1[d1] = -n[d1]; s2[d2] = -n[d2]; sp[pl] = 1;
                        r.op = PROC_EDGE;
                        if (!neighbour(p, 1, x, y, z, s1[0], s1[1], s1[2], READ_RHO | READ_U, r.nb[0]) ||
                            !neighbour(p, 1, x, y, z, s2[0], s2[1], s2[2], READ_RHO | READ_U, r.nb[1]) ||
                            !neighbour(p, 1, x, y, z, sp[0], sp[1], sp[2], READ_U, r.nb[2]) ||
                            !neighbour(p, 1, x, y, z, -sp[0], -sp[1], -sp[2], READ_U, r.nb[3]))
                            return fail(PLBAMD_EINVAL, err);
                        phase[1].push_back(r);
                    } else if (P.shape == SHAPE_CORNER) {
                        r.op = PROC_CORNER;
                        if (!neighbour(p, 3, x, y, z, -n[0], 0, 0, READ_RHO | READ_U, r.nb[0]) ||
                            !neighbour(p, 3, x, y, z, 0, -n[1], 0, READ_RHO | READ_U, r.nb[1]) ||
                            !neighbour(p, 3, x, y, z, 0, 0, -n[2], READ_RHO | READ_U, r.nb[2]))
                            return fail(PLBAMD_EINVAL, err);
                        phase[3].push_back(r);
                    }
                }
    }
    (void)boxIndex;
    uint32_t k = 0;
    for (int ph = 0; ph < 4; ++ph) {
        L.procPhase[ph] = k;
        out.insert(out.end(), phase[ph].begin(), phase[ph].end());
        k += (uint32_t)phase[ph].size();
    }
    L.procPhase[4] = k;
    return PLBAMD_OK;
}

// Builds the device tags, the special-cell list and the porosity array of every
// block from the domain-wide description (tags, planes, porosity).
int rebuild(Lattice& L) {
    if (!L.dirty) return PLBAMD_OK;
    if (L.anyCopy && L.dValues) {
        // the copy processors rewrote the stored values on the device: keep them
        for (size_t p = 0; p < L.planeValueBase.size() && p < L.planes.size(); ++p)
            PLBAMD_HIP(hipMemcpy(L.planes[p].values.data(), L.dValues + L.planeValueBase[p],
                                 L.planes[p].values.size() * sizeof(double), hipMemcpyDeviceToHost));
    }
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
                        const int prev = planeOf[at(x, y, z)];
                        if (prev >= 0 && (L.planes[prev].shape != SHAPE_PLANE || L.planes[p].shape != SHAPE_PLANE))
                            return fail(PLBAMD_EINVAL, "boundary objects overlap at an edge or corner cell");
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
                    if (p >= 0 && L.planes[p].shape != SHAPE_PLANE) {
                        // edge or corner: StoreVelocityDynamics over the background dynamics
                        s.info = (uint16_t)SK_STORE_VELOCITY | bulkBit;
                        s.value = planeValueBase[p] + boxIndex[c] * 3u;
                        dev[slot] = BT_SKIP;
                        special.push_back(s);
                    } else if (p >= 0) {
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
    std::vector<ProcCell> procs;
    PLBAMD_TRY(buildProcessors(L, planeOf, boxIndex, special, procs));
    hipFree(L.dProc); L.dProc = nullptr;
    if (!procs.empty()) {
        PLBAMD_HIP(hipMalloc(&L.dProc, procs.size() * sizeof(ProcCell)));
        PLBAMD_HIP(hipMemcpy(L.dProc, procs.data(), procs.size() * sizeof(ProcCell), hipMemcpyHostToDevice));
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
        // in the arithmetic type of the storage (FP32 below FP64 storage): the
        // kernels use the porosity converted to it, so the values are the same
        if (L.desc.storage == PLBAMD_F64) {
            PLBAMD_HIP(hipMalloc(&L.dAlpha, alpha.size() * sizeof(double)));
            PLBAMD_HIP(hipMemcpy(L.dAlpha, alpha.data(), alpha.size() * sizeof(double), hipMemcpyHostToDevice));
        } else {
            std::vector<float> a32(alpha.begin(), alpha.end());
            PLBAMD_HIP(hipMalloc(&L.dAlpha, a32.size() * sizeof(float)));
            PLBAMD_HIP(hipMemcpy(L.dAlpha, a32.data(), a32.size() * sizeof(float), hipMemcpyHostToDevice));
        }
    }
    L.dirty = false;
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

// The recording step of the adjoint (FP64 and 16-bit storage).
int recordStep(Lattice& L, bool odd, const Recording& r) {
    const bool bgk = L.desc.dynamics == PLBAMD_BGK;
    switch (L.desc.storage) {
    case PLBAMD_F64: return bgk ? launchStepTyped<double, 0>(L, odd, false, &r) : launchStepTyped<double, 1>(L, odd, false, &r);
    case PLBAMD_F16: return bgk ? launchStepTyped<f16s, 0>(L, odd, false, &r) : launchStepTyped<f16s, 1>(L, odd, false, &r);
    case PLBAMD_BF16: return bgk ? launchStepTyped<bf16s, 0>(L, odd, false, &r) : launchStepTyped<bf16s, 1>(L, odd, false, &r);
    default: return fail(PLBAMD_EINVAL, "adjoint: FP64 or 16-bit storage");
    }
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
    // the device address of the fixed scratch of plbamd_reduce_moments_noalloc
    return scratchSetup(device);
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

// Blocks of a tiling that run as one: on one device a dense tiling needs no
// envelopes, so blocks are merged while every merged block stays below 2^32
// cells (the per-block index limit) -- one block whenever the domain fits. Above
// that, z-adjacent blocks with the same x and y extents merge first (their
// envelopes are the costly z-face ones), then y, then x.
// PLBAMD_MAX_BLOCK_CELLS lowers the limit (at least 64 cells), so that the
// merging of tilings above it can be exercised on small lattices.
static std::vector<plbamd_box> coalesceBlocks(const plbamd_lattice_desc& d, std::vector<plbamd_box> b) {
    uint64_t limit = 1ull << 32;
    if (const char* e = getenv("PLBAMD_MAX_BLOCK_CELLS"))
        limit = std::max<uint64_t>(64, std::min<uint64_t>(limit, strtoull(e, nullptr, 10)));
    if ((uint64_t)d.nx * d.ny * d.nz < limit) return {plbamd_box{0, d.nx - 1, 0, d.ny - 1, 0, d.nz - 1}};
    if (b.size() > 4096) return b;   // pairwise search: large block lists stay as given
    auto lo = [](const plbamd_box& x, int a) { return a == 0 ? x.x0 : (a == 1 ? x.y0 : x.z0); };
    auto hi = [](const plbamd_box& x, int a) { return a == 0 ? x.x1 : (a == 1 ? x.y1 : x.z1); };
    auto cells = [&](const plbamd_box& x) {
        return (uint64_t)(x.x1 - x.x0 + 1) * (uint64_t)(x.y1 - x.y0 + 1) * (uint64_t)(x.z1 - x.z0 + 1);
    };
    for (bool merged = true; merged;) {
        merged = false;
        for (int axis = 2; axis >= 0 && !merged; --axis)
            for (size_t i = 0; i < b.size() && !merged; ++i)
                for (size_t j = 0; j < b.size() && !merged; ++j) {
                    if (i == j || hi(b[i], axis) + 1 != lo(b[j], axis)) continue;
                    bool same = true;
                    for (int a = 0; a < 3; ++a)
                        if (a != axis && (lo(b[i], a) != lo(b[j], a) || hi(b[i], a) != hi(b[j], a))) same = false;
                    if (!same || cells(b[i]) + cells(b[j]) >= limit) continue;
                    if (axis == 0) b[i].x1 = b[j].x1;
                    else if (axis == 1) b[i].y1 = b[j].y1;
                    else b[i].z1 = b[j].z1;
                    b.erase(b.begin() + (std::ptrdiff_t)j);
                    merged = true;
                }
    }
    return b;
}

int plbamd_multiblock_create(const plbamd_lattice_desc* desc, const plbamd_box* blocks, int nblocks, int* handle) {
    return plbamd_multiblock_create_ex(desc, blocks, nblocks, 0, handle);
}

int plbamd_multiblock_create_ex(const plbamd_lattice_desc* desc, const plbamd_box* blocks, int nblocks, int flags,
                                int* handle) {
    if (!desc || !handle || !blocks || nblocks <= 0) return fail(PLBAMD_EINVAL, "multiblock_create: bad argument");
    if (flags & ~PLBAMD_KEEP_BLOCKS) return fail(PLBAMD_EINVAL, "multiblock_create: unknown flags");
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
    std::vector<plbamd_box> run(blocks, blocks + nblocks);
    if (!(flags & PLBAMD_KEEP_BLOCKS)) run = coalesceBlocks(*desc, std::move(run));
    auto L = std::make_unique<Lattice>();
    L->desc = *desc;
    L->multi = run.size() > 1;
    L->cells = cells;
    L->bulk = std::move(run);
    PLBAMD_TRY(setup(*L));
    return registerLattice(std::move(L), handle);
}

// What the external-moment processors support here: a single-block lattice of
// bulk BGK / RBGK cells with BounceBack and NoDynamics cells.
static int externalSupported(Lattice& L, const char* who) {
    PLBAMD_TRY(rebuild(L));
    if (L.multi || L.nSpecial || L.dWalls || L.dAlpha || L.dProc || L.desc.force[0] != 0 || L.desc.force[1] != 0 ||
        L.desc.force[2] != 0)
        return fail(PLBAMD_EINVAL, std::string(who) + ": external rhoBar / j need a single-block lattice of bulk, "
                                                      "BounceBack and NoDynamics cells (no boundary objects, "
                                                      "momentum-exchange nodes, link walls, body force or porosity)");
    return PLBAMD_OK;
}

static int externalFieldOp(int handle, const plbamd_box* box, double* rhoBar, double* j, int op, const char* who) {
    Lattice* L = find(handle);
    if (!L) return fail(PLBAMD_EINVAL, std::string(who) + ": unknown handle");
    if (!box || !boxValid(*L, *box)) return fail(PLBAMD_EINVAL, std::string(who) + ": box outside the lattice");
    if (op != 0 && (!rhoBar || !j)) return fail(PLBAMD_EINVAL, std::string(who) + ": null array");
    PLBAMD_TRY(externalSupported(*L, who));
    switch (L->desc.storage) {
    case PLBAMD_F64: return externalField<double>(*L, *box, rhoBar, j, op);
    case PLBAMD_F32: return externalField<float>(*L, *box, rhoBar, j, op);
    case PLBAMD_F16: return externalField<f16s>(*L, *box, rhoBar, j, op);
    default: return externalField<bf16s>(*L, *box, rhoBar, j, op);
    }
}

int plbamd_compute_rhobar_j(int handle, const plbamd_box* box) {
    return externalFieldOp(handle, box, nullptr, nullptr, 0, "compute_rhobar_j");
}

int plbamd_get_rhobar_j(int handle, const plbamd_box* box, double* rhoBar, double* j) {
    return externalFieldOp(handle, box, rhoBar, j, 1, "get_rhobar_j");
}

int plbamd_set_rhobar_j(int handle, const plbamd_box* box, const double* rhoBar, const double* j) {
    return externalFieldOp(handle, box, const_cast<double*>(rhoBar), const_cast<double*>(j), 2, "set_rhobar_j");
}

int plbamd_external_collide_and_stream(int handle, int nsteps) {
    Lattice* L = find(handle);
    if (!L) return fail(PLBAMD_EINVAL, "external_collide_and_stream: unknown handle");
    if (nsteps < 0) return fail(PLBAMD_EINVAL, "external_collide_and_stream: negative step count");
    PLBAMD_TRY(externalSupported(*L, "external_collide_and_stream"));
    // at a non-periodic edge Palabos's processor streams through the block's
    // envelope layer (populations enter from cells outside the domain)
    if (!L->desc.periodic[0] || !L->desc.periodic[1] || !L->desc.periodic[2])
        return fail(PLBAMD_EINVAL, "external_collide_and_stream: periodic lattices only (at a non-periodic edge "
                                   "ExternalRhoJcollideAndStream3D streams through the block envelope; use "
                                   "BounceBack cells as walls)");
    switch (L->desc.storage) {
    case PLBAMD_F64: return externalSteps<double>(*L, nsteps);
    case PLBAMD_F32: return externalSteps<float>(*L, nsteps);
    case PLBAMD_F16: return externalSteps<f16s>(*L, nsteps);
    default: return externalSteps<bf16s>(*L, nsteps);
    }
}

int plbamd_get_block_count(int handle, int* count) {
    if (!count) return fail(PLBAMD_EINVAL, "get_block_count: null output");
    Lattice* L = find(handle);
    if (!L) return fail(PLBAMD_EINVAL, "get_block_count: unknown handle");
    *count = (int)L->bulk.size();
    return PLBAMD_OK;
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

static int addBoundaryObject(Lattice& L, const Plane& p, int* id) {
    if (L.desc.force[0] != 0 || L.desc.force[1] != 0 || L.desc.force[2] != 0)
        return fail(PLBAMD_EINVAL, "boundaries are not combined with a body force");
    L.planes.push_back(p);
    if (id) *id = (int)L.planes.size() - 1;
    L.dirty = true;
    return PLBAMD_OK;
}

int plbamd_add_velocity_edge(int handle, int plane, int normal1, int normal2, const plbamd_box* box, int* id) {
    Lattice* L = find(handle);
    if (!L) return fail(PLBAMD_EINVAL, "add_velocity_edge: unknown handle");
    if (!box || !boxValid(*L, *box)) return fail(PLBAMD_EINVAL, "add_velocity_edge: box outside the lattice");
    if (plane < 0 || plane > 2 || (normal1 != -1 && normal1 != 1) || (normal2 != -1 && normal2 != 1))
        return fail(PLBAMD_EINVAL, "add_velocity_edge: plane in 0..2, normals -1 or +1");
    const int lo[3] = {box->x0, box->y0, box->z0}, hi[3] = {box->x1, box->y1, box->z1};
    const int d1 = (plane + 1) % 3, d2 = (plane + 2) % 3;
    if (lo[d1] != hi[d1] || lo[d2] != hi[d2])
        return fail(PLBAMD_EINVAL, "add_velocity_edge: the box must be one cell thick across the edge");
    Plane p;
    p.family = PLBAMD_REGULARIZED; p.kind = PLBAMD_VELOCITY; p.dir = plane; p.ori = 0; p.box = *box; p.wrapped = false;
    p.shape = SHAPE_EDGE;
    p.normal[d1] = normal1;
    p.normal[d2] = normal2;
    p.values.assign(boxCells(*box) * 3, 0.0);
    return addBoundaryObject(*L, p, id);
}

int plbamd_add_velocity_corner(int handle, int x_normal, int y_normal, int z_normal, int x, int y, int z, int* id) {
    Lattice* L = find(handle);
    if (!L) return fail(PLBAMD_EINVAL, "add_velocity_corner: unknown handle");
    const plbamd_box b{x, x, y, y, z, z};
    if (!boxValid(*L, b)) return fail(PLBAMD_EINVAL, "add_velocity_corner: cell outside the lattice");
    const int n[3] = {x_normal, y_normal, z_normal};
    for (int a = 0; a < 3; ++a)
        if (n[a] != -1 && n[a] != 1) return fail(PLBAMD_EINVAL, "add_velocity_corner: normals -1 or +1");
    Plane p;
    p.family = PLBAMD_REGULARIZED; p.kind = PLBAMD_VELOCITY; p.dir = 0; p.ori = 0; p.box = b; p.wrapped = false;
    p.shape = SHAPE_CORNER;
    for (int a = 0; a < 3; ++a) p.normal[a] = n[a];
    p.values.assign(3, 0.0);
    return addBoundaryObject(*L, p, id);
}

int plbamd_set_boundary_type(int handle, int id, int bc_type) {
    Lattice* L = find(handle);
    if (!L) return fail(PLBAMD_EINVAL, "set_boundary_type: unknown handle");
    if (id < 0 || id >= (int)L->planes.size()) return fail(PLBAMD_EINVAL, "set_boundary_type: unknown boundary");
    if (bc_type < PLBAMD_BC_DIRICHLET || bc_type > PLBAMD_BC_NORMAL_OUTFLOW)
        return fail(PLBAMD_EINVAL, "set_boundary_type: unknown boundary type");
    L->planes[id].bcType = bc_type;
    L->dirty = true;
    return PLBAMD_OK;
}

int plbamd_execute_processors(int handle) {
    Lattice* L = find(handle);
    if (!L) return fail(PLBAMD_EINVAL, "execute_processors: unknown handle");
    PLBAMD_TRY(rebuild(*L));
    if (!L->dProc) return PLBAMD_OK;
    const int parity = (int)(L->time & 1);
    switch (L->desc.storage) {
    case PLBAMD_F64: PLBAMD_TRY(launchProcessors<double>(*L, parity)); break;
    case PLBAMD_F32: PLBAMD_TRY(launchProcessors<float>(*L, parity)); break;
    case PLBAMD_F16: PLBAMD_TRY(launchProcessors<f16s>(*L, parity)); break;
    default: PLBAMD_TRY(launchProcessors<bf16s>(*L, parity)); break;
    }
    PLBAMD_HIP(hipStreamSynchronize(L->stream));
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

int plbamd_set_adjoint_record_bytes(int handle, int64_t bytes) {
    Lattice* L = find(handle);
    if (!L || bytes < -1) return fail(PLBAMD_EINVAL, "set_adjoint_record_bytes: bad argument");
    L->adjointRecordBytes = bytes;
    return PLBAMD_OK;
}

int plbamd_get_adjoint_record_info(int handle, int* kept_segments, int64_t* bytes) {
    Lattice* L = find(handle);
    if (!L || !kept_segments || !bytes) return fail(PLBAMD_EINVAL, "get_adjoint_record_info: bad argument");
    *kept_segments = L->adjointKept;
    *bytes = L->adjointKeptBytes;
    return PLBAMD_OK;
}

int plbamd_get_time(int handle, int64_t* t) {
    Lattice* L = find(handle);
    if (!L || !t) return fail(PLBAMD_EINVAL, "get_time: bad argument");
    *t = L->time;
    return PLBAMD_OK;
}

}  // extern "C"
