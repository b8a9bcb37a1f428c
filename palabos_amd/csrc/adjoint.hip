// palabos_amd: the adjoint gradient (adjoint.h) — forward T steps with a
// checkpoint every K steps, then segments backwards: recompute the segment from
// its checkpoint with the forward kernels, recording per step the pre-collision
// moments (rhoBar, j) of every cell and the populations of the plane cells, then
// K reverse steps. The last segments, as many as the recording budget holds,
// are recorded during the forward sweep instead and need no recompute (the
// same kernels on the same states: the same records).
#include "engine.h"

namespace plbamd {

int adjointGradient(Lattice& L, int T, int K, const plbamd_objective& obj, double* Jout, double* gradAlpha,
                    double* gradScale, double* timing) {
    const int st = L.desc.storage;
    if (st != PLBAMD_F64 && st != PLBAMD_F16 && st != PLBAMD_BF16)
        return fail(PLBAMD_EINVAL, "adjoint: FP64 or 16-bit storage");
    const bool forced = L.desc.force[0] != 0 || L.desc.force[1] != 0 || L.desc.force[2] != 0;
    if (!L.hostTags.empty()) {
        // link-wise walls only: their reflection is the transposed streaming's,
        // and their solid cells take no part
        for (uint8_t t : L.hostTags)
            if (t != PLBAMD_TAG_BULK && t != PLBAMD_TAG_WALL_LINK)
                return fail(PLBAMD_EINVAL, "adjoint: of the per-cell tags only link-wise walls are differentiated "
                                           "(not BounceBack, NoDynamics or momentum exchange)");
        if (!L.planes.empty()) return fail(PLBAMD_EINVAL, "adjoint: link-wise walls with boundary planes are not differentiated");
    }
    if (T <= 0 || K <= 0 || (K & 1)) return fail(PLBAMD_EINVAL, "adjoint: steps > 0 and an even checkpoint interval");
    if (obj.first_step < 0 || obj.first_step >= T) return fail(PLBAMD_EINVAL, "adjoint: first_step in [0, nsteps)");
    if (L.time & 1) return fail(PLBAMD_ESTATE, "adjoint: start after an even number of steps");
    if (obj.inlet_x < 0 || obj.inlet_x >= L.desc.nx || obj.outlet_x < 0 || obj.outlet_x >= L.desc.nx)
        return fail(PLBAMD_EINVAL, "adjoint: measuring planes outside the lattice");
    for (const Plane& p : L.planes)
        if (p.shape != SHAPE_PLANE || p.bcType != PLBAMD_BC_DIRICHLET)
            return fail(PLBAMD_EINVAL, "adjoint: boundary edges, corners and non-Dirichlet boundary types are not differentiated");
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
    // The porosity gradient: FP64 total; a 16-bit adjoint accumulates each segment
    // in FP32 (gA32, 8 bytes per cell and step less) and adds it to the total.
    char *ck = nullptr, *work = nullptr, *lamA = nullptr, *lamB = nullptr, *mom = nullptr;
    double *pf = nullptr, *gA = nullptr, *gS = nullptr, *acc = nullptr, *sum = nullptr;
    float* gA32 = nullptr;
    char* keepMom = nullptr;    // records of the segments recorded in the forward sweep
    double* keepPf = nullptr;
    uint16_t* dAdjBlock = nullptr;
    uint32_t* dAdjBegin = nullptr;
    auto release = [&]() {
        hipFree(ck); hipFree(work); hipFree(mom); hipFree(pf); hipFree(lamA); hipFree(lamB); hipFree(gA); hipFree(gA32);
        hipFree(keepMom); hipFree(keepPf);
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
    if (err == hipSuccess && !lam64) err = hipMalloc(&gA32, cells * sizeof(float));
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
    // Segments recorded in the forward sweep: the last nKeep, whose records
    // (len per segment, one more for f^T) fit the budget. Record r of kept
    // segment c is keep[keepFirst[c] + r].
    const uint64_t recBytes = momElems * lamSize + pfElems * sizeof(double);
    uint64_t budget = 0;
    if (L.adjointRecordBytes >= 0) {
        budget = (uint64_t)L.adjointRecordBytes;
    } else {
        size_t freeB = 0, totalB = 0;
        if (hipMemGetInfo(&freeB, &totalB) == hipSuccess) {
            const uint64_t reserve = std::max<uint64_t>(4ull << 30, totalB / 32);
            budget = freeB > reserve ? freeB - reserve : 0;
        } else {
            (void)hipGetLastError();
        }
    }
    int nKeep = 0;
    uint64_t keepRecs = 1;
    while (nKeep < nCk) {
        const int c = nCk - 1 - nKeep;
        const uint64_t more = keepRecs + (uint64_t)std::min(K, T - c * K);
        if (more * recBytes > budget) break;
        keepRecs = more;
        ++nKeep;
    }
    if (nKeep > 0) {
        if (hipMalloc(&keepMom, keepRecs * momElems * lamSize) != hipSuccess ||
            hipMalloc(&keepPf, keepRecs * pfElems * sizeof(double)) != hipSuccess) {
            // the budget was not there after all: recompute everything
            (void)hipGetLastError();
            hipFree(keepMom);
            hipFree(keepPf);
            keepMom = nullptr;
            keepPf = nullptr;
            nKeep = 0;
        }
    }
    const int keepFirst = nCk - nKeep;   // first kept segment
    auto keepMomAt = [&](int c, int r) {
        return keepMom + ((uint64_t)(c - keepFirst) * K + (uint64_t)r) * momElems * lamSize;
    };
    auto keepPfAt = [&](int c, int r) { return keepPf + ((uint64_t)(c - keepFirst) * K + (uint64_t)r) * pfElems; };
    L.adjointKept = nKeep;
    L.adjointKeptBytes = nKeep > 0 ? (int64_t)(keepRecs * recBytes) : 0;
    if (err == hipSuccess) err = hipMemsetAsync(gA, 0, cells * sizeof(double), L.stream);
    if (err == hipSuccess && gA32) err = hipMemsetAsync(gA32, 0, cells * sizeof(float), L.stream);
    if (err == hipSuccess) err = hipMemsetAsync(gS, 0, std::max<uint32_t>(1, L.nSpecial) * sizeof(double), L.stream);
    if (err == hipSuccess) err = hipMemsetAsync(acc, 0, nAcc * sizeof(double), L.stream);
    // cells the odd steps skip (outermost envelope rows) keep zero moments
    if (err == hipSuccess) err = hipMemsetAsync(mom, 0, momElems * (K + 1) * lamSize, L.stream);
    const bool saveStats = L.statsOn;
    L.statsOn = false;
    float msF = 0, msB = 0;
    if (err == hipSuccess && nKeep > 0) err = hipMemsetAsync(keepMom, 0, keepRecs * momElems * lamSize, L.stream);
    // forward, with checkpoints (even time: the AA array is in natural layout);
    // the kept segments record their steps instead
    if (err == hipSuccess) err = hipEventRecord(L.ev0, L.stream);
    for (int c = 0; c < nCk && err == hipSuccess; ++c) {
        const int len = std::min(K, T - c * K);
        if (c < keepFirst) {
            err = hipMemcpyAsync(ck + (uint64_t)c * stateBytes, L.arena, stateBytes, hipMemcpyDeviceToDevice, L.stream);
            if (err != hipSuccess) break;
            if (step(L, len) != PLBAMD_OK) err = hipErrorUnknown;
            continue;
        }
        for (int m = 0; m < len && err == hipSuccess; ++m) {
            Recording r;
            r.moments = keepMomAt(c, m);
            r.momStride = cells;
            r.specialF = L.nSpecial ? keepPfAt(c, m) : nullptr;
            if (recordStep(L, (L.time & 1) != 0, r) != PLBAMD_OK) err = hipErrorUnknown;
            L.time += 1;
        }
    }
    if (err == hipSuccess && nKeep > 0) {
        // the record of f^T, on a copy (the lattice stays at step T)
        err = hipMemcpyAsync(work, L.arena, stateBytes, hipMemcpyDeviceToDevice, L.stream);
        Recording r;
        r.moments = keepMomAt(nCk - 1, T - (nCk - 1) * K);
        r.momStride = cells;
        r.specialF = L.nSpecial ? keepPfAt(nCk - 1, T - (nCk - 1) * K) : nullptr;
        void* const real = L.arena;
        L.arena = work;
        if (err == hipSuccess && recordStep(L, (L.time & 1) != 0, r) != PLBAMD_OK) err = hipErrorUnknown;
        L.arena = real;
    }
    if (err == hipSuccess) err = hipEventRecord(L.ev1, L.stream);
    if (err == hipSuccess) err = hipEventSynchronize(L.ev1);
    if (err == hipSuccess) err = hipEventElapsedTime(&msF, L.ev0, L.ev1);
    // backward
    AdjointArgs a{};
    a.momStride = cells;
    a.alpha = L.dAlpha;
    a.gAlpha = lam64 ? (void*)gA : (void*)gA32;
    a.tags = L.dTags;
    a.walls = L.dWalls;
    a.solidSkip = L.hostTags.empty() ? 0 : 1;
    // the forward's relaxation rate: 16-bit BGK relaxes with omega' (exact_rates)
    a.omega = (st != PLBAMD_F64 && L.desc.dynamics == PLBAMD_BGK) ? (double)exact_rates((float)L.desc.omega).omega
                                                                  : L.desc.omega;
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
    const bool walls = a.walls != nullptr;
#define PLBAMD_ADJ(LT, W) \
    do { \
        if (rbgk) { \
            if (forced) hipLaunchKernelGGL((adjoint_bulk<LT, 1, true, W>), gr, ablock, 0, L.stream, args); \
            else hipLaunchKernelGGL((adjoint_bulk<LT, 1, false, W>), gr, ablock, 0, L.stream, args); \
        } else { \
            if (forced) hipLaunchKernelGGL((adjoint_bulk<LT, 0, true, W>), gr, ablock, 0, L.stream, args); \
            else hipLaunchKernelGGL((adjoint_bulk<LT, 0, false, W>), gr, ablock, 0, L.stream, args); \
        } \
    } while (0)
    auto reverseKernel = [&](const AdjointArgs& args, const dim3& gr) {
        if (lam64) {
            if (walls) PLBAMD_ADJ(double, true);
            else PLBAMD_ADJ(double, false);
        } else {
            if (walls) PLBAMD_ADJ(float, true);
            else PLBAMD_ADJ(float, false);
        }
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
        // the segment's records: kept from the forward sweep, or recomputed
        char* segMom = mom;
        double* segPf = pf;
        if (c >= keepFirst) {
            segMom = keepMomAt(c, 0);
            segPf = keepPfAt(c, 0);
        } else {
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
        }
        if (err == hipSuccess && c == nCk - 1) {
            a.mom = segMom + (uint64_t)len * momElems * lamSize;
            a.explicitTerm = 1;
            a.lam = lamNext;
            if (lam64) hipLaunchKernelGGL(adjoint_final<double>, agrid, ablock, 0, L.stream, a);
            else hipLaunchKernelGGL(adjoint_final<float>, agrid, ablock, 0, L.stream, a);
            err = hipGetLastError();
            exchangeLam(lamNext);
        }
        for (int m = len - 1; m >= 0 && err == hipSuccess; --m) {
            const int t = t0 + m;
            a.mom = segMom + (uint64_t)m * momElems * lamSize;
            a.planeF = segPf + (uint64_t)m * pfElems;
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
        if (err == hipSuccess && gA32) {
            hipLaunchKernelGGL(flush_gradient, dim3((unsigned)((cells + 255) / 256)), dim3(256), 0, L.stream, gA32, gA,
                               cells);
            err = hipGetLastError();
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

}  // namespace plbamd
