// palabos_amd: the step launches of one storage type and one dynamics,
// compiled once per (PLBAMD_S, PLBAMD_DYN) so that the kernel variants build in
// parallel translation units.
#include "engine.h"

#if !defined(PLBAMD_S) || !defined(PLBAMD_DYN)
#error "compile with -DPLBAMD_S=<storage type> -DPLBAMD_DYN=<0 BGK | 1 RBGK>"
#endif

namespace plbamd {

template<typename S, int DYN>
int launchStepTyped(Lattice& L, bool odd, bool stats, const Recording* record) {
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
        if (O && L.staged && L.cachedZ)                                                                         \
            hipLaunchKernelGGL((bulk_odd_staged<double, DYN, TG, false, FO, true, true>), grid, L.bulkBlock,     \
                               L.stagedLds, L.stream, a);                                                       \
        else if (O && L.staged)                                                                                 \
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
        else if (O && L.cachedZ)                                                                                \
            hipLaunchKernelGGL((bulk_pairs<S, DYN, O, TG, false, FO, true, false, false, O>), grid,           \
                               L.bulkBlock, L.pairsLds, L.stream, a);                                           \
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
        if (L.packed && O && L.cachedZ)                                                                         \
            hipLaunchKernelGGL((bulk_pairs<S, DYN, O, TG, ST, FO, false, true, false, O>), grid, L.bulkBlock, \
                               L.pairsLds, L.stream, a);                                                        \
        else if (L.packed)                                                                                      \
            hipLaunchKernelGGL((bulk_pairs<S, DYN, O, TG, ST, FO, false, true>), grid, L.bulkBlock,             \
                               O ? L.pairsLds : 0, L.stream, a);                                                \
        else if (O && L.zWrap && L.cachedZ)                                                                     \
            hipLaunchKernelGGL((bulk_pairs<S, DYN, O, TG, ST, FO, false, false, O, O>), grid, L.bulkBlock, \
                               L.pairsLds, L.stream, a);                                                        \
        else if (O && L.zWrap)                                                                                  \
            hipLaunchKernelGGL((bulk_pairs<S, DYN, O, TG, ST, FO, false, false, true>), grid, L.bulkBlock,      \
                               L.pairsLds, L.stream, a);                                                        \
        else if (O && L.cachedZ)                                                                                \
            hipLaunchKernelGGL((bulk_pairs<S, DYN, O, TG, ST, FO, false, false, false, O>), grid,             \
                               L.bulkBlock, L.pairsLds, L.stream, a);                                           \
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
        if (O && L.staged && L.cachedZ)                                                                         \
            hipLaunchKernelGGL((bulk_odd_staged<S, DYN, TG, ST, FO, false, true>), grid, L.bulkBlock,            \
                               L.stagedLds, L.stream, a);                                                       \
        else if (O && L.staged)                                                                                 \
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
    // data processors after the stream: Neumann-type copies, edge and corner closures
    if (L.dProc && !record) PLBAMD_TRY(launchProcessors<S>(L, odd ? 0 : 1));
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

template int launchStepTyped<PLBAMD_S, PLBAMD_DYN>(Lattice&, bool, bool, const Recording*);

}  // namespace plbamd
