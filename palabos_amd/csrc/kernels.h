// Device kernels of the engine: AA-pattern collide-and-stream over a table of
// blocks, the special-cell kernel, the envelope exchange, transfers, analysis.
//
// Data layout in HBM. A lattice is one or more blocks in one arena. A block
// stores its cells (bulk plus envelope) slot-major, f[slot][x][y][z], z
// fastest, each slot padded to a multiple of 64 cells, in the storage type
// (double | float | f16 bits | bf16 bits) of shifted populations f - t_i
// (roundOffPolicy.h:73-90). Per-cell tags (uint8) and porosity (FP64) follow
// the same cell order. A single-block lattice has no envelope and wraps its
// periodic axes in place; a multi-block lattice gives every block a two-cell
// envelope towards its neighbours and periodic images, refreshed by the
// exchange kernel every second step.
//
// Streaming is the in-place AA pattern, one array:
//   even step  each cell reads its own slots i, collides, writes f*_i to its
//              own slot opp(i)                        (all accesses local)
//   odd step   each cell gathers f_i from slot opp(i) of cell x - c_i, collides,
//              scatters f*_i to slot i of cell x + c_i (natural layout again)
// A source or destination outside a non-periodic edge of the block (or a
// link-wise wall) is replaced by the cell's own opposite slot: Palabos's
// same-step edge reflection (atomicBlock/blockLattice3D.hh:403-427). Every
// memory slot is read and written by exactly one cell per step, so the update
// is race-free and the cell order is free, as it is in Palabos.
//
// Envelopes. An even step is local, so it corrupts nothing; an odd step
// corrupts the outermost layer (its gathers leave the block) and, through its
// scatters, the layer inside it. A two-cell envelope refreshed after every odd
// step therefore keeps every bulk cell bit-identical to a single-block run, as
// Palabos's one-cell envelope refreshed every step does
// (multiBlock/multiBlockLattice3D.hh:318-360).
#pragma once

#include <type_traits>

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/palabos_amd.h"
#include "boundary_device.h"
#include "d3q19_device.h"
#include "storage.h"

namespace plbamd {

// BT_WALL_NEAR: background dynamics next to a link-wise wall; the per-cell link
// mask (bit i: neighbour x + c_i is a wall) selects the reflected populations.
enum BulkTag : uint8_t { BT_BULK = 0, BT_BOUNCE_BACK = 1, BT_NO_DYNAMICS = 2, BT_SKIP = 3, BT_WALL_NEAR = 4 };

// One block of a lattice. Coordinates are block-local (0..n-1 over the
// allocated cells); the bulk (cells that own their values and take
// statistics) is [b0, b1] inclusive.
struct BlockGeom {
    int nx, ny, nz;           // allocated extent
    int px, py, pz;           // wrap inside the block (single-block periodic axes)
    int bx0, by0, bz0;        // bulk begin, local
    int bx1, by1, bz1;        // bulk end, local, inclusive
    int gx0, gy0, gz0;        // global coordinate of local (0,0,0), before periodic wrap
    int nzChunks, nyGroups;   // bulk-launch decomposition: row segments, y groups
    int rowsPerWG;            // packed-row launches: whole rows per workgroup
    uint64_t offset;          // element offset of slot 0 in the arena
    uint64_t slot;            // elements per slot
    uint64_t cellOffset;      // offset of the block's cells in per-cell arrays
    uint32_t wgBegin;         // first workgroup of the block in the bulk launch
    uint32_t evenBegin;       // first workgroup of the block in the linear even-step launch
};

__device__ __forceinline__ uint64_t lin(const BlockGeom& g, int x, int y, int z) {
    return ((uint64_t)x * g.ny + y) * g.nz + z;
}

__device__ __forceinline__ bool shift(int v, int d, int n, int periodic, int& out) {
    int w = v + d;
    if (w < 0) { if (!periodic) return false; w += n; }
    else if (w >= n) { if (!periodic) return false; w -= n; }
    out = w;
    return true;
}

// A row on the outermost layer of a block's envelope along x or y (not a
// domain edge: those blocks have no envelope on that side).
__device__ __forceinline__ bool outer_envelope_row(const BlockGeom& g, int x, int y) {
    return (g.bx0 > 0 && x == 0) || (g.bx1 < g.nx - 1 && x == g.nx - 1) ||
           (g.by0 > 0 && y == 0) || (g.by1 < g.ny - 1 && y == g.ny - 1);
}

__device__ __forceinline__ bool in_bulk(const BlockGeom& g, int x, int y, int z) {
    return x >= g.bx0 && x <= g.bx1 && y >= g.by0 && y <= g.by1 && z >= g.bz0 && z <= g.bz1;
}

// Position of f_i of cell (x,y,z) in the swapped (after-even-step) layout:
// slot opp(i) at x - c_i, or slot i at the cell itself when the source reflects.
__device__ __forceinline__ uint64_t swapped_index(const BlockGeom& g, int x, int y, int z, int i,
                                                  uint32_t wallMask, uint64_t cell) {
    if (i == 0) return cell;
    int sx, sy, sz;
    bool ok = shift(x, -cx(i), g.nx, g.px, sx) && shift(y, -cy(i), g.ny, g.py, sy) &&
              shift(z, -cz(i), g.nz, g.pz, sz) && !((wallMask >> opp(i)) & 1u);
    if (ok) return (uint64_t)opp(i) * g.slot + lin(g, sx, sy, sz);
    return (uint64_t)i * g.slot + cell;
}

__device__ __forceinline__ uint64_t pop_index(const BlockGeom& g, int parity, int x, int y, int z, int i,
                                              uint32_t wall, uint64_t cell) {
    if (parity == 0) return (uint64_t)i * g.slot + cell;
    return swapped_index(g, x, y, z, i, wall, cell);
}

// Block of a workgroup of the bulk launch (binary search over wgBegin).
__device__ __forceinline__ int find_block(const BlockGeom* blocks, int nBlocks, uint32_t wg) {
    int lo = 0, hi = nBlocks - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (blocks[mid].wgBegin <= wg) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Block reduction of an 8-double record (fixed order: deterministic): a
// shuffle tree inside each wavefront, then the wavefront sums in order. Slot 2
// is a maximum, the others are sums. Up to 1024 threads per workgroup. Only
// the first N entries are reduced; the others are known to be zero and are
// written as zero.
template<int N = 8>
__device__ __forceinline__ void block_reduce_store(double v[8], double* out) {
    __shared__ double red[16][8];
    const int t = threadIdx.x + threadIdx.y * blockDim.x;
    const int n = blockDim.x * blockDim.y;
    const int lane = t & 63, wave = t >> 6, nWaves = (n + 63) >> 6;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
        for (int k = 0; k < N; ++k) {
            const double o = __shfl_down(v[k], off, 64);
            if (lane + off < 64 && t + off < n) v[k] = (k == 2) ? fmax(v[k], o) : v[k] + o;
        }
    }
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < N; ++k) red[wave][k] = v[k];
    }
    __syncthreads();
    if (t == 0) {
        double r[8];
#pragma unroll
        for (int k = 0; k < N; ++k) r[k] = red[0][k];
        for (int w = 1; w < nWaves; ++w) {
#pragma unroll
            for (int k = 0; k < N; ++k) r[k] = (k == 2) ? fmax(r[k], red[w][k]) : r[k] + red[w][k];
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) out[k] = k < N ? r[k] : 0.0;
    }
    __syncthreads();
}

template<typename S>
__device__ __forceinline__ S* at(S* base, uint32_t byteOffset) {
    return reinterpret_cast<S*>(reinterpret_cast<char*>(base) + byteOffset);
}

// ---------------------------------------------------------------------------
// Bulk kernel. One launch covers every block: a workgroup takes one segment
// of zChunk cells of a row (x, y) of one block (TY rows when the row is
// shorter than a workgroup). A wavefront always lies in one (x, y) row, so
// the nine neighbouring rows and their validity are wavefront-uniform and
// live in scalar registers: a load or store is a row base plus one of three
// per-lane z offsets. Wavefronts touching an edge of the block take the
// general path with per-lane edge selection.

struct StepArgs {
    void* arena;
    const BlockGeom* blocks;
    int nBlocks;
    int zChunk;
    double omega;
    double force[3];
    const uint8_t* tags;     // BulkTag per cell (all blocks) or null
    const void* alpha;       // Brinkman porosity per cell (all blocks), in the arithmetic type, or null
    double* partials;        // 8 doubles per workgroup when statistics are on
    void* moments;           // MOMENTS: rhoBar, jx, jy, jz of the pre-collision state (MomentType<S>),
    uint64_t momStride;      //          [4][momStride] by cell (all blocks)
    int xcdRemap;            // 1: consecutive work items on the same XCD
    const uint16_t* wgBlock; // block of every work item (multi-block), or null (one block)
    const uint32_t* walls;   // link-wall mask per cell (all blocks), read for BT_WALL_NEAR cells
    const uint16_t* evenBlock; // linear even-step launch: block of every workgroup, or null
    int ldsPitch;            // packed-row launches: LDS words per staged direction
};

// Block of a work item: one table load (the binary search over the block table
// costs a chain of dependent loads at the start of every workgroup).
__device__ __forceinline__ int block_of(const StepArgs& a, uint32_t wg) {
    return a.wgBlock ? (int)a.wgBlock[wg] : 0;
}

// The dispatcher hands workgroup b to XCD b % 8. With the remap, XCD k takes
// the k-th contiguous eighth of the work items (the z-chunks of a row and the
// neighbouring rows share that XCD's L2); the tail beyond a multiple of 8 maps
// to itself. A bijection on [0, n).
__device__ __forceinline__ uint32_t xcd_item(uint32_t b, uint32_t n) {
    const uint32_t q = n / 8;
    if (b >= 8 * q) return b;
    return (b % 8) * q + b / 8;
}

// Pre-collision moments rhoBar = sum_i f_i, j = sum_i c_i f_i (plain order, FP64
// sums), recorded for the adjoint's reverse sweep in the type of the adjoint
// state: FP64 for an FP64 forward, FP32 for the others (whose populations are
// FP32 values).
template<typename S>
using MomentType = typename std::conditional<std::is_same<S, double>::value, double, float>::type;

template<typename T, typename M>
__device__ __forceinline__ void moments_of(const T* fc, M out[4]) {
    double rb = 0, jx = 0, jy = 0, jz = 0;
#pragma unroll
    for (int i = 0; i < 19; ++i) {
        const double v = (double)fc[i];
        rb += v;
        jx += cx(i) * v;
        jy += cy(i) * v;
        jz += cz(i) * v;
    }
    out[0] = (M)rb;
    out[1] = (M)jx;
    out[2] = (M)jy;
    out[3] = (M)jz;
}

template<typename T, typename M>
__device__ __forceinline__ void store_moments(const T* fc, M* m, uint64_t stride, uint64_t idx) {
    double rb = 0, jx = 0, jy = 0, jz = 0;
#pragma unroll
    for (int i = 0; i < 19; ++i) {
        const double v = (double)fc[i];
        rb += v;
        jx += cx(i) * v;
        jy += cy(i) * v;
        jz += cz(i) * v;
    }
    m[idx] = (M)rb;
    m[stride + idx] = (M)jx;
    m[2 * stride + idx] = (M)jy;
    m[3 * stride + idx] = (M)jz;
}

// The records of a pair of cells (idx even, stride even: 16-bit lattices),
// computed one cell at a time: one 8-byte store per moment where both cells
// record, so that a wavefront's stores cover whole lines.
__device__ __forceinline__ void store_moments_pair(const float ra[4], const float rb[4], bool recA, bool recB,
                                                   float* m, uint64_t stride, uint64_t idx) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        float* p = m + k * stride + idx;
        if (recA && recB) *reinterpret_cast<float2*>(p) = make_float2(ra[k], rb[k]);
        else if (recA) p[0] = ra[k];
        else if (recB) p[1] = rb[k];
    }
}

// Collision of one cell by its bulk tag: background dynamics (with moments
// recording and statistics), bounce-back swap, or nothing (NoDynamics).
template<typename S, int DYN, bool STATS, bool FORCED, bool MOMENTS>
__device__ __forceinline__ void collide_cell(const StepArgs& a, const BlockGeom& g, uint64_t cell, uint8_t tag,
                                             bool inBulk, typename Store<S>::T fc[19], double st[8],
                                             uint32_t wall = 0, double* wf = nullptr) {
    using T = typename Store<S>::T;
    if constexpr (MOMENTS) {
        if (tag == BT_BULK || tag == BT_WALL_NEAR)
            store_moments(fc, reinterpret_cast<MomentType<S>*>(a.moments), a.momStride, g.cellOffset + cell);
    }
    if (tag == BT_BULK || tag == BT_WALL_NEAR) {
        const T omega = (T)a.omega;
        const T F[3] = {(T)a.force[0], (T)a.force[1], (T)a.force[2]};
        T rhoBar;
        T drag[3] = {0, 0, 0};
        const T alpha = (FORCED && a.alpha) ? reinterpret_cast<const T*>(a.alpha)[g.cellOffset + cell] : (T)0;
        const T uSqr = collide_bulk<T, DYN, FORCED, sizeof(S) == 2>(fc, omega, F, alpha, rhoBar, drag);
        if constexpr (STATS) {
            if (inBulk) {
                st[0] += (double)rhoBar;
                st[1] += (double)uSqr;
                st[2] = fmax(st[2], (double)uSqr);
                st[3] += 1.0;
                if constexpr (FORCED) {
                    st[4] += (double)drag[0];
                    st[5] += (double)drag[1];
                    st[6] += (double)drag[2];
                }
                // link-wise wall: each outgoing population towards the wall carries
                // 2 c_i f*_i (the stored, storage-rounded value is what is reflected)
                if (wall) {
#pragma unroll
                    for (int i = 1; i < 19; ++i) {
                        if (!((wall >> i) & 1u)) continue;
                        const T fo = Store<S>::round(fc[i]);
#pragma unroll
                        for (int c = 0; c < 3; ++c) wf[c] += (double)((T)2 * ci(i, c) * fo);
                    }
                }
            }
        }
    } else if (tag == BT_BOUNCE_BACK) {
#pragma unroll
        for (int i = 1; i <= 9; ++i) { T t = fc[i]; fc[i] = fc[i + 9]; fc[i + 9] = t; }
    }
}

// Statistics of a bulk workgroup: two records, [rhoBar uSqr maxUSqr count Px Py Pz -]
// and the link-wall force [Fx Fy - Fz - - - -].
// Without link walls (walls: the lattice's wall masks, null when it has none)
// the force record is zero and written as such, without a reduction; without
// a force (FORCED false) so are the porous-drag entries 4-6.
template<bool STATS, bool FORCED>
__device__ __forceinline__ void store_bulk_stats(double st[8], const double wf[3], double* partials, uint32_t wg,
                                                 const void* walls) {
    if constexpr (STATS) {
        block_reduce_store<FORCED ? 7 : 4>(st, partials + 16 * (uint64_t)wg);
        if (walls) {
            double w[8] = {wf[0], wf[1], 0, wf[2], 0, 0, 0, 0};
            block_reduce_store(w, partials + 16 * (uint64_t)wg + 8);
        } else if (threadIdx.x + threadIdx.y == 0) {
#pragma unroll
            for (int k = 0; k < 8; ++k) partials[16 * (uint64_t)wg + 8 + k] = 0.0;
        }
    }
}

template<typename S, int DYN, bool ODD, bool TAGS, bool STATS, bool FORCED, bool MOMENTS = false>
__global__ void __launch_bounds__(512) bulk_step(StepArgs a) {
    using T = typename Store<S>::T;
    const uint32_t wg = a.xcdRemap ? xcd_item(blockIdx.x, gridDim.x) : blockIdx.x;
    const int b = block_of(a, wg);
    const BlockGeom g = a.blocks[b];
    S* __restrict__ f = reinterpret_cast<S*>(a.arena) + g.offset;
    const uint32_t local = wg - g.wgBegin;
    const int chunk = (int)(local % (uint32_t)g.nzChunks);
    const uint32_t rest = local / (uint32_t)g.nzChunks;
    const int yGroup = (int)(rest % (uint32_t)g.nyGroups);
    const int x = (int)(rest / (uint32_t)g.nyGroups);
    const int y = __builtin_amdgcn_readfirstlane(yGroup * (int)blockDim.y + (int)threadIdx.y);
    const int zBegin = chunk * a.zChunk;
    const int zEnd = min(g.nz, zBegin + a.zChunk);
    double st[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    double wf[3] = {0, 0, 0};

    if (y < g.ny) {
        int xm = 0, xp = 0, ym = 0, yp = 0;
        const bool vxm = shift(x, -1, g.nx, g.px, xm), vxp = shift(x, 1, g.nx, g.px, xp);
        const bool vym = shift(y, -1, g.ny, g.py, ym), vyp = shift(y, 1, g.ny, g.py, yp);
        const int xs[3] = {xm, x, xp};
        const int ys[3] = {ym, y, yp};
        const bool vx[3] = {vxm, true, vxp};
        const bool vy[3] = {vym, true, vyp};
        const bool rowsValid = vxm && vxp && vym && vyp;
        const bool rowInBulkXY = x >= g.bx0 && x <= g.bx1 && y >= g.by0 && y <= g.by1;
        const uint64_t rowBase = ((uint64_t)x * g.ny + y) * g.nz;
        S* rows[3][3];
#pragma unroll
        for (int ax = 0; ax < 3; ++ax)
#pragma unroll
            for (int ay = 0; ay < 3; ++ay)
                rows[ax][ay] = f + ((uint64_t)xs[ax] * g.ny + ys[ay]) * g.nz;

        for (int z = zBegin + (int)threadIdx.x; z < zEnd; z += blockDim.x) {
            const uint64_t cell = rowBase + z;
            uint8_t tag = BT_BULK;
            uint32_t wall = 0;
            if constexpr (TAGS) {
                tag = a.tags[g.cellOffset + cell];
                // the link mask is zero except on wall-near cells: loaded without
                // testing the tag. Odd steps leave skipped cells here (their lanes
                // take no part in the ballot below); even steps test the tag after
                // their population loads are issued.
                if (a.walls) wall = a.walls[g.cellOffset + cell];
                if (ODD && tag == BT_SKIP) continue;
            }
            int zm = 0, zp = 0;
            const bool vzm = shift(z, -1, g.nz, g.pz, zm), vzp = shift(z, 1, g.nz, g.pz, zp);
            const uint32_t zb[3] = {(uint32_t)zm * (uint32_t)sizeof(S), (uint32_t)z * (uint32_t)sizeof(S),
                                    (uint32_t)zp * (uint32_t)sizeof(S)};
            bool fast = true;
            if constexpr (ODD) {
                const uint64_t active = __ballot(1);
                fast = rowsValid && (__ballot(vzm && vzp && wall == 0) == active);
            }

            T fc[19];
            if constexpr (!ODD) {
#pragma unroll
                for (int i = 0; i < 19; ++i) fc[i] = Store<S>::load(at(rows[1][1] + (uint64_t)i * g.slot, zb[1]));
                if constexpr (TAGS) {
                    if (tag == BT_SKIP) continue;
                }
            } else if (fast) {
#pragma unroll
                for (int i = 0; i < 19; ++i)
                    fc[i] = Store<S>::load(at(rows[1 - cx(i)][1 - cy(i)] + (uint64_t)opp(i) * g.slot, zb[1 - cz(i)]));
            } else {
                const int zs[3] = {zm, z, zp};
                const bool vz[3] = {vzm, true, vzp};
                fc[0] = Store<S>::load(f + cell);
#pragma unroll
                for (int i = 1; i < 19; ++i) {
                    const int ax = 1 - cx(i), ay = 1 - cy(i), az = 1 - cz(i);
                    const bool ok = vx[ax] && vy[ay] && vz[az] && !((wall >> opp(i)) & 1u);
                    const uint64_t src = ok ? (uint64_t)opp(i) * g.slot +
                                                  (((uint64_t)xs[ax] * g.ny + ys[ay]) * g.nz + zs[az])
                                            : (uint64_t)i * g.slot + cell;
                    fc[i] = Store<S>::load(f + src);
                }
            }

            collide_cell<S, DYN, STATS, FORCED, MOMENTS>(a, g, cell, tag, rowInBulkXY && z >= g.bz0 && z <= g.bz1,
                                                       fc, st, wall, wf);

            if constexpr (!ODD) {
#pragma unroll
                for (int i = 0; i < 19; ++i) Store<S>::store(at(rows[1][1] + (uint64_t)opp(i) * g.slot, zb[1]), fc[i]);
            } else if (fast) {
#pragma unroll
                for (int i = 0; i < 19; ++i)
                    Store<S>::store(at(rows[1 + cx(i)][1 + cy(i)] + (uint64_t)i * g.slot, zb[1 + cz(i)]), fc[i]);
            } else {
                const int zs[3] = {zm, z, zp};
                const bool vz[3] = {vzm, true, vzp};
                Store<S>::store(f + cell, fc[0]);
#pragma unroll
                for (int i = 1; i < 19; ++i) {
                    const int ax = 1 + cx(i), ay = 1 + cy(i), az = 1 + cz(i);
                    const bool ok = vx[ax] && vy[ay] && vz[az] && !((wall >> i) & 1u);
                    const uint64_t dst = ok ? (uint64_t)i * g.slot +
                                                  (((uint64_t)xs[ax] * g.ny + ys[ay]) * g.nz + zs[az])
                                            : (uint64_t)opp(i) * g.slot + cell;
                    Store<S>::store(f + dst, fc[i]);
                }
            }
        }
    }
    store_bulk_stats<STATS, FORCED>(st, wf, a.partials, wg, a.walls);
}

// ---------------------------------------------------------------------------
// Even step over a block's cells in memory order. The even step is local (each
// cell reads and writes only its own slots), so a workgroup takes a run of
// consecutive cells regardless of rows: every lane is busy even when the
// block's rows are short. 16-bit storage moves two cells per thread with
// 32-bit pair accesses (every block's cell count is even then).

template<typename S, int DYN, bool TAGS, bool STATS, bool FORCED, bool MOMENTS = false>
__global__ void __launch_bounds__(512) bulk_even_linear(StepArgs a) {
    using T = typename Store<S>::T;
    constexpr int CPT = sizeof(typename Store<S>::R) == 2 ? 2 : 1;
    const uint32_t wg = blockIdx.x;
    const int b = a.evenBlock ? (int)a.evenBlock[wg] : 0;
    const BlockGeom g = a.blocks[b];
    S* __restrict__ f = reinterpret_cast<S*>(a.arena) + g.offset;
    const uint64_t cells = (uint64_t)g.nx * g.ny * g.nz;
    const uint64_t c0 = ((uint64_t)(wg - g.evenBegin) * blockDim.x + threadIdx.x) * CPT;
    double st[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    double wf[3] = {0, 0, 0};
    if (c0 < cells) {
        T fc[CPT][19];
        if constexpr (CPT == 2) {
#pragma unroll
            for (int i = 0; i < 19; ++i) {
                const uint32_t v = __builtin_nontemporal_load(reinterpret_cast<const uint32_t*>(f + (uint64_t)i * g.slot + c0));
                fc[0][i] = Store<S>::from_raw((uint16_t)(v & 0xffffu));
                fc[1][i] = Store<S>::from_raw((uint16_t)(v >> 16));
            }
        } else {
#pragma unroll
            for (int i = 0; i < 19; ++i) fc[0][i] = Store<S>::load(f + (uint64_t)i * g.slot + c0);
        }
        bool comp[CPT];
#pragma unroll
        for (int h = 0; h < CPT; ++h) {
            const uint64_t cell = c0 + h;
            uint8_t tag = BT_BULK;
            uint32_t wall = 0;
            if constexpr (TAGS) {
                tag = a.tags[g.cellOffset + cell];
                if constexpr (STATS) {
                    if (tag == BT_WALL_NEAR) wall = a.walls[g.cellOffset + cell];
                }
            }
            comp[h] = tag != BT_SKIP;
            bool inBulk = false;
            if constexpr (STATS) {
                const uint32_t plane = (uint32_t)g.ny * (uint32_t)g.nz;
                const uint32_t cc = (uint32_t)cell;
                const int x = (int)(cc / plane), rem = (int)(cc % plane);
                const int y = rem / g.nz, z = rem % g.nz;
                inBulk = x >= g.bx0 && x <= g.bx1 && y >= g.by0 && y <= g.by1 && z >= g.bz0 && z <= g.bz1;
            }
            if (comp[h]) collide_cell<S, DYN, STATS, FORCED, MOMENTS>(a, g, cell, tag, inBulk, fc[h], st, wall, wf);
        }
        // 16-bit pairs: every pair is written back, a cell not collided here with the
        // value of slot opp(i) it was loaded with (its own thread is the only writer
        // of the pair in an even step; the special-cell kernel runs after this one):
        // the stores leave no holes in the lines they write, which cost a
        // read-modify-write at the memory (porous medium with 7 % solid cells: even
        // step 377 -> 334 us). One cell per thread keeps its conditional stores (the
        // write-back would keep the loaded values live through the collision: 75 ->
        // 103 registers).
        if constexpr (CPT == 2) {
#pragma unroll
            for (int i = 0; i < 19; ++i) {
                const uint32_t lo = Store<S>::to_raw(comp[0] ? fc[0][i] : fc[0][opp(i)]);
                const uint32_t hi = Store<S>::to_raw(comp[1] ? fc[1][i] : fc[1][opp(i)]);
                __builtin_nontemporal_store(lo | (hi << 16), reinterpret_cast<uint32_t*>(f + (uint64_t)opp(i) * g.slot + c0));
            }
        } else if (comp[0]) {
#pragma unroll
            for (int i = 0; i < 19; ++i) Store<S>::store(f + (uint64_t)opp(i) * g.slot + c0, fc[0][i]);
        }
    }
    store_bulk_stats<STATS, FORCED>(st, wf, a.partials, wg, a.walls);
}

// ---------------------------------------------------------------------------
// Odd (streaming) step with the z-moving populations staged through LDS.
//
// In the AA odd step a cell gathers slot opp(i) from x - c_i and scatters to
// slot i at x + c_i. For the 10 directions with c_z != 0 those accesses are
// shifted by one element along z, so every wavefront access straddles one
// more cache line than it uses (one line in two for 16-bit storage). Here a
// workgroup takes one z-segment of one row (one cell per thread): it loads
// the segment of each z-moving source row aligned into LDS (plus one halo
// element at each end), reads the shifted values from LDS, and writes the
// shifted outputs back through LDS so that the global stores are aligned too.
// The 9 other directions are aligned already and go straight to registers.
// Rows at a non-periodic x/y edge of the block take the per-lane general path.

// One cell of an odd step by per-lane neighbour tests (rows at block edges).
template<typename S, int DYN, bool STATS, bool FORCED, bool MOMENTS>
__device__ __forceinline__ void odd_cell_general(const StepArgs& a, const BlockGeom& g, S* f, int z, uint64_t cell,
                                                 uint8_t tag, bool inBulk, const int xs[3], const int ys[3],
                                                 const bool vx[3], const bool vy[3], double st[8], uint32_t wall,
                                                 double wf[3]) {
    using T = typename Store<S>::T;
    int zm = 0, zp = 0;
    const bool vzm = shift(z, -1, g.nz, g.pz, zm), vzp = shift(z, 1, g.nz, g.pz, zp);
    const int zs[3] = {zm, z, zp};
    const bool vz[3] = {vzm, true, vzp};
    T fc[19];
    fc[0] = Store<S>::load(f + cell);
#pragma unroll
    for (int i = 1; i < 19; ++i) {
        const int ax = 1 - cx(i), ay = 1 - cy(i), az = 1 - cz(i);
        const bool ok = vx[ax] && vy[ay] && vz[az] && !((wall >> opp(i)) & 1u);
        const uint64_t src = ok ? (uint64_t)opp(i) * g.slot + (((uint64_t)xs[ax] * g.ny + ys[ay]) * g.nz + zs[az])
                                : (uint64_t)i * g.slot + cell;
        fc[i] = Store<S>::load(f + src);
    }
    collide_cell<S, DYN, STATS, FORCED, MOMENTS>(a, g, cell, tag, inBulk, fc, st, wall, wf);
    Store<S>::store(f + cell, fc[0]);
#pragma unroll
    for (int i = 1; i < 19; ++i) {
        const int ax = 1 + cx(i), ay = 1 + cy(i), az = 1 + cz(i);
        const bool ok = vx[ax] && vy[ay] && vz[az] && !((wall >> i) & 1u);
        const uint64_t dst = ok ? (uint64_t)i * g.slot + (((uint64_t)xs[ax] * g.ny + ys[ay]) * g.nz + zs[az])
                                : (uint64_t)opp(i) * g.slot + cell;
        Store<S>::store(f + dst, fc[i]);
    }
}

// Index of direction i among the z-moving directions, or -1.
__host__ __device__ constexpr int zmove_index(int i) {
    int k = 0;
    for (int j = 0; j < i; ++j) k += cz(j) != 0 ? 1 : 0;
    return cz(i) != 0 ? k : -1;
}
constexpr int kZMoving = 10;

// The m-th (0-4) direction with c_z == SZ and its z-moving index, for a
// per-lane m (a chain of selects: no indexed array).
template<int SZ>
__device__ __forceinline__ void zdir(int m, int& dir, int& k) {
    dir = 0;
    k = 0;
    int n = 0;
#pragma unroll
    for (int i = 1; i < 19; ++i) {
        if (cz(i) == SZ) {
            if (m == n) { dir = i; k = zmove_index(i); }
            ++n;
        }
    }
}

// FP32 statistics variants without forcing: asking for 7 waves per SIMD keeps
// them at 72 VGPRs (2-6 spilled; 99 unconstrained, 4 waves), which the
// short-row lattices need (one 128-lane workgroup per row): the sphere
// channel's FP32 odd step with statistics took 131 us at 6 waves (79 VGPRs, no
// spills) and takes 126 us at 7. Every other instantiation is unconstrained
// (1 = the default).
//
// CACHEDZ: the z-moving outputs (staged rows, halo elements, row-end
// reflections) are written with ordinary stores instead of nontemporal ones.
// Chosen for lattices small enough that those lines can stay in the device's
// last-level cache until the next step reads them (the lines that two
// workgroups write in parts also merge there).
template<typename S, int DYN, bool TAGS, bool STATS, bool FORCED, bool MOMENTS = false, bool CACHEDZ = false>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(
    (sizeof(S) == 4 && STATS && !FORCED) ? 7 : 1))) bulk_odd_staged(StepArgs a) {
    using T = typename Store<S>::T;
    using R = typename Store<S>::R;
    // LDS words of at least 32 bits: 16-bit values would put two lanes on one bank word
    using L = typename std::conditional<(sizeof(R) < 4), uint32_t, R>::type;
    extern __shared__ __align__(16) unsigned char ldsRaw[];
    const int W = (int)blockDim.x;            // segment length: one cell per thread
    const int P = W + 2;                      // pitch: low halo, segment, high halo
    L* buf = reinterpret_cast<L*>(ldsRaw);
    // TAGS only: per cell of the segment, the directions whose output it does not
    // send along c_i (its link-wall mask; all bits when it is not collided here)
    uint32_t* blocked = reinterpret_cast<uint32_t*>(ldsRaw + (size_t)kZMoving * P * sizeof(L));

    const uint32_t wg = a.xcdRemap ? xcd_item(blockIdx.x, gridDim.x) : blockIdx.x;
    const int b = block_of(a, wg);
    const BlockGeom g = a.blocks[b];
    S* __restrict__ f = reinterpret_cast<S*>(a.arena) + g.offset;
    const uint32_t local = wg - g.wgBegin;
    const int chunk = (int)(local % (uint32_t)g.nzChunks);
    const uint32_t rest = local / (uint32_t)g.nzChunks;
    const int y = (int)(rest % (uint32_t)g.nyGroups);
    const int x = (int)(rest / (uint32_t)g.nyGroups);
    const int z0 = chunk * W;
    const int len = min(g.nz - z0, W);
    const int t = (int)threadIdx.x;
    const int z = z0 + t;
    const bool active = t < len;
    const uint64_t rowBase = ((uint64_t)x * g.ny + y) * g.nz;
    const uint64_t cell = rowBase + z;
    double st[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    double wf[3] = {0, 0, 0};

    int xm = 0, xp = 0, ym = 0, yp = 0;
    const bool vxm = shift(x, -1, g.nx, g.px, xm), vxp = shift(x, 1, g.nx, g.px, xp);
    const bool vym = shift(y, -1, g.ny, g.py, ym), vyp = shift(y, 1, g.ny, g.py, yp);
    const int xs[3] = {xm, x, xp};
    const int ys[3] = {ym, y, yp};
    const bool vx[3] = {vxm, true, vxp};
    const bool vy[3] = {vym, true, vyp};
    const bool rowsValid = vxm && vxp && vym && vyp;
    const bool inBulk = x >= g.bx0 && x <= g.bx1 && y >= g.by0 && y <= g.by1 && z >= g.bz0 && z <= g.bz1;

    uint8_t tag = BT_BULK;
    uint32_t wall = 0;
    if constexpr (TAGS) {
        // the link mask is zero except on wall-near cells: loaded without testing
        // the tag, so that no load waits for another before the population loads
        if (active) tag = a.tags[g.cellOffset + cell];
        if (active && a.walls) wall = a.walls[g.cellOffset + cell];
    }
    const bool compute = active && tag != BT_SKIP;
    T fc[19];

    if (!rowsValid) {
        // general path: per-lane neighbour validity in x, y and z. Rows in the
        // outermost envelope layer of a block are skipped: nothing they would write
        // reaches the bulk, and the envelope exchange after this step replaces them.
        if (compute && !outer_envelope_row(g, x, y))
            odd_cell_general<S, DYN, STATS, FORCED, MOMENTS>(a, g, f, z, cell, tag, inBulk, xs, ys, vx, vy, st, wall, wf);
    } else {
        S* rows[3][3];
#pragma unroll
        for (int ax = 0; ax < 3; ++ax)
#pragma unroll
            for (int ay = 0; ay < 3; ++ay)
                rows[ax][ay] = f + ((uint64_t)xs[ax] * g.ny + ys[ay]) * g.nz;
        int zlo = 0, zhi = 0;
        const bool vlo = shift(z0, -1, g.nz, g.pz, zlo);
        const bool vhi = shift(z0 + len - 1, 1, g.nz, g.pz, zhi);
        // A whole periodic row in this segment: the elements beyond its ends are
        // its other end's, already in LDS. LDS position 1 + j holds cell j, so the
        // positions 0 and len + 1 wrap to len and 1 (no halo loads or stores).
        const bool wrapRow = g.pz && len == g.nz;
        auto wrapPos = [&](int s) { return !wrapRow ? s : (s == 0 ? len : (s == len + 1 ? 1 : s)); };

        // 1. loads, all issued before any is consumed: the 9 aligned directions
        // to registers, the 10 z-moving source-row elements to registers and then
        // to LDS. Inactive lanes (past the row end) reload the segment's first
        // cell and keep their results to themselves.
        const int zc = active ? z : z0;
        R raw[kZMoving];
#pragma unroll
        for (int i = 0; i < 19; ++i) {
            const S* src = rows[1 - cx(i)][1 - cy(i)] + (uint64_t)opp(i) * g.slot;
            if (cz(i) == 0) fc[i] = Store<S>::load(src + zc);
            else raw[zmove_index(i)] = Store<S>::load_raw(src + zc);
        }
        if constexpr (TAGS) {
            // a link wall at x - c_i: f_i comes back from the cell's own slot i
            // (after every unconditional load is issued)
            if (wall) {
#pragma unroll
                for (int i = 0; i < 19; ++i)
                    if (cz(i) == 0 && ((wall >> opp(i)) & 1u)) fc[i] = Store<S>::load(f + (uint64_t)i * g.slot + cell);
            }
        }
        if constexpr (TAGS) {
            // z-moving directions with a link wall at x - c_i: the cell's own slot
            // i, issued before the barrier so that its latency overlaps it
            if (compute && wall) {
#pragma unroll
                for (int i = 0; i < 19; ++i)
                    if (cz(i) != 0 && ((wall >> opp(i)) & 1u)) fc[i] = Store<S>::load(f + (uint64_t)i * g.slot + cell);
            }
        }
        // Non-periodic row ends: the edge cell's five inputs from beyond the end
        // are its own slots i (reflection). Lanes 0-4 load the low end's, lanes
        // 5-9 the high end's: one load instruction for the ten, into the halo
        // positions that the edge cells read after the barrier. (One lane loading
        // all five per end measured slower than the edge cell reading its own
        // slots: 14.2 against 14.6 GLUPS on the FP64 sphere channel.)
        const bool loEdge = !vlo && t < 5, hiEdge = !vhi && t >= 5 && t < 10;
        int eDir = 0, eK = 0;
        R eVal = R();
        if (loEdge || hiEdge) {
            int dHi = 0, kHi = 0;
            zdir<1>(t, eDir, eK);
            zdir<-1>(t - 5, dHi, kHi);
            if (hiEdge) { eDir = dHi; eK = kHi; }
            eVal = Store<S>::load_raw(f + (uint64_t)eDir * g.slot + rowBase + (loEdge ? z0 : z0 + len - 1));
        }
        // Halo elements inside a periodic or segmented row: the source elements
        // beyond the segment ends.
        R halo[kZMoving];
        const bool loLane = t == 0 && vlo && !wrapRow, hiLane = t == W - 1 && vhi && !wrapRow;
        if (loLane || hiLane) {
            const int zh = loLane ? zlo : zhi;
#pragma unroll
            for (int i = 0; i < 19; ++i)
                if (cz(i) != 0)
                    halo[zmove_index(i)] = Store<S>::load_raw(rows[1 - cx(i)][1 - cy(i)] + (uint64_t)opp(i) * g.slot + zh);
        }
        if (active) {
#pragma unroll
            for (int k = 0; k < kZMoving; ++k) buf[k * P + t + 1] = raw[k];
        }
        if (loLane || hiLane) {
            const int hp = loLane ? 0 : len + 1;
#pragma unroll
            for (int k = 0; k < kZMoving; ++k) buf[k * P + hp] = halo[k];
        }
        if (loEdge || hiEdge) buf[eK * P + (loEdge ? 0 : len + 1)] = eVal;
        if constexpr (TAGS) {
            if (active) blocked[t] = compute ? wall : 0xffffffffu;
        }
        __syncthreads();
        if (compute) {
#pragma unroll
            for (int i = 0; i < 19; ++i) {
                if (cz(i) != 0 && !((wall >> opp(i)) & 1u))
                    fc[i] = Store<S>::from_raw((R)buf[zmove_index(i) * P + wrapPos(t + 1 - cz(i))]);
            }
        }
        __syncthreads();   // every LDS input is read before the outputs reuse the buffer

        // 2. collision, aligned outputs to global, z-moving outputs to LDS
        if (compute) {
            collide_cell<S, DYN, STATS, FORCED, MOMENTS>(a, g, cell, tag, inBulk, fc, st, wall, wf);
#pragma unroll
            for (int i = 0; i < 19; ++i) {
                const bool reflect = (wall >> i) & 1u;   // link wall at x + c_i
                if (cz(i) == 0) {
                    S* p = reflect ? f + (uint64_t)opp(i) * g.slot + cell
                                   : rows[1 + cx(i)][1 + cy(i)] + (uint64_t)i * g.slot + z;
                    Store<S>::store(p, fc[i]);
                } else {
                    const int k = zmove_index(i);
                    int zd = 0;
                    if (!shift(z, cz(i), g.nz, g.pz, zd))
                        buf[k * P + (cz(i) < 0 ? len + 1 : 0)] = Store<S>::to_raw(fc[i]);   // row end: for the edge lanes
                    else if (!reflect)
                        buf[k * P + wrapPos(t + 1 + cz(i))] = Store<S>::to_raw(fc[i]);
                    else
                        Store<S>::store(f + (uint64_t)opp(i) * g.slot + cell, fc[i]);   // link-wall reflection
                }
            }
        }
        __syncthreads();

        // 3. aligned stores of the z-moving destination rows; position q takes the
        // output of cell q - c_z when that cell is in this segment and was computed
        auto storeZ = [](S* p, R v) {
            if constexpr (CACHEDZ) Store<S>::store_raw_cached(p, v);
            else Store<S>::store_raw(p, v);
        };
#pragma unroll
        for (int k = 0; k < kZMoving; ++k) raw[k] = (R)buf[k * P + t + 1];
        if (loLane || hiLane) {
            const int hp = loLane ? 0 : len + 1;
#pragma unroll
            for (int k = 0; k < kZMoving; ++k) halo[k] = (R)buf[k * P + hp];
        }
#pragma unroll
        for (int i = 0; i < 19; ++i) {
            if (cz(i) != 0) {
                const int k = zmove_index(i);
                S* dst = rows[1 + cx(i)][1 + cy(i)] + (uint64_t)i * g.slot;
                int p = t - cz(i);
                if (wrapRow) p = p < 0 ? p + len : (p >= len ? p - len : p);
                bool ok = active && p >= 0 && p < len;
                if constexpr (TAGS) ok = ok && !((blocked[p < 0 ? 0 : (p >= len ? 0 : p)] >> i) & 1u);
                if (ok) storeZ(dst + z, raw[k]);
                // halo position: the low one takes cell 0's output for c_z < 0,
                // the high one cell len-1's output for c_z > 0
                bool h = cz(i) < 0 ? loLane : hiLane;
                if constexpr (TAGS) h = h && !((blocked[cz(i) < 0 ? 0 : len - 1] >> i) & 1u);
                if (h) storeZ(dst + (cz(i) < 0 ? zlo : zhi), halo[k]);
            }
        }
        // Row-end reflections, one store instruction: lanes 0-4 write the low edge
        // cell's outputs along c_z = -1 (staged at position len + 1), lanes 5-9 the
        // high edge cell's along c_z = +1 (position 0), each into the cell's own
        // slot opp(i), when the edge cell was computed here.
        if (loEdge || hiEdge) {
            int dOut = 0, kOut = 0, dHi = 0, kHi = 0;
            zdir<-1>(t, dOut, kOut);
            zdir<1>(t - 5, dHi, kHi);
            if (hiEdge) { dOut = dHi; kOut = kHi; }
            const int ce = loEdge ? 0 : len - 1;
            bool done = true;
            if constexpr (TAGS) done = blocked[ce] != 0xffffffffu;
            if (done) storeZ(f + (uint64_t)opp(dOut) * g.slot + rowBase + z0 + ce, (R)buf[kOut * P + (loEdge ? len + 1 : 0)]);
        }
    }
    store_bulk_stats<STATS, FORCED>(st, wf, a.partials, wg, a.walls);
}

// ---------------------------------------------------------------------------
// 16-bit storage, two cells per thread. A thread owns the cell pair (z, z+1),
// z even, and moves both 16-bit values of a slot with one aligned 32-bit
// access: half the memory instructions of one cell per thread. Even steps are
// purely local. Odd steps stage the z-moving source and destination rows as
// 32-bit pair words in LDS (the neighbouring pair supplies the shifted half),
// so every global access stays a 32-bit aligned pair, except single halves at
// a segment's ends and where one cell of a pair belongs to the special-cell
// kernel. Requires nz even for every block. Rows at a non-periodic x/y edge
// take the per-cell general path.

__device__ __forceinline__ uint32_t ld_pair(const void* p) {
    return __builtin_nontemporal_load(reinterpret_cast<const uint32_t*>(p));
}
__device__ __forceinline__ void st_pair(void* p, uint32_t v) {
    __builtin_nontemporal_store(v, reinterpret_cast<uint32_t*>(p));
}
__device__ __forceinline__ uint16_t lo16(uint32_t v) { return (uint16_t)(v & 0xffffu); }
__device__ __forceinline__ uint16_t hi16(uint32_t v) { return (uint16_t)(v >> 16); }
__device__ __forceinline__ uint32_t pack16(uint16_t lo, uint16_t hi) { return (uint32_t)lo | ((uint32_t)hi << 16); }

// PACKED: a workgroup takes g.rowsPerWG whole rows of a block (rows shorter than
// half the workgroup: the lanes a one-row segment would leave idle take the next
// rows; every row holds at least two pairs, so that its low and high halo words
// come from different lanes). Without it a workgroup takes one segment of up to
// 2W cells of one row and the row is wavefront-uniform. ZWRAP: every block is
// periodic along z (no z edge: the edge handling is compiled out, which keeps
// the kernel within 80 registers).
// CACHEDZ (odd steps): ordinary stores for the z-moving rows, the single halves
// and the reflections, as bulk_odd_staged's CACHEDZ.
template<typename S, int DYN, bool ODD, bool TAGS, bool STATS, bool FORCED, bool MOMENTS = false, bool PACKED = false,
         bool ZWRAP = false, bool CACHEDZ = false>
__global__ void __launch_bounds__(512) bulk_pairs(StepArgs a) {
    static_assert(sizeof(typename Store<S>::R) == 2, "pairs of 16-bit values");
    using T = typename Store<S>::T;
    extern __shared__ __align__(16) unsigned char ldsRaw[];
    const int W = (int)blockDim.x;            // threads per workgroup
    // pitch in pair words per staged direction: per segment a low halo word, the
    // segment, a high halo word
    const int P = PACKED ? a.ldsPitch : W + 2;
    uint32_t* buf = reinterpret_cast<uint32_t*>(ldsRaw);
    // ODD && TAGS: per cell, the directions whose output it does not send along c_i
    // (its link-wall mask; all bits when it is not collided here), 2W words
    uint32_t* blocked = reinterpret_cast<uint32_t*>(ldsRaw + (size_t)kZMoving * P * 4);

    const uint32_t wg = a.xcdRemap ? xcd_item(blockIdx.x, gridDim.x) : blockIdx.x;
    const int b = block_of(a, wg);
    const BlockGeom g = a.blocks[b];
    S* __restrict__ f = reinterpret_cast<S*>(a.arena) + g.offset;
    const uint32_t local = wg - g.wgBegin;
    const int t = (int)threadIdx.x;
    int x, y, z0, len, p, lanes, sb;
    bool rowOk;
    if constexpr (PACKED) {
        // segment = whole row; lanes t = r * units + p, r the row within the workgroup
        const int units = g.nz / 2;
        const int r = t / units;
        p = t - r * units;
        const uint32_t rowIdx = local * (uint32_t)g.rowsPerWG + (uint32_t)r;
        rowOk = r < g.rowsPerWG && rowIdx < (uint32_t)g.nx * (uint32_t)g.ny;
        x = rowOk ? (int)(rowIdx / (uint32_t)g.ny) : 0;
        y = rowOk ? (int)(rowIdx % (uint32_t)g.ny) : 0;
        z0 = 0;
        len = g.nz;
        lanes = units;
        sb = r * (units + 2);
    } else {
        const int chunk = (int)(local % (uint32_t)g.nzChunks);
        const uint32_t rest = local / (uint32_t)g.nzChunks;
        y = (int)(rest % (uint32_t)g.nyGroups);
        x = (int)(rest / (uint32_t)g.nyGroups);
        z0 = chunk * 2 * W;
        len = min(g.nz - z0, 2 * W);          // even
        p = t;
        lanes = W;
        sb = 0;
        rowOk = true;
    }
    const int nPairs = len / 2;
    const int cb = 2 * (t - p);               // first entry of the segment in `blocked`
    const bool active = rowOk && p < nPairs;
    const int zA = z0 + 2 * p, zB = zA + 1;
    const uint64_t rowBase = ((uint64_t)x * g.ny + y) * g.nz;
    const uint64_t cellA = rowBase + zA, cellB = cellA + 1;
    const bool rowInBulk = x >= g.bx0 && x <= g.bx1 && y >= g.by0 && y <= g.by1;
    const bool inBulkA = rowInBulk && zA >= g.bz0 && zA <= g.bz1;
    const bool inBulkB = rowInBulk && zB >= g.bz0 && zB <= g.bz1;
    double st[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    double wf[3] = {0, 0, 0};

    uint8_t tagA = BT_BULK, tagB = BT_BULK;
    uint32_t wallA = 0, wallB = 0;
    if constexpr (TAGS) {
        if (active) {
            // one 16-bit tag load and one 64-bit link-mask load for the pair (cellA is
            // even in every block, and so is every block's cell offset). The mask is
            // zero except on wall-near cells, so it is loaded without testing the tag:
            // no load waits for another before the population loads are issued.
            const uint16_t tg = *reinterpret_cast<const uint16_t*>(a.tags + g.cellOffset + cellA);
            tagA = (uint8_t)(tg & 0xffu);
            tagB = (uint8_t)(tg >> 8);
            if (a.walls) {
                const uint64_t w = *reinterpret_cast<const uint64_t*>(a.walls + g.cellOffset + cellA);
                wallA = (uint32_t)w;
                wallB = (uint32_t)(w >> 32);
            }
        }
    }
    const bool compA = active && tagA != BT_SKIP, compB = active && tagB != BT_SKIP;
    T fa[19], fb[19];

    if constexpr (!ODD) {
        // even step: each cell reads its own slots and writes slot opp(i)
        S* own = f + cellA;
        if (active) {
#pragma unroll
            for (int i = 0; i < 19; ++i) {
                const uint32_t v = ld_pair(own + (uint64_t)i * g.slot);
                fa[i] = Store<S>::from_raw(lo16(v));
                fb[i] = Store<S>::from_raw(hi16(v));
            }
        }
        // tagged lattices: the records of the pair stored together (collide_cell
        // records one cell; measured faster only with tags)
        constexpr bool PAIR_REC = MOMENTS && TAGS;
        const bool recA = PAIR_REC && compA && (tagA == BT_BULK || tagA == BT_WALL_NEAR);
        const bool recB = PAIR_REC && compB && (tagB == BT_BULK || tagB == BT_WALL_NEAR);
        float ra[4] = {0, 0, 0, 0}, rb[4] = {0, 0, 0, 0};
        if (recA) moments_of(fa, ra);
        if (compA) collide_cell<S, DYN, STATS, FORCED, MOMENTS && !PAIR_REC>(a, g, cellA, tagA, inBulkA, fa, st, wallA, wf);
        if (recB) moments_of(fb, rb);
        if (compB) collide_cell<S, DYN, STATS, FORCED, MOMENTS && !PAIR_REC>(a, g, cellB, tagB, inBulkB, fb, st, wallB, wf);
        if constexpr (PAIR_REC)
            store_moments_pair(ra, rb, recA, recB, reinterpret_cast<float*>(a.moments), a.momStride,
                               g.cellOffset + cellA);
        // whole pair words, also for pairs of two cells not collided here (no holes in
        // the lines written): such a cell gets back the value of slot opp(i) it was
        // loaded with (only this thread writes the pair in an even step)
        if (active) {
#pragma unroll
            for (int i = 0; i < 19; ++i)
                st_pair(own + (uint64_t)opp(i) * g.slot, pack16(Store<S>::to_raw(compA ? fa[i] : fa[opp(i)]),
                                                                Store<S>::to_raw(compB ? fb[i] : fb[opp(i)])));
        }
    } else {
        int xm = 0, xp = 0, ym = 0, yp = 0;
        const bool vxm = shift(x, -1, g.nx, g.px, xm), vxp = shift(x, 1, g.nx, g.px, xp);
        const bool vym = shift(y, -1, g.ny, g.py, ym), vyp = shift(y, 1, g.ny, g.py, yp);
        const int xs[3] = {xm, x, xp};
        const int ys[3] = {ym, y, yp};
        const bool vx[3] = {vxm, true, vxp};
        const bool vy[3] = {vym, true, vyp};
        const bool allValid = vxm && vxp && vym && vyp;
        // rows on a block edge: per-lane edge selection, no LDS (the outermost
        // envelope layer is not collided: see bulk_odd_staged)
        if (rowOk && !allValid && !outer_envelope_row(g, x, y)) {
                if (compA)
                odd_cell_general<S, DYN, STATS, FORCED, MOMENTS>(a, g, f, zA, cellA, tagA, inBulkA, xs, ys, vx, vy, st,
                                                                 wallA, wf);
            if (compB)
                odd_cell_general<S, DYN, STATS, FORCED, MOMENTS>(a, g, f, zB, cellB, tagB, inBulkB, xs, ys, vx, vy, st,
                                                                 wallB, wf);
        }
        // staged rows. Unpacked, the row is workgroup-uniform; packed, every thread
        // takes part in the barriers and the edge rows' lanes are masked off (their
        // segments in LDS are used by no other row).
        if (PACKED || allValid) {
            const bool sAct = active && allValid;
            const bool sCompA = compA && allValid, sCompB = compB && allValid;
            S* rows[3][3];
#pragma unroll
            for (int ax = 0; ax < 3; ++ax)
#pragma unroll
                for (int ay = 0; ay < 3; ++ay)
                    rows[ax][ay] = f + ((uint64_t)xs[ax] * g.ny + ys[ay]) * g.nz;
            // halo pair words: (z0-2, z0-1) and (z0+len, z0+len+1)
            int zlo = 0, zhi = 0;
            const bool vlo = shift(z0, -2, g.nz, g.pz, zlo);
            const bool vhi = shift(z0 + len - 1, 1, g.nz, g.pz, zhi);
            // A whole periodic row in this segment: the pair words beyond its ends
            // are its other end's, already in LDS (pair j at sb + 1 + j): read and
            // stored from there, with no halo loads and no single-half stores.
            const bool wrapRow = g.pz && len == g.nz;
            const int qLo = (wrapRow && p == 0) ? sb + nPairs : sb + p;          // LDS word of pair p - 1
            const int qHi = (wrapRow && p == nPairs - 1) ? sb + 1 : sb + p + 2;  // LDS word of pair p + 1
            const int zc = active ? zA : z0;
            uint32_t raw[kZMoving];
            if (sAct) {
#pragma unroll
                for (int i = 0; i < 19; ++i) {
                    const S* src = rows[1 - cx(i)][1 - cy(i)] + (uint64_t)opp(i) * g.slot;
                    if (cz(i) == 0) {
                        const uint32_t v = ld_pair(src + zc);
                        fa[i] = Store<S>::from_raw(lo16(v));
                        fb[i] = Store<S>::from_raw(hi16(v));
                    } else {
                        raw[zmove_index(i)] = ld_pair(src + zc);
                    }
                }
            }
            // TAGS: wavefronts with a link wall take the wall loads (most take none)
            bool waveWalls = false;
            if constexpr (TAGS) waveWalls = __ballot(sAct && (wallA | wallB) != 0u) != 0;
            if (waveWalls) {
                // link walls at x - c_i: f_i comes back from the cell's own slot i
                // (after every unconditional load is issued)
                // (one pair load for both cells)
                if (sAct && (wallA | wallB)) {
#pragma unroll
                    for (int i = 0; i < 19; ++i) {
                        if (cz(i) != 0) continue;
                        if (((wallA | wallB) >> opp(i)) & 1u) {
                            const uint32_t v = ld_pair(f + (uint64_t)i * g.slot + cellA);
                            if ((wallA >> opp(i)) & 1u) fa[i] = Store<S>::from_raw(lo16(v));
                            if ((wallB >> opp(i)) & 1u) fb[i] = Store<S>::from_raw(hi16(v));
                        }
                    }
                }
            }
            // TAGS: the own slot-i pair where a link wall reflects a z-moving
            // direction, loaded with the other loads (not after the barrier)
            uint32_t wref[kZMoving];
#pragma unroll
            for (int k = 0; k < kZMoving; ++k) wref[k] = 0;
            if (waveWalls) {
#pragma unroll
                for (int i = 0; i < 19; ++i) {
                    if (cz(i) == 0) continue;
                    if (sAct && (((wallA | wallB) >> opp(i)) & 1u))
                        wref[zmove_index(i)] = ld_pair(f + (uint64_t)i * g.slot + cellA);
                }
            }
            // Halo words. Inside a periodic or segmented row: the source pairs beyond
            // the segment ends. At a non-periodic row end: the end cell's own slot-i
            // value in the half its shifted read takes (the high half of the low word
            // for c_z = +1, the low half of the high word for c_z = -1), so that the
            // edge reflection is read from LDS like any other input.
            uint32_t halo[kZMoving];
            const bool segOk = rowOk && allValid;
            const bool loLane = segOk && p == 0 && vlo && !wrapRow, hiLane = segOk && p == lanes - 1 && vhi && !wrapRow;
            const bool loEdge = !ZWRAP && segOk && p == 0 && !vlo, hiEdge = !ZWRAP && segOk && p == lanes - 1 && !vhi;
            const bool edge = loEdge || hiEdge;
            if (loLane || hiLane || edge) {
                // edge lanes load every z-moving direction's end pair; the shifted
                // read takes only the directions arriving from beyond the end
                const int zh = edge ? (loEdge ? z0 : z0 + len - 2) : (loLane ? zlo : zhi);
#pragma unroll
                for (int i = 0; i < 19; ++i)
                    if (cz(i) != 0) {
                        const S* src = edge ? f + rowBase + (uint64_t)i * g.slot
                                            : rows[1 - cx(i)][1 - cy(i)] + (uint64_t)opp(i) * g.slot;
                        const uint32_t v = ld_pair(src + zh);
                        halo[zmove_index(i)] = !edge ? v : (loEdge ? (v << 16) : (v >> 16));
                    }
            }
            const int q = sb + p;                 // this lane's pair word in its segment, minus 1
            if (sAct) {
#pragma unroll
                for (int k = 0; k < kZMoving; ++k) buf[k * P + q + 1] = raw[k];
            }
            if (loLane || hiLane || edge) {
                const int hp = sb + ((loLane || loEdge) ? 0 : nPairs + 1);
#pragma unroll
                for (int k = 0; k < kZMoving; ++k) buf[k * P + hp] = halo[k];
            }
            if constexpr (TAGS) {
                if (sAct) {
                    blocked[2 * t] = compA ? wallA : 0xffffffffu;
                    blocked[2 * t + 1] = compB ? wallB : 0xffffffffu;
                }
            }
            __syncthreads();
            // shifted inputs: c_z = +1 reads z - 1, c_z = -1 reads z + 1
            const bool vzA = ZWRAP || zA > 0 || g.pz, vzB = ZWRAP || zB < g.nz - 1 || g.pz;
            if (sAct) {
#pragma unroll
                for (int i = 0; i < 19; ++i) {
                    if (cz(i) == 0) continue;
                    const uint32_t* row = buf + zmove_index(i) * P;
                    const bool wA = (wallA >> opp(i)) & 1u, wB = (wallB >> opp(i)) & 1u;
                    uint16_t ra, rb;
                    if (cz(i) > 0) {
                        ra = hi16(row[qLo]);
                        rb = lo16(row[q + 1]);
                    } else {
                        ra = hi16(row[q + 1]);
                        rb = lo16(row[qHi]);
                    }
                    if constexpr (TAGS) {
                        // link wall at x - c_i: f_i comes back from the cell's own slot i
                        if (wA) ra = lo16(wref[zmove_index(i)]);
                        if (wB) rb = hi16(wref[zmove_index(i)]);
                    }
                    fa[i] = Store<S>::from_raw(ra);
                    fb[i] = Store<S>::from_raw(rb);
                }
            }
            // A plain wavefront: every lane's two cells collided here with no link
            // wall, and so the cells next to its pair that its stores take outputs
            // from (segment ends excepted: those are not this segment's). Its stores
            // below are those of a lattice without tags; the others test every
            // direction. (With tags, the tag and wall tests cost a third of the odd
            // step even where no cell is near a wall.)
            bool plainWave = true;
            if constexpr (TAGS) {
                bool lanePlain = true;
                if (sAct)
                    lanePlain = compA && compB && (wallA | wallB) == 0u && (p == 0 || blocked[cb + 2 * p - 1] == 0u) &&
                                (2 * p + 2 >= len || blocked[cb + 2 * p + 2] == 0u);
                plainWave = __ballot(lanePlain) == __ballot(1);
            }
            __syncthreads();   // every LDS input is read before the outputs reuse the buffer
            // MOMENTS with tags: the records of the pair stored together (collide_cell
            // records one cell; measured faster only with tags)
            constexpr bool PAIR_REC = MOMENTS && TAGS;
            const bool recA = PAIR_REC && sCompA && (tagA == BT_BULK || tagA == BT_WALL_NEAR);
            const bool recB = PAIR_REC && sCompB && (tagB == BT_BULK || tagB == BT_WALL_NEAR);
            float ra[4] = {0, 0, 0, 0}, rb[4] = {0, 0, 0, 0};
            if (recA) moments_of(fa, ra);
            if (sCompA)
                collide_cell<S, DYN, STATS, FORCED, MOMENTS && !PAIR_REC>(a, g, cellA, tagA, inBulkA, fa, st, wallA, wf);
            if (recB) moments_of(fb, rb);
            if (sCompB)
                collide_cell<S, DYN, STATS, FORCED, MOMENTS && !PAIR_REC>(a, g, cellB, tagB, inBulkB, fb, st, wallB, wf);
            if constexpr (PAIR_REC)
                store_moments_pair(ra, rb, recA, recB, reinterpret_cast<float*>(a.moments), a.momStride,
                                   g.cellOffset + cellA);
            // outputs: aligned directions to HBM, z-moving ones to LDS. PL: a plain
            // wavefront (or no tags): every active cell collided here, no walls.
            // reflections and z-moving rows: nontemporal, or ordinary (CACHEDZ)
            auto storeZ = [](S* p, T v) {
                if constexpr (CACHEDZ) Store<S>::store_raw_cached(p, Store<S>::to_raw(v));
                else Store<S>::store(p, v);
            };
            auto storeRawZ = [](S* p, uint16_t v) {
                if constexpr (CACHEDZ) Store<S>::store_raw_cached(p, v);
                else Store<S>::store_raw(p, v);
            };
            auto storePairZ = [](S* p, uint32_t v) {
                if constexpr (CACHEDZ) *reinterpret_cast<uint32_t*>(p) = v;
                else st_pair(p, v);
            };
            auto emitOutputs = [&](auto plainTag) {
                constexpr bool PL = decltype(plainTag)::value;
                if (!sAct) return;
#pragma unroll
                for (int i = 0; i < 19; ++i) {
                    // link walls at x + c_i: the output goes back to the cell's own slot opp(i)
                    const bool rA = !PL && compA && ((wallA >> i) & 1u), rB = !PL && compB && ((wallB >> i) & 1u);
                    const bool sA = PL || (compA && !rA), sB = PL || (compB && !rB);
                    if (rA) storeZ(f + (uint64_t)opp(i) * g.slot + cellA, fa[i]);
                    if (rB) storeZ(f + (uint64_t)opp(i) * g.slot + cellB, fb[i]);
                    if (cz(i) == 0) {
                        S* dst = rows[1 + cx(i)][1 + cy(i)] + (uint64_t)i * g.slot + zA;
                        if (sA && sB) st_pair(dst, pack16(Store<S>::to_raw(fa[i]), Store<S>::to_raw(fb[i])));
                        else {
                            if (sA) Store<S>::store(dst, fa[i]);
                            if (sB) Store<S>::store(dst + 1, fb[i]);
                        }
                    } else {
                        buf[zmove_index(i) * P + q + 1] = pack16(Store<S>::to_raw(fa[i]), Store<S>::to_raw(fb[i]));
                        // edge reflection where the destination leaves a non-periodic z edge
                        if (cz(i) > 0 && !vzB && sB) storeZ(f + (uint64_t)opp(i) * g.slot + cellB, fb[i]);
                        if (cz(i) < 0 && !vzA && sA) storeZ(f + (uint64_t)opp(i) * g.slot + cellA, fa[i]);
                    }
                }
            };
            if (plainWave) emitOutputs(std::true_type{});
            else emitOutputs(std::integral_constant<bool, !TAGS>{});
            __syncthreads();
            // destination pair words: position (zA, zB) takes, for c_z = +1, the
            // outputs of cells zA - 1 and zA; for c_z = -1, of cells zB and zB + 1. Half
            // words are stored where a cell sent no output there: beyond a non-periodic
            // z edge (that half is another row's edge reflection) and at the ends of a
            // row split into segments (the halves beyond a segment end are the
            // neighbouring segment's).
            const bool segmented = len < g.nz;
            auto emitRows = [&](auto plainTag) {
                constexpr bool PL = decltype(plainTag)::value;
                // segment-local cell c was collided here and sent its direction-i output
                // along c_i; `near`: c is this lane's or next to its pair, which a plain
                // wavefront has checked
                auto computed = [&](int c, int i, bool near) {
                    if (c < 0 || c >= len) return false;
                    if constexpr (TAGS) {
                        if (!(PL && near)) return ((blocked[cb + c] >> i) & 1u) == 0;
                    }
                    return true;
                };
                // the same as words, once per lane: bit i set where the cell sent its
                // direction-i output along c_i
                auto sent = [&](int c, bool near) -> uint32_t {
                    if (c < 0 || c >= len) return 0u;
                    if constexpr (TAGS) {
                        if (!(PL && near)) return ~blocked[cb + c];
                    }
                    return ~0u;
                };
                const uint32_t sentA = sent(2 * p, true), sentB = sent(2 * p + 1, true);
                const uint32_t sentLo = (wrapRow && p == 0) ? sent(len - 1, false) : sent(2 * p - 1, true);
                const uint32_t sentHi = (wrapRow && p == nPairs - 1) ? sent(0, false) : sent(2 * p + 2, true);
#pragma unroll
                for (int i = 0; i < 19; ++i) {
                    if (cz(i) == 0) continue;
                    const uint32_t* row = buf + zmove_index(i) * P;
                    S* dst = rows[1 + cx(i)][1 + cy(i)] + (uint64_t)i * g.slot;
                    if (sAct) {
                        uint16_t lo, hi;
                        bool okLo, okHi;
                        if (cz(i) > 0) {
                            lo = hi16(row[qLo]);
                            okLo = (sentLo >> i) & 1u;
                            hi = lo16(row[q + 1]);
                            okHi = (sentA >> i) & 1u;
                        } else {
                            lo = hi16(row[q + 1]);
                            okLo = (sentB >> i) & 1u;
                            hi = lo16(row[qHi]);
                            okHi = (sentHi >> i) & 1u;
                        }
                        if (okLo && okHi) storePairZ(dst + zA, pack16(lo, hi));
                        else if (okLo || okHi) storeRawZ(dst + (okLo ? zA : zB), okLo ? lo : hi);
                    }
                }
                // single halves beyond the segment: c_z = +1 sends the last cell's
                // output to z0 + len; c_z = -1 sends the first cell's to z0 - 1 (the
                // end cell may be another wavefront's: always tested)
                if (!segmented) return;
                int zdHi = 0, zdLo = 0;
                const bool beyondHi = segOk && p == lanes - 1 && shift(z0 + len - 1, 1, g.nz, g.pz, zdHi);
                const bool beyondLo = segOk && p == 0 && shift(z0, -1, g.nz, g.pz, zdLo);
                if (!beyondHi && !beyondLo) return;
#pragma unroll
                for (int i = 0; i < 19; ++i) {
                    if (cz(i) == 0) continue;
                    const uint32_t* row = buf + zmove_index(i) * P;
                    S* dst = rows[1 + cx(i)][1 + cy(i)] + (uint64_t)i * g.slot;
                    if (cz(i) > 0) {
                        if (beyondHi && computed(len - 1, i, false)) storeRawZ(dst + zdHi, hi16(row[sb + nPairs]));
                    } else {
                        if (beyondLo && computed(0, i, false)) storeRawZ(dst + zdLo, lo16(row[sb + 1]));
                    }
                }
            };
            if (plainWave) emitRows(std::true_type{});
            else emitRows(std::integral_constant<bool, !TAGS>{});
        }
    }
    store_bulk_stats<STATS, FORCED>(st, wf, a.partials, wg, a.walls);
}

// ---------------------------------------------------------------------------
// Special-cell kernel: boundary planes, momentum-exchange nodes, the cells of
// external velocity edges and corners. Same AA addressing as the bulk kernel,
// with the wall mask added to the edge test.
//
// SK_STORE_VELOCITY: StoreVelocityDynamics over the background dynamics
// (boundaryCondition/boundaryDynamics.hh:171-209), the dynamics of the edge and
// corner cells of an external velocity condition: collides as the background
// dynamics, reports its stored velocity; the finite-difference edge and corner
// processors (fd_boundary.h) rewrite its populations after every stream.

enum SpecialKind : uint32_t { SK_PLANE = 0, SK_MOMENTUM_EXCHANGE = 1, SK_STORE_VELOCITY = 2 };

struct SpecialCell {
    uint32_t cell;        // block-local linear cell index
    uint16_t x, y, z;     // block-local coordinates
    uint16_t info;        // kind | family << 2 | bckind << 3 | dir << 4 | (ori > 0) << 6 | wrapped << 7
                          //      | bulk << 8 (the cell takes statistics)
    uint32_t value;       // index into the plane value array (velocity: 3 doubles, pressure: 1)
    uint32_t wall;        // bit i: neighbour x + c_i is a link-wise wall
    uint32_t fluid;       // bit i: fluid direction of a momentum-exchange node
    uint16_t block;       // block index
    uint16_t pad0;
    uint32_t pad1;
};
static_assert(sizeof(SpecialCell) == 32, "special record is 32 bytes");

struct SpecialArgs {
    void* arena;
    const BlockGeom* blocks;
    double omega;
    double force[3];
    const void* alpha;    // Brinkman porosity per cell (all blocks), in the arithmetic type, or null
    const SpecialCell* cells;
    uint32_t count;
    const double* values;
    double* partials;     // 16 doubles per workgroup when statistics are on:
                          // rhoBar uSqr maxUSqr count Fx Fy Fz -, then Px Py - Pz - - - -
    void* moments;        // or null: pre-collision moments (layout and type of StepArgs::moments)
    uint64_t momStride;
    double* dumpF;        // or null: pre-completion populations, 19 per record
};

template<typename S, int DYN, bool ODD, bool STATS, bool FORCED>
__global__ void __launch_bounds__(256) special_step(SpecialArgs a) {
    using T = typename Store<S>::T;
    const T omega = (T)a.omega;
    const T F[3] = {(T)a.force[0], (T)a.force[1], (T)a.force[2]};
    double st[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    double sp[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < a.count) {
        const SpecialCell sc = a.cells[k];
        const BlockGeom g = a.blocks[sc.block];
        S* __restrict__ f = reinterpret_cast<S*>(a.arena) + g.offset;
        const int x = sc.x, y = sc.y, z = sc.z;
        const uint64_t cell = sc.cell;
        const uint32_t kind = sc.info & 3u;
        const bool bulk = (sc.info >> 8) & 1u;
        T fc[19];
        if constexpr (!ODD) {
#pragma unroll
            for (int i = 0; i < 19; ++i) fc[i] = Store<S>::load(f + (uint64_t)i * g.slot + cell);
        } else {
#pragma unroll
            for (int i = 0; i < 19; ++i) fc[i] = Store<S>::load(f + swapped_index(g, x, y, z, i, sc.wall, cell));
        }
        if (a.moments) store_moments(fc, reinterpret_cast<MomentType<S>*>(a.moments), a.momStride, g.cellOffset + cell);
        if (a.dumpF) {
#pragma unroll
            for (int i = 0; i < 19; ++i) a.dumpF[19 * (uint64_t)k + i] = (double)fc[i];
        }

        if (kind == SK_MOMENTUM_EXCHANGE) {
            // bounceBackModels.hh:98-135
            T mom[3] = {T(), T(), T()};
#pragma unroll
            for (int i = 1; i < 19; ++i) {
                if (!((sc.fluid >> i) & 1u)) continue;
#pragma unroll
                for (int c = 0; c < 3; ++c) mom[c] += (T)2 * ci(i, c) * fc[opp(i)];
            }
            if (bulk) {
                st[4] += (double)(-mom[0]);
                st[5] += (double)(-mom[1]);
                st[6] += (double)(-mom[2]);
            }
#pragma unroll
            for (int i = 1; i <= 9; ++i) { T t = fc[i]; fc[i] = fc[i + 9]; fc[i + 9] = t; }
        } else {
            if (kind == SK_PLANE) complete_plane(fc, sc.info, a.values + sc.value);
            T rhoBar, drag[3] = {0, 0, 0};
            const T alpha = (FORCED && a.alpha) ? reinterpret_cast<const T*>(a.alpha)[g.cellOffset + cell] : (T)0;
            const T uSqr = collide_bulk<T, DYN, FORCED, sizeof(S) == 2>(fc, omega, F, alpha, rhoBar, drag);
            if (bulk) {
                st[0] += (double)rhoBar;
                st[1] += (double)uSqr;
                st[2] = fmax(st[2], (double)uSqr);
                st[3] += 1.0;
                if constexpr (FORCED) {
                    sp[0] += (double)drag[0];
                    sp[1] += (double)drag[1];
                    sp[3] += (double)drag[2];
                }
                // link-wise wall: each outgoing population towards the wall carries
                // 2 c_i f*_i (the stored, storage-rounded value is what is reflected)
#pragma unroll
                for (int i = 1; i < 19; ++i) {
                    if (!((sc.wall >> i) & 1u)) continue;
                    const T fo = Store<S>::round(fc[i]);
#pragma unroll
                    for (int c = 0; c < 3; ++c) st[4 + c] += (double)((T)2 * ci(i, c) * fo);
                }
            }
        }

        if constexpr (!ODD) {
#pragma unroll
            for (int i = 0; i < 19; ++i) Store<S>::store(f + (uint64_t)opp(i) * g.slot + cell, fc[i]);
        } else {
            Store<S>::store(f + cell, fc[0]);
#pragma unroll
            for (int i = 1; i < 19; ++i) {
                int dx, dy, dz;
                const bool ok = shift(x, cx(i), g.nx, g.px, dx) && shift(y, cy(i), g.ny, g.py, dy) &&
                                shift(z, cz(i), g.nz, g.pz, dz) && !((sc.wall >> i) & 1u);
                const uint64_t dst = ok ? (uint64_t)i * g.slot + lin(g, dx, dy, dz)
                                        : (uint64_t)opp(i) * g.slot + cell;
                Store<S>::store(f + dst, fc[i]);
            }
        }
    }
    if constexpr (STATS) {
        block_reduce_store(st, a.partials + 16 * (uint64_t)blockIdx.x);
        block_reduce_store(sp, a.partials + 16 * (uint64_t)blockIdx.x + 8);
    }
}

// Final statistics reduction: one workgroup, fixed order. out[0..7]:
// sumRhoBar sumUSqr maxUSqr count Fx Fy Fz -, out[8..15]: Px Py - Pz (slot 2 is a max).
// First level of the statistics reduction: workgroup w sums the bulk records
// w*256 + t, strided by the grid, into one record of the same 16-double layout
// (statistics in 0..7, the link-wall force in 8, 9, 11).
constexpr uint32_t kStatsLevel1 = 256;
static __global__ void __launch_bounds__(256) reduce_bulk_records(const double* bulk, uint64_t nBulk, double* out) {
    double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    double wf[3] = {0, 0, 0};
    for (uint64_t b = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; b < nBulk; b += (uint64_t)gridDim.x * blockDim.x) {
        const double* p = bulk + 16 * b;
        v[0] += p[0]; v[1] += p[1]; v[2] = fmax(v[2], p[2]); v[3] += p[3];
        v[4] += p[4]; v[5] += p[5]; v[6] += p[6];
        wf[0] += p[8]; wf[1] += p[9]; wf[2] += p[11];
    }
    block_reduce_store(v, out + 16 * (uint64_t)blockIdx.x);
    double q[8] = {wf[0], wf[1], 0, wf[2], 0, 0, 0, 0};
    block_reduce_store(q, out + 16 * (uint64_t)blockIdx.x + 8);
}

static __global__ void __launch_bounds__(256) reduce_partials(const double* bulk, uint64_t nBulk,
                                                       const double* special, uint64_t nSpecial,
                                                       double* out) {
    double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    double wf[3] = {0, 0, 0};
    for (uint64_t b = threadIdx.x; b < nBulk; b += blockDim.x) {
        const double* p = bulk + 16 * b;
        v[0] += p[0]; v[1] += p[1]; v[2] = fmax(v[2], p[2]); v[3] += p[3];
        v[4] += p[4]; v[5] += p[5]; v[6] += p[6];
        wf[0] += p[8]; wf[1] += p[9]; wf[2] += p[11];
    }
    double sp[3] = {0, 0, 0};
    for (uint64_t b = threadIdx.x; b < nSpecial; b += blockDim.x) {
        const double* p = special + 16 * b;
        s[0] += p[0]; s[1] += p[1]; s[2] = fmax(s[2], p[2]); s[3] += p[3];
        s[4] += p[4]; s[5] += p[5]; s[6] += p[6];
        sp[0] += p[8]; sp[1] += p[9]; sp[2] += p[11];
    }
    double r[8] = {v[0] + s[0], v[1] + s[1], fmax(v[2], s[2]), v[3] + s[3], s[4] + wf[0], s[5] + wf[1], s[6] + wf[2], 0};
    block_reduce_store(r, out);
    double q[8] = {v[4] + sp[0], v[5] + sp[1], 0, v[6] + sp[2], 0, 0, 0, 0};
    block_reduce_store(q, out + 8);
}

// ---------------------------------------------------------------------------
// Envelope exchange (multi-block): copies the natural-layout populations of
// the bulk cells that make up each envelope, segment by segment
// (SerialBlockCommunicator3D::duplicateOverlaps, serialBlockCommunicator3D.cpp:43-95,
// BlockLatticeDataTransfer3D::attribute_static, blockLattice3D.hh:873-885).

struct CopySegment {
    uint16_t src, dst;            // block indices
    uint16_t pad[2];
    int sx, sy, sz;               // block-local origin in the source
    int dx, dy, dz;               // block-local origin in the destination
    int lx, ly, lz;               // extent
    uint64_t cellBegin;           // first cell of the segment in the launch
};

// One workgroup per task: 256 consecutive cells of one segment.
struct CopyTask {
    uint32_t seg;
    uint32_t first;
};

template<typename S>
__global__ void __launch_bounds__(256) exchange_envelopes(void* arenaV, const BlockGeom* blocks,
                                                          const CopySegment* segs, const CopyTask* tasks) {
    S* arena = reinterpret_cast<S*>(arenaV);
    const CopyTask task = tasks[blockIdx.x];
    const CopySegment s = segs[task.seg];
    const uint64_t r = (uint64_t)task.first + threadIdx.x;
    if (r < (uint64_t)s.lx * s.ly * s.lz) {
        const int lz = (int)(r % (uint64_t)s.lz), ly = (int)((r / s.lz) % (uint64_t)s.ly);
        const int lx = (int)(r / ((uint64_t)s.lz * s.ly));
        const BlockGeom& gs = blocks[s.src];
        const BlockGeom& gd = blocks[s.dst];
        const S* src = arena + gs.offset + lin(gs, s.sx + lx, s.sy + ly, s.sz + lz);
        S* dst = arena + gd.offset + lin(gd, s.dx + lx, s.dy + ly, s.dz + lz);
        // plain (cached) accesses: the source cells were just written by the step
        // kernel and the next step reads the envelope back
        S v[19];
#pragma unroll
        for (int i = 0; i < 19; ++i) v[i] = src[(uint64_t)i * gs.slot];
#pragma unroll
        for (int i = 0; i < 19; ++i) dst[(uint64_t)i * gd.slot] = v[i];
    }
}

// ---------------------------------------------------------------------------
// Transfers and analysis on one block: the cells of the global box I that the
// block owns, staged at their position in the caller's box G.

struct BoxArgs {
    BlockGeom g;
    int block;
    int parity;
    int ix0, iy0, iz0, ilx, ily, ilz;   // iteration box, block-local
    int ox, oy, oz;                     // offset of the iteration box inside the staging box
    int sly, slz;                       // staging box extents along y, z
    const SpecialCell* special;         // sorted by (block, cell)
    uint32_t nSpecial;
    const uint32_t* walls;              // link-wall mask per cell (all blocks), or null
};

__device__ __forceinline__ int find_special(const BoxArgs& b, uint64_t cell) {
    const uint64_t key = ((uint64_t)b.block << 32) | cell;
    uint32_t lo = 0, hi = b.nSpecial;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) / 2;
        const uint64_t mk = ((uint64_t)b.special[mid].block << 32) | b.special[mid].cell;
        if (mk < key) lo = mid + 1; else hi = mid;
    }
    if (lo < b.nSpecial && b.special[lo].block == b.block && b.special[lo].cell == cell) return (int)lo;
    return -1;
}

// Link-wall mask of a cell: bulk cells next to a wall keep it in the per-cell
// array, special cells (planes, momentum-exchange nodes) in their record.
__device__ __forceinline__ uint32_t wall_mask_of(const BoxArgs& b, uint64_t cell) {
    if (b.walls && b.walls[b.g.cellOffset + cell]) return b.walls[b.g.cellOffset + cell];
    const int s = find_special(b, cell);
    return s >= 0 ? b.special[s].wall : 0u;
}

struct BoxCell {
    int x, y, z;          // block-local
    uint64_t cell;        // block-local linear
    uint64_t staged;      // index in the staging box
};

__device__ __forceinline__ BoxCell box_cell(const BoxArgs& b, uint64_t k) {
    BoxCell c;
    int lx, ly, lz;
    if (k < (1ull << 32)) {
        // 32-bit division (a 64-bit one is a long instruction sequence per cell)
        const uint32_t k32 = (uint32_t)k, row = k32 / (uint32_t)b.ilz;
        lz = (int)(k32 - row * (uint32_t)b.ilz);
        lx = (int)(row / (uint32_t)b.ily);
        ly = (int)(row - (uint32_t)lx * (uint32_t)b.ily);
    } else {
        lz = (int)(k % (uint64_t)b.ilz);
        ly = (int)((k / b.ilz) % (uint64_t)b.ily);
        lx = (int)(k / ((uint64_t)b.ilz * b.ily));
    }
    c.x = b.ix0 + lx; c.y = b.iy0 + ly; c.z = b.iz0 + lz;
    c.cell = lin(b.g, c.x, c.y, c.z);
    c.staged = ((uint64_t)(b.ox + lx) * b.sly + (b.oy + ly)) * b.slz + (b.oz + lz);
    return c;
}

#define PLBAMD_BOX_LOOP(b)                                                                      \
    const uint64_t n_ = (uint64_t)(b).ilx * (b).ily * (b).ilz;                                  \
    for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < n_;                  \
         k += (uint64_t)gridDim.x * blockDim.x)

// Transfers between the arena (slot-major) and a staging buffer of cell
// records ([cell][19], C order): 256 cells per workgroup pass through LDS, so
// that the staging side is read or written as consecutive elements (a lane per
// element) and the arena side as slots (a lane per cell). H is the staging
// element (FP64 / FP32 values, or uint16 raw bits when RAW).
template<typename S, typename H, bool RAW>
__device__ __forceinline__ void upload_box_t(void* arena, BoxArgs b, const H* host) {
    using T = typename Store<S>::T;
    __shared__ H tile[256 * 19];
    __shared__ uint64_t stg[256];
    S* f = reinterpret_cast<S*>(arena) + b.g.offset;
    const uint64_t n = (uint64_t)b.ilx * b.ily * b.ilz;
    for (uint64_t k0 = (uint64_t)blockIdx.x * 256; k0 < n; k0 += (uint64_t)gridDim.x * 256) {
        const uint32_t cnt = (uint32_t)min((uint64_t)256, n - k0);
        const uint64_t k = k0 + threadIdx.x;
        BoxCell c{};
        if (threadIdx.x < cnt) {
            c = box_cell(b, k);
            stg[threadIdx.x] = c.staged;
        }
        __syncthreads();
        for (uint32_t e = threadIdx.x; e < cnt * 19; e += 256) {
            const uint32_t j = e / 19;
            tile[e] = host[stg[j] * 19 + (e - j * 19)];
        }
        __syncthreads();
        if (threadIdx.x < cnt) {
            const uint32_t wall = b.parity ? wall_mask_of(b, c.cell) : 0;
            for (int i = 0; i < 19; ++i) {
                const uint64_t at = pop_index(b.g, b.parity, c.x, c.y, c.z, i, wall, c.cell);
                if constexpr (RAW) reinterpret_cast<uint16_t*>(f)[at] = (uint16_t)tile[threadIdx.x * 19 + i];
                else Store<S>::store(f + at, (T)tile[threadIdx.x * 19 + i]);
            }
        }
        __syncthreads();
    }
}

template<typename S, typename H, bool RAW>
__device__ __forceinline__ void download_box_t(const void* arena, BoxArgs b, H* host) {
    __shared__ H tile[256 * 19];
    __shared__ uint64_t stg[256];
    const S* f = reinterpret_cast<const S*>(arena) + b.g.offset;
    const uint64_t n = (uint64_t)b.ilx * b.ily * b.ilz;
    for (uint64_t k0 = (uint64_t)blockIdx.x * 256; k0 < n; k0 += (uint64_t)gridDim.x * 256) {
        const uint32_t cnt = (uint32_t)min((uint64_t)256, n - k0);
        if (threadIdx.x < cnt) {
            const BoxCell c = box_cell(b, k0 + threadIdx.x);
            stg[threadIdx.x] = c.staged;
            const uint32_t wall = b.parity ? wall_mask_of(b, c.cell) : 0;
            for (int i = 0; i < 19; ++i) {
                const uint64_t at = pop_index(b.g, b.parity, c.x, c.y, c.z, i, wall, c.cell);
                if constexpr (RAW) tile[threadIdx.x * 19 + i] = (H)reinterpret_cast<const uint16_t*>(f)[at];
                else tile[threadIdx.x * 19 + i] = (H)Store<S>::load(f + at);
            }
        }
        __syncthreads();
        for (uint32_t e = threadIdx.x; e < cnt * 19; e += 256) {
            const uint32_t j = e / 19;
            host[stg[j] * 19 + (e - j * 19)] = tile[e];
        }
        __syncthreads();
    }
}

template<typename S, typename H>
__global__ void __launch_bounds__(256) upload_box(void* arena, BoxArgs b, const H* host) {
    upload_box_t<S, H, false>(arena, b, host);
}
template<typename S>
__global__ void __launch_bounds__(256) upload_box_raw(void* arena, BoxArgs b, const uint16_t* host) {
    upload_box_t<S, uint16_t, true>(arena, b, host);
}
template<typename S, typename H>
__global__ void __launch_bounds__(256) download_box(const void* arena, BoxArgs b, H* host) {
    download_box_t<S, H, false>(arena, b, host);
}
template<typename S>
__global__ void __launch_bounds__(256) download_box_raw(const void* arena, BoxArgs b, uint16_t* host) {
    download_box_t<S, uint16_t, true>(arena, b, host);
}

// Completed snapshot of wrapped-factory plane cells: the populations the
// wrapped post-stream processor leaves in Palabos's lattice
// (wrappedLocalBoundaryProcessor3D.hh:41-62).
template<typename S, typename H>
__global__ void complete_download(const void* arena, BoxArgs b, const double* values, H* host, int raw) {
    using T = typename Store<S>::T;
    const S* f = reinterpret_cast<const S*>(arena) + b.g.offset;
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < b.nSpecial; k += gridDim.x * blockDim.x) {
        const SpecialCell sc = b.special[k];
        if (sc.block != b.block || (sc.info & 3u) != SK_PLANE || !((sc.info >> 7) & 1u)) continue;
        const int x = sc.x, y = sc.y, z = sc.z;
        if (x < b.ix0 || x >= b.ix0 + b.ilx || y < b.iy0 || y >= b.iy0 + b.ily || z < b.iz0 || z >= b.iz0 + b.ilz)
            continue;
        T fc[19];
        for (int i = 0; i < 19; ++i)
            fc[i] = Store<S>::load(f + pop_index(b.g, b.parity, x, y, z, i, b.parity ? sc.wall : 0u, sc.cell));
        complete_plane(fc, sc.info, values + sc.value);
        const uint64_t o = (((uint64_t)(b.ox + x - b.ix0) * b.sly + (b.oy + y - b.iy0)) * b.slz + (b.oz + z - b.iz0)) * 19;
        for (int i = 0; i < 19; ++i) {
            if (raw) reinterpret_cast<uint16_t*>(host)[o + i] = Store<S>::bits(fc[i]);
            else host[o + i] = (H)Store<S>::round(fc[i]);
        }
    }
}

// Equilibrium initialisation (core/cell.hh:124-132 iniCellAtEquilibrium) in the
// compute precision of the lattice.
template<typename S>
__global__ void equilibrium_box(void* arena, BoxArgs b, const double* rho, const double* u) {
    using T = typename Store<S>::T;
    S* f = reinterpret_cast<S*>(arena) + b.g.offset;
    PLBAMD_BOX_LOOP(b) {
        const BoxCell c = box_cell(b, k);
        const uint32_t wall = b.parity ? wall_mask_of(b, c.cell) : 0;
        const T density = (T)rho[c.staged];
        const T j[3] = {(T)u[3 * c.staged] * density, (T)u[3 * c.staged + 1] * density, (T)u[3 * c.staged + 2] * density};
        const T jSqr = j[0] * j[0] + j[1] * j[1] + j[2] * j[2];
        const T rhoBar = density - (T)1;
        const T ir = inv_rho(rhoBar);
        for (int i = 0; i < 19; ++i)
            Store<S>::store(f + pop_index(b.g, b.parity, c.x, c.y, c.z, i, wall, c.cell), feq(i, rhoBar, ir, j, jSqr));
    }
}

struct MomentsArgs {
    const uint8_t* tags;
    const double* values;
    double force[3];
    int forced;              // Guo force: the velocity is j / rho + F / 2
    double* fields;          // or null: per staged cell the 7 analysis values
    double solidRho[3];      // density parameter of BounceBack, NoDynamics, momentum-exchange nodes
};

// Per-cell density and velocity as Palabos's analysis functions see them
// (plane_macroscopic_any and the rules listed there), and their FP64 sums.
// Bulk cells: rho = fullRho(get_rhoBar) with the sequential sum
// (momentTemplates3D.h:375-380, BasicBulkDynamics::computeDensity,
// core/dynamics.hh:278-281); u = get_j * invRho(get_rhoBar) (compute_uLb,
// momentTemplates3D.h:138-144), or with a Guo force j * invRho + F / 2 where
// rhoBar comes with j from get_rhoBar_j (ExternalForceDynamics::computeVelocity,
// basicDynamics/externalForceDynamics.hh:59-74). BounceBack, NoDynamics and
// momentum-exchange nodes (core/dynamics.hh:814-880, 994-1060;
// bounceBackModels.hh:160-223): u = 0, rho = the dynamics' density parameter,
// rhoBar = rho - 1, except NoDynamics, whose computeRhoBar is
// 1 - SkordosFactor = 0 (core/dynamics.hh:1055-1058). Link-wall solids (no
// Palabos class): rho = 1, rhoBar = 0, u = 0. With m.fields, the analysis fields
// of dataAnalysisWrapper3D.h:695-777 in T: computeDensity, computeRhoBar,
// computeVelocity (3), computeKineticEnergy = normSqr(u) / 2 and
// computeVelocityNorm = sqrt(normSqr(u)) (dataAnalysisFunctional3D.hh:668-719).
// PAR: the parity of the layout (b.parity), a template parameter so that the
// natural layout's addressing is a plain slot offset.
template<typename S, int PAR>
__global__ void __launch_bounds__(256) moments_box(const void* arena, BoxArgs b, MomentsArgs m, double* rhoOut,
                                                   double* uOut, double* partials) {
    using T = typename Store<S>::T;
    const S* f = reinterpret_cast<const S*>(arena) + b.g.offset;
    double st[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const uint64_t n = (uint64_t)b.ilx * b.ily * b.ilz;
    for (uint64_t k0 = (uint64_t)blockIdx.x * 256; k0 < n; k0 += (uint64_t)gridDim.x * 256) {
        const uint32_t cnt = (uint32_t)min((uint64_t)256, n - k0);
        if (threadIdx.x < cnt) {
            const uint64_t k = k0 + threadIdx.x;
            const BoxCell c = box_cell(b, k);
            const uint8_t tag = m.tags ? m.tags[b.g.cellOffset + c.cell] : (uint8_t)BT_BULK;
            const int sp = (tag == BT_SKIP) ? find_special(b, c.cell) : -1;
            const uint32_t wall = !PAR ? 0u
                                  : (tag == BT_WALL_NEAR && b.walls) ? b.walls[b.g.cellOffset + c.cell]
                                  : (sp >= 0 ? b.special[sp].wall : 0u);
            T rhoBar = T(), u[3] = {T(), T(), T()};
            const bool solid = tag == BT_BOUNCE_BACK || tag == BT_NO_DYNAMICS ||
                               (tag == BT_SKIP && (sp < 0 || (b.special[sp].info & 3u) == SK_MOMENTUM_EXCHANGE));
            T rho = T(1);
            if (solid) {
                if (tag == BT_BOUNCE_BACK) {
                    rho = (T)m.solidRho[0];
                    rhoBar = rho - (T)1;
                } else if (tag == BT_NO_DYNAMICS) {
                    rho = (T)m.solidRho[1];
                } else if (sp >= 0) {
                    rho = (T)m.solidRho[2];
                    rhoBar = rho - (T)1;
                }
            } else if (sp >= 0 && ((b.special[sp].info & 3u) == SK_PLANE ||
                                   (b.special[sp].info & 3u) == SK_STORE_VELOCITY)) {
                continue;   // plane, edge and corner cells: moments_planes (their closures would
                            // triple this kernel's registers)
            } else {
                T fc[19];
                for (int i = 0; i < 19; ++i) fc[i] = Store<S>::load(f + pop_index(b.g, PAR, c.x, c.y, c.z, i, wall, c.cell));
                T rb, j[3];
                rhobar_j(fc, rb, j);
                rhoBar = rhobar_sum(fc);
                const T ir = inv_rho(m.forced ? rb : rhoBar);
                for (int a2 = 0; a2 < 3; ++a2) u[a2] = m.forced ? j[a2] * ir + (T)m.force[a2] / (T)2 : j[a2] * ir;
                rho = rhoBar + (T)1;
            }
            if (rhoOut) rhoOut[c.staged] = (double)rho;
            if (uOut) {
                uOut[3 * c.staged] = (double)u[0];
                uOut[3 * c.staged + 1] = (double)u[1];
                uOut[3 * c.staged + 2] = (double)u[2];
            }
            const T uSqr = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
            if (m.fields) {
                // (measured: direct stores beat an LDS transpose here, whose
                // barriers cost more at this kernel's occupancy of 3 waves)
                double* o = m.fields + 7 * c.staged;
                o[0] = (double)rho;
                o[1] = (double)rhoBar;
                o[2] = (double)u[0];
                o[3] = (double)u[1];
                o[4] = (double)u[2];
                o[5] = (double)(uSqr / (T)2);
                o[6] = (double)sqrt(uSqr);
            }
            st[0] += (double)rhoBar;
            st[1] += (double)uSqr;
            st[3] += (double)u[0];
            st[4] += (double)u[1];
            st[5] += (double)u[2];
        }
    }
    if (partials) block_reduce_store(st, partials + 8 * (uint64_t)blockIdx.x);
}

// The plane cells of moments_box (velocity / pressure planes): their density and
// velocity from the closure of their boundary dynamics (plane_macroscopic_any),
// the same outputs and per-workgroup FP64 records. Edge and corner cells
// (StoreVelocityDynamics): computeDensity = fullRho(get_rhoBar) of the
// populations, computeRhoBar = rhoBar of that density, computeVelocity = the
// stored velocity (boundaryDynamics.hh:82-87, :324-344).
template<typename S>
__global__ void __launch_bounds__(256) moments_planes(const void* arena, BoxArgs b, MomentsArgs m, double* rhoOut,
                                                      double* uOut, double* partials) {
    using T = typename Store<S>::T;
    const S* f = reinterpret_cast<const S*>(arena) + b.g.offset;
    double st[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < b.nSpecial; k += gridDim.x * blockDim.x) {
        const SpecialCell sc = b.special[k];
        const uint32_t kind = sc.info & 3u;
        if (sc.block != b.block || (kind != SK_PLANE && kind != SK_STORE_VELOCITY)) continue;
        const int x = sc.x, y = sc.y, z = sc.z;
        if (x < b.ix0 || x >= b.ix0 + b.ilx || y < b.iy0 || y >= b.iy0 + b.ily || z < b.iz0 || z >= b.iz0 + b.ilz)
            continue;
        T fc[19];
        for (int i = 0; i < 19; ++i)
            fc[i] = Store<S>::load(f + pop_index(b.g, b.parity, x, y, z, i, b.parity ? sc.wall : 0u, sc.cell));
        T rhoBar, u[3], rho;
        if (kind == SK_PLANE) {
            plane_macroscopic_any(fc, sc.info, m.values + sc.value, rhoBar, u);
            rho = rhoBar + (T)1;
        } else {
            rho = rhobar_sum(fc) + (T)1;
            rhoBar = rho - (T)1;
            for (int c = 0; c < 3; ++c) u[c] = (T)m.values[sc.value + c];
        }
        const uint64_t staged = ((uint64_t)(b.ox + x - b.ix0) * b.sly + (b.oy + y - b.iy0)) * b.slz + (b.oz + z - b.iz0);
        if (rhoOut) rhoOut[staged] = (double)rho;
        if (uOut) {
            uOut[3 * staged] = (double)u[0];
            uOut[3 * staged + 1] = (double)u[1];
            uOut[3 * staged + 2] = (double)u[2];
        }
        const T uSqr = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
        if (m.fields) {
            double* o = m.fields + 7 * staged;
            o[0] = (double)rho;
            o[1] = (double)rhoBar;
            o[2] = (double)u[0];
            o[3] = (double)u[1];
            o[4] = (double)u[2];
            o[5] = (double)(uSqr / (T)2);
            o[6] = (double)sqrt(uSqr);
        }
        st[0] += (double)rhoBar;
        st[1] += (double)uSqr;
        st[3] += (double)u[0];
        st[4] += (double)u[1];
        st[5] += (double)u[2];
    }
    if (partials) block_reduce_store(st, partials + 8 * (uint64_t)blockIdx.x);
}

// Sum of n 8-double records (slot 2 unused) in a fixed order. One workgroup
// per gridDim.x: workgroup w sums records w, w + gridDim.x, ... into out[w];
// a launch of many workgroups followed by one of a single workgroup over
// their outputs is the two-level form.
static __global__ void __launch_bounds__(256) sum_records(const double* p, uint32_t n, double* out) {
    double r[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < n; b += gridDim.x * blockDim.x)
        for (int k = 0; k < 8; ++k) if (k != 2) r[k] += p[8 * b + k];
    block_reduce_store(r, out + 8 * (uint64_t)blockIdx.x);
}

}  // namespace plbamd
