// palabos_amd engine: D3Q19 collide-and-stream on gfx950 (MI355X).
//
// Data layout in HBM (one lattice = one Palabos block, envelope included):
//   populations  SoA, slot-major: f[slot][x][y][z], z fastest, each slot padded
//                to a multiple of 64 cells; storage type double | float |
//                f16 bits | bf16 bits (f-bar = f - t_i, roundOffPolicy.h:73-90)
//   tags         uint8 per cell, only when the lattice has non-bulk cells:
//                0 bulk, 1 BounceBack node, 2 NoDynamics, 3 handled by the
//                special-cell list (planes, momentum-exchange nodes, cells next
//                to link-wise walls) or a link-wise wall cell (never updated)
//   special list 32-byte records, sorted by cell index
//
// Streaming is the in-place AA pattern, one array, no second lattice:
//   even step  each cell reads its own slots i, collides, writes f*_i to its
//              own slot opp(i)                        (all accesses local)
//   odd step   each cell gathers f_i from slot opp(i) of cell x - c_i, collides,
//              scatters f*_i to slot i of cell x + c_i (back to natural layout)
// A source or destination outside a non-periodic edge (or a link-wise wall)
// is replaced by the cell's own opposite slot: that is Palabos's same-step
// edge reflection (atomicBlock/blockLattice3D.hh:403-427). Every memory slot
// is read and written by exactly one cell per step, so the update is race-free
// and the cell order is free, as it is in Palabos.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/palabos_amd.h"
#include "d3q19_device.h"

namespace plbamd {

// ---------------------------------------------------------------------------
// Storage types

struct f16s { uint16_t b; };
struct bf16s { uint16_t b; };

template<typename S> struct Store;
template<> struct Store<double> {
    using T = double;
    static __device__ __forceinline__ double load(const double* p) { return __builtin_nontemporal_load(p); }
    static __device__ __forceinline__ void store(double* p, double v) { __builtin_nontemporal_store(v, p); }
    static __device__ __forceinline__ double round(double v) { return v; }
    static __device__ __forceinline__ uint16_t bits(double) { return 0; }
};
template<> struct Store<float> {
    using T = float;
    static __device__ __forceinline__ float load(const float* p) { return __builtin_nontemporal_load(p); }
    static __device__ __forceinline__ void store(float* p, float v) { __builtin_nontemporal_store(v, p); }
    static __device__ __forceinline__ float round(float v) { return v; }
    static __device__ __forceinline__ uint16_t bits(float) { return 0; }
};
template<> struct Store<f16s> {
    using T = float;
    static __device__ __forceinline__ float load(const f16s* p) {
        uint16_t b = __builtin_nontemporal_load(reinterpret_cast<const uint16_t*>(p));
        _Float16 h;
        __builtin_memcpy(&h, &b, 2);
        return (float)h;
    }
    static __device__ __forceinline__ void store(f16s* p, float v) {
        _Float16 h = (_Float16)v;  // round to nearest even
        uint16_t b;
        __builtin_memcpy(&b, &h, 2);
        __builtin_nontemporal_store(b, reinterpret_cast<uint16_t*>(p));
    }
    static __device__ __forceinline__ float round(float v) { return (float)(_Float16)v; }
    static __device__ __forceinline__ uint16_t bits(float v) {
        _Float16 h = (_Float16)v;
        uint16_t b;
        __builtin_memcpy(&b, &h, 2);
        return b;
    }
};
__device__ __forceinline__ uint16_t bf16_round(float v) {
    uint32_t x = __float_as_uint(v);
    if ((x & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((x >> 16) | 0x40u);
    x += 0x7fffu + ((x >> 16) & 1u);
    return (uint16_t)(x >> 16);
}
template<> struct Store<bf16s> {
    using T = float;
    static __device__ __forceinline__ float load(const bf16s* p) {
        uint16_t b = __builtin_nontemporal_load(reinterpret_cast<const uint16_t*>(p));
        return __uint_as_float((uint32_t)b << 16);
    }
    static __device__ __forceinline__ void store(bf16s* p, float v) {
        __builtin_nontemporal_store(bf16_round(v), reinterpret_cast<uint16_t*>(p));
    }
    static __device__ __forceinline__ float round(float v) { return __uint_as_float((uint32_t)bf16_round(v) << 16); }
    static __device__ __forceinline__ uint16_t bits(float v) { return bf16_round(v); }
};

// ---------------------------------------------------------------------------
// Kernel arguments

enum BulkTag : uint8_t { BT_BULK = 0, BT_BOUNCE_BACK = 1, BT_NO_DYNAMICS = 2, BT_SKIP = 3 };

struct Geometry {
    int nx, ny, nz;
    int px, py, pz;          // periodic flags
    uint64_t slot;           // elements per population slot
};

struct StepArgs {
    void* f;
    Geometry g;
    int zChunk;              // z cells per workgroup row segment
    int nzChunks;            // row segments per row
    double omega;
    double force[3];
    const uint8_t* tags;     // BulkTag per cell or null
    const double* alpha;     // Brinkman porosity per cell or null
    double* partials;        // 8 doubles per workgroup when statistics are on
};

enum SpecialKind : uint32_t { SK_PLANE = 0, SK_MOMENTUM_EXCHANGE = 1, SK_NEAR_WALL = 2 };

struct SpecialCell {
    uint32_t cell;
    uint16_t x, y, z;
    uint16_t info;        // kind | family << 2 | bckind << 3 | dir << 4 | (ori > 0) << 6 | wrapped << 7
    uint32_t value;       // index into the plane value array (velocity: 3 doubles, pressure: 1)
    uint32_t wall;        // bit i: neighbour x + c_i is a link-wise wall
    uint32_t fluid;       // bit i: fluid direction of a momentum-exchange node
    uint32_t pad[2];
};
static_assert(sizeof(SpecialCell) == 32, "special record is 32 bytes");

struct SpecialArgs {
    void* f;
    Geometry g;
    double omega;
    const SpecialCell* cells;
    uint32_t count;
    const double* values;
    double* partials;     // 8 doubles per workgroup when statistics are on
};

// ---------------------------------------------------------------------------
// Neighbour arithmetic

__device__ __forceinline__ bool shift(int v, int d, int n, int periodic, int& out) {
    int w = v + d;
    if (w < 0) { if (!periodic) return false; w += n; }
    else if (w >= n) { if (!periodic) return false; w -= n; }
    out = w;
    return true;
}

__device__ __forceinline__ uint64_t lin(const Geometry& g, int x, int y, int z) {
    return ((uint64_t)x * g.ny + y) * g.nz + z;
}

// Position of f_i of cell (x,y,z) in the swapped (after-even-step) layout:
// slot opp(i) at x - c_i, or slot i at the cell itself when the source reflects.
__device__ __forceinline__ uint64_t swapped_index(const Geometry& g, int x, int y, int z, int i,
                                                  uint32_t wallMask, uint64_t cell) {
    if (i == 0) return cell;
    int sx, sy, sz;
    bool ok = shift(x, -cx(i), g.nx, g.px, sx) && shift(y, -cy(i), g.ny, g.py, sy) &&
              shift(z, -cz(i), g.nz, g.pz, sz) && !((wallMask >> opp(i)) & 1u);
    if (ok) return (uint64_t)opp(i) * g.slot + lin(g, sx, sy, sz);
    return (uint64_t)i * g.slot + cell;
}

// ---------------------------------------------------------------------------
// Block reduction of the 8-double statistics record (fixed order: deterministic)

__device__ __forceinline__ void block_reduce_store(double v[8], double* out) {
    __shared__ double red[256][8];
    const int t = threadIdx.x + threadIdx.y * blockDim.x;
    const int n = blockDim.x * blockDim.y;
#pragma unroll
    for (int k = 0; k < 8; ++k) red[t][k] = v[k];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (t < s && t + s < n) {
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                if (k == 2) red[t][k] = fmax(red[t][k], red[t + s][k]);
                else red[t][k] += red[t + s][k];
            }
        }
        __syncthreads();
    }
    if (t == 0) {
#pragma unroll
        for (int k = 0; k < 8; ++k) out[k] = red[0][k];
    }
}

// ---------------------------------------------------------------------------
// Bulk kernel: one workgroup per (x, rows y..y+TY-1), threads stride along z.
//
// A wavefront always lies in one (x, y) row (the z extent of a workgroup is a
// multiple of 64), so the nine neighbouring rows (x-1..x+1, y-1..y+1) and
// their validity are wavefront-uniform and live in scalar registers. A load or
// store is then a scalar row base (slot + row start) plus one of three 32-bit
// per-lane byte offsets (z-1, z, z+1): the global_load/store scalar-base form,
// with no per-population vector address arithmetic. Wavefronts touching a
// non-periodic edge (or a periodic wrap in x/y that is not a plain row shift)
// take the general path with per-lane edge selection.

template<typename S>
__device__ __forceinline__ S* at(S* base, uint32_t byteOffset) {
    return reinterpret_cast<S*>(reinterpret_cast<char*>(base) + byteOffset);
}

template<typename S, int DYN, bool ODD, bool TAGS, bool STATS, bool FORCED>
__global__ void __launch_bounds__(256) bulk_step(StepArgs a) {
    using T = typename Store<S>::T;
    S* __restrict__ f = reinterpret_cast<S*>(a.f);
    const Geometry g = a.g;
    const int x = blockIdx.y;
    const int zChunk = a.zChunk;
    const int chunk = blockIdx.x % a.nzChunks;
    const int y = __builtin_amdgcn_readfirstlane((blockIdx.x / a.nzChunks) * blockDim.y + threadIdx.y);
    const int zBegin = chunk * zChunk;
    const int zEnd = min(g.nz, zBegin + zChunk);
    const T omega = (T)a.omega;
    const T F[3] = {(T)a.force[0], (T)a.force[1], (T)a.force[2]};
    double st[8] = {0, 0, 0, 0, 0, 0, 0, 0};

    if (y < g.ny) {
        // Row neighbours (wavefront-uniform): x-1, x+1, y-1, y+1.
        int xm = 0, xp = 0, ym = 0, yp = 0;
        const bool vxm = shift(x, -1, g.nx, g.px, xm), vxp = shift(x, 1, g.nx, g.px, xp);
        const bool vym = shift(y, -1, g.ny, g.py, ym), vyp = shift(y, 1, g.ny, g.py, yp);
        const int xs[3] = {xm, x, xp};
        const int ys[3] = {ym, y, yp};
        const bool vx[3] = {vxm, true, vxp};
        const bool vy[3] = {vym, true, vyp};
        const bool rowsValid = vxm && vxp && vym && vyp;
        const uint64_t rowBase = ((uint64_t)x * g.ny + y) * g.nz;
        // slot-0 pointer of each neighbouring row
        S* rows[3][3];
#pragma unroll
        for (int ax = 0; ax < 3; ++ax)
#pragma unroll
            for (int ay = 0; ay < 3; ++ay)
                rows[ax][ay] = f + ((uint64_t)xs[ax] * g.ny + ys[ay]) * g.nz;

        for (int z = zBegin + threadIdx.x; z < zEnd; z += blockDim.x) {
            const uint64_t cell = rowBase + z;
            uint8_t tag = BT_BULK;
            if constexpr (TAGS) {
                tag = a.tags[cell];
                if (tag == BT_SKIP) continue;
            }
            int zm = 0, zp = 0;
            const bool vzm = shift(z, -1, g.nz, g.pz, zm), vzp = shift(z, 1, g.nz, g.pz, zp);
            const uint32_t zb[3] = {(uint32_t)zm * (uint32_t)sizeof(S), (uint32_t)z * (uint32_t)sizeof(S),
                                    (uint32_t)zp * (uint32_t)sizeof(S)};
            bool fast = true;
            if constexpr (ODD) {
                const uint64_t active = __ballot(1);
                fast = rowsValid && (__ballot(vzm && vzp) == active);
            }

            T fc[19];
            if constexpr (!ODD) {
#pragma unroll
                for (int i = 0; i < 19; ++i) fc[i] = Store<S>::load(at(rows[1][1] + (uint64_t)i * g.slot, zb[1]));
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
                    // source x - c_i
                    const int ax = 1 - cx(i), ay = 1 - cy(i), az = 1 - cz(i);
                    const bool ok = vx[ax] && vy[ay] && vz[az];
                    const uint64_t src = ok ? (uint64_t)opp(i) * g.slot +
                                                  (((uint64_t)xs[ax] * g.ny + ys[ay]) * g.nz + zs[az])
                                            : (uint64_t)i * g.slot + cell;
                    fc[i] = Store<S>::load(f + src);
                }
            }

            if (tag == BT_BULK) {
                T rhoBar;
                T drag[3] = {0, 0, 0};
                const T alpha = (FORCED && a.alpha) ? (T)a.alpha[cell] : (T)0;
                const T uSqr = collide_bulk<T, DYN, FORCED>(fc, omega, F, alpha, rhoBar, drag);
                if constexpr (STATS) {
                    st[0] += (double)rhoBar;
                    st[1] += (double)uSqr;
                    st[2] = fmax(st[2], (double)uSqr);
                    st[3] += 1.0;
                    if constexpr (FORCED) {
                        st[4] += (double)drag[0];
                        st[5] += (double)drag[1];
                        st[6] += (double)drag[2];
                    }
                }
            } else if (tag == BT_BOUNCE_BACK) {
#pragma unroll
                for (int i = 1; i <= 9; ++i) { T t = fc[i]; fc[i] = fc[i + 9]; fc[i + 9] = t; }
            }

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
                    // destination x + c_i
                    const int ax = 1 + cx(i), ay = 1 + cy(i), az = 1 + cz(i);
                    const bool ok = vx[ax] && vy[ay] && vz[az];
                    const uint64_t dst = ok ? (uint64_t)i * g.slot +
                                                  (((uint64_t)xs[ax] * g.ny + ys[ay]) * g.nz + zs[az])
                                            : (uint64_t)opp(i) * g.slot + cell;
                    Store<S>::store(f + dst, fc[i]);
                }
            }
        }
    }
    if constexpr (STATS) {
        block_reduce_store(st, a.partials + 8 * (uint64_t)(blockIdx.y * gridDim.x + blockIdx.x));
    }
}

// ---------------------------------------------------------------------------
// Plane completions (dynamics-based composite: completion, then base collision)

// VelocityDirichletBoundaryDynamics::computeRhoBar/computeRhoBarJ
// (boundaryCondition/boundaryDynamics.hh:643-700), velIsJ() false.
template<typename T, int DIR, int ORI>
__device__ __forceinline__ void velocity_moments(const T* f, const T u[3], T& rhoBar, T j[3]) {
    T rhoOnWall = T();
#pragma unroll
    for (int i = 0; i < 19; ++i) if (ci(i, DIR) == 0) rhoOnWall += f[i];
    T rhoNormal = T();
#pragma unroll
    for (int i = 0; i < 19; ++i) if (ci(i, DIR) == ORI) rhoNormal += f[i];
    const T velNormal = (T)ORI * u[DIR];
    rhoBar = ((T)2 * rhoNormal + rhoOnWall - velNormal) / ((T)1 + velNormal);
    const T rho = rhoBar + (T)1;
#pragma unroll
    for (int c = 0; c < 3; ++c) j[c] = rho * u[c];
}

// DensityDirichletBoundaryDynamics::computeJ (boundaryDynamics.hh:776-800).
template<typename T, int DIR, int ORI>
__device__ __forceinline__ void pressure_moments(const T* f, T rhoBar, T j[3]) {
    j[0] = j[1] = j[2] = T();
    T rhoOnWall = T();
#pragma unroll
    for (int i = 0; i < 19; ++i) if (ci(i, DIR) == 0) rhoOnWall += f[i];
    T rhoNormal = T();
#pragma unroll
    for (int i = 0; i < 19; ++i) if (ci(i, DIR) == ORI) rhoNormal += f[i];
    j[DIR] = (T)ORI * ((T)2 * rhoNormal + rhoOnWall - rhoBar);
}

// ZouHeClosure (boundaryCondition/zouHeDynamics.hh:44-107).
template<typename T, int DIR, int ORI>
__device__ __forceinline__ void zouhe_close(T* f, T rhoBar, const T j[3]) {
    const T jSqr = j[0] * j[0] + j[1] * j[1] + j[2] * j[2];
    const T ir = inv_rho(rhoBar);
    // missing populations: opposite of c[DIR] == ORI, in index order of the latter
#pragma unroll
    for (int n = 0; n < 19; ++n) {
        if (ci(n, DIR) != ORI) continue;
        const int p = opp(n), o = n;
        f[p] = f[o] - feq(o, rhoBar, ir, j, jSqr) + feq(p, rhoBar, ir, j, jSqr);
    }
    T fRhoBar, fJ[3];
    rhobar_j(f, fRhoBar, fJ);
    const T fir = inv_rho(fRhoBar);
    const T falseRho = fRhoBar + (T)1;
    T diff[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) diff[c] = (j[c] - falseRho * (fJ[c] * fir)) / (T)4;
    // the four diagonal missing populations (all but the normal one)
#pragma unroll
    for (int n = 0; n < 19; ++n) {
        if (ci(n, DIR) != ORI) continue;
        const int p = opp(n);
        if (cx(p) * cx(p) + cy(p) * cy(p) + cz(p) * cz(p) == 1) continue;
#pragma unroll
        for (int iDim = 1; iDim < 3; ++iDim) {
            const int a2 = (DIR + iDim) % 3;
            f[p] += ci(p, a2) * diff[a2];
        }
    }
}

// Regularized completion: boundaryTemplates::compute_PiNeq
// (boundaryCondition/boundaryTemplates.h:42-90) and
// IsoThermalBulkDynamics::regularize (basicDynamics/isoThermalDynamics.hh:55-69).
template<typename T, int DIR, int ORI>
__device__ __forceinline__ void regularized_close(T* f, T rhoBar, const T j[3]) {
    const T jSqr = j[0] * j[0] + j[1] * j[1] + j[2] * j[2];
    const T ir = inv_rho(rhoBar);
    T fNeq[19];
#pragma unroll
    for (int i = 0; i < 19; ++i) {
        fNeq[i] = T();
        if (ci(i, DIR) == 0) fNeq[i] = f[i] - feq(i, rhoBar, ir, j, jSqr);
        else if (ci(i, DIR) == ORI) fNeq[i] = f[i] - feq(i, rhoBar, ir, j, jSqr);
    }
    T pi[6];
    int iPi = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
#pragma unroll
        for (int b = a; b < 3; ++b) {
            pi[iPi] = T();
#pragma unroll
            for (int i = 0; i < 19; ++i)
                if (ci(i, DIR) == 0) pi[iPi] += ci(i, a) * ci(i, b) * fNeq[i];
#pragma unroll
            for (int i = 0; i < 19; ++i)
                if (ci(i, DIR) == ORI) pi[iPi] += (T)2 * ci(i, a) * ci(i, b) * fNeq[i];
            ++iPi;
        }
    }
    f[0] = feq(0, rhoBar, ir, j, jSqr) + pi_to_fneq_generic(0, pi);
#pragma unroll
    for (int i = 1; i <= 9; ++i) {
        f[i] = feq(i, rhoBar, ir, j, jSqr);
        f[i + 9] = feq(i + 9, rhoBar, ir, j, jSqr);
        const T fn = pi_to_fneq_generic(i, pi);
        f[i] += fn;
        f[i + 9] += fn;
    }
}

template<typename T, int DIR, int ORI>
__device__ __forceinline__ void complete_plane_dir(T* f, int family, int bckind, const double* v) {
    T rhoBar, j[3];
    if (bckind == PLBAMD_VELOCITY) {
        const T u[3] = {(T)v[0], (T)v[1], (T)v[2]};
        velocity_moments<T, DIR, ORI>(f, u, rhoBar, j);
    } else {
        rhoBar = (T)v[0] - (T)1;  // StoreDensityDynamics::defineDensity, boundaryDynamics.hh:251-255
        pressure_moments<T, DIR, ORI>(f, rhoBar, j);
    }
    if (family == PLBAMD_ZOUHE) zouhe_close<T, DIR, ORI>(f, rhoBar, j);
    else regularized_close<T, DIR, ORI>(f, rhoBar, j);
}

template<typename T>
__device__ void complete_plane(T* f, uint32_t info, const double* v) {
    const int family = (info >> 2) & 1, bckind = (info >> 3) & 1, dir = (info >> 4) & 3;
    const bool pos = (info >> 6) & 1;
    switch (dir * 2 + (pos ? 1 : 0)) {
    case 0: complete_plane_dir<T, 0, -1>(f, family, bckind, v); break;
    case 1: complete_plane_dir<T, 0, 1>(f, family, bckind, v); break;
    case 2: complete_plane_dir<T, 1, -1>(f, family, bckind, v); break;
    case 3: complete_plane_dir<T, 1, 1>(f, family, bckind, v); break;
    case 4: complete_plane_dir<T, 2, -1>(f, family, bckind, v); break;
    default: complete_plane_dir<T, 2, 1>(f, family, bckind, v); break;
    }
}

// ---------------------------------------------------------------------------
// Special-cell kernel: planes, momentum-exchange nodes, cells next to link-wise
// walls. Same AA addressing as the bulk kernel, with the wall mask added to the
// edge test.

template<typename S, int DYN, bool ODD, bool STATS>
__global__ void __launch_bounds__(256) special_step(SpecialArgs a) {
    using T = typename Store<S>::T;
    S* __restrict__ f = reinterpret_cast<S*>(a.f);
    const Geometry g = a.g;
    const T omega = (T)a.omega;
    double st[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < a.count) {
        const SpecialCell sc = a.cells[k];
        const int x = sc.x, y = sc.y, z = sc.z;
        const uint64_t cell = sc.cell;
        const uint32_t kind = sc.info & 3u;
        T fc[19];
        if constexpr (!ODD) {
#pragma unroll
            for (int i = 0; i < 19; ++i) fc[i] = Store<S>::load(f + (uint64_t)i * g.slot + cell);
        } else {
#pragma unroll
            for (int i = 0; i < 19; ++i) fc[i] = Store<S>::load(f + swapped_index(g, x, y, z, i, sc.wall, cell));
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
            st[4] += (double)(-mom[0]);
            st[5] += (double)(-mom[1]);
            st[6] += (double)(-mom[2]);
#pragma unroll
            for (int i = 1; i <= 9; ++i) { T t = fc[i]; fc[i] = fc[i + 9]; fc[i + 9] = t; }
        } else {
            if (kind == SK_PLANE) complete_plane(fc, sc.info, a.values + sc.value);
            T rhoBar, drag[3];
            const T F0[3] = {0, 0, 0};
            const T uSqr = collide_bulk<T, DYN, false>(fc, omega, F0, (T)0, rhoBar, drag);
            st[0] += (double)rhoBar;
            st[1] += (double)uSqr;
            st[2] = fmax(st[2], (double)uSqr);
            st[3] += 1.0;
            // link-wise wall: each outgoing population towards the wall carries 2 c_i f*_i
            // (the stored, storage-rounded value is what is reflected)
#pragma unroll
            for (int i = 1; i < 19; ++i) {
                if (!((sc.wall >> i) & 1u)) continue;
                const T fo = Store<S>::round(fc[i]);
#pragma unroll
                for (int c = 0; c < 3; ++c) st[4 + c] += (double)((T)2 * ci(i, c) * fo);
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
    if constexpr (STATS) block_reduce_store(st, a.partials + 8 * (uint64_t)blockIdx.x);
}

// Final statistics reduction: one workgroup, fixed order.
__global__ void __launch_bounds__(256) reduce_partials(const double* bulk, uint64_t nBulk,
                                                       const double* special, uint64_t nSpecial,
                                                       double* out) {
    double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (uint64_t b = threadIdx.x; b < nBulk; b += blockDim.x) {
        const double* p = bulk + 8 * b;
        v[0] += p[0]; v[1] += p[1]; v[2] = fmax(v[2], p[2]); v[3] += p[3];
        v[4] += p[4]; v[5] += p[5]; v[6] += p[6];
    }
    for (uint64_t b = threadIdx.x; b < nSpecial; b += blockDim.x) {
        const double* p = special + 8 * b;
        s[0] += p[0]; s[1] += p[1]; s[2] = fmax(s[2], p[2]); s[3] += p[3];
        s[4] += p[4]; s[5] += p[5]; s[6] += p[6];
    }
    // record: sumRho sumU maxU count Fx Fy Fz (force from special) / Px Py Pz (porous from bulk)
    double r[8] = {v[0] + s[0], v[1] + s[1], fmax(v[2], s[2]), v[3] + s[3], s[4], s[5], s[6], 0};
    block_reduce_store(r, out);
    // porous drag sums; slot 2 of a record is a max, so the z sum goes to slot 3
    double q[8] = {v[4], v[5], 0, v[6], 0, 0, 0, 0};
    block_reduce_store(q, out + 8);
}

// ---------------------------------------------------------------------------
// Transfers and analysis (natural-layout view of the current parity)

template<typename S>
__device__ __forceinline__ uint64_t pop_index(const Geometry& g, int parity, int x, int y, int z, int i,
                                              uint32_t wall, uint64_t cell) {
    if (parity == 0) return (uint64_t)i * g.slot + cell;
    return swapped_index(g, x, y, z, i, wall, cell);
}

struct BoxArgs {
    Geometry g;
    int parity;
    int x0, y0, z0, lx, ly, lz;
    const SpecialCell* special;   // sorted by cell, to find wall masks
    uint32_t nSpecial;
};

__device__ __forceinline__ uint32_t wall_mask_of(const BoxArgs& b, uint64_t cell) {
    uint32_t lo = 0, hi = b.nSpecial;
    while (lo < hi) {
        uint32_t mid = (lo + hi) / 2;
        if (b.special[mid].cell < cell) lo = mid + 1; else hi = mid;
    }
    if (lo < b.nSpecial && b.special[lo].cell == cell) return b.special[lo].wall;
    return 0;
}

template<typename S, typename H>
__global__ void upload_box(void* fv, BoxArgs b, const H* host) {
    using T = typename Store<S>::T;
    S* f = reinterpret_cast<S*>(fv);
    const uint64_t n = (uint64_t)b.lx * b.ly * b.lz;
    for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < n; k += (uint64_t)gridDim.x * blockDim.x) {
        const int lz = k % b.lz, ly = (k / b.lz) % b.ly, lx = k / ((uint64_t)b.lz * b.ly);
        const int x = b.x0 + lx, y = b.y0 + ly, z = b.z0 + lz;
        const uint64_t cell = lin(b.g, x, y, z);
        const uint32_t wall = b.parity ? wall_mask_of(b, cell) : 0;
        for (int i = 0; i < 19; ++i) {
            const T v = (T)host[k * 19 + i];
            Store<S>::store(f + pop_index<S>(b.g, b.parity, x, y, z, i, wall, cell), v);
        }
    }
}

// Raw 16-bit upload/download (bit patterns equal to the storage type).
template<typename S>
__global__ void upload_box_raw(void* fv, BoxArgs b, const uint16_t* host) {
    uint16_t* f = reinterpret_cast<uint16_t*>(fv);
    const uint64_t n = (uint64_t)b.lx * b.ly * b.lz;
    for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < n; k += (uint64_t)gridDim.x * blockDim.x) {
        const int lz = k % b.lz, ly = (k / b.lz) % b.ly, lx = k / ((uint64_t)b.lz * b.ly);
        const int x = b.x0 + lx, y = b.y0 + ly, z = b.z0 + lz;
        const uint64_t cell = lin(b.g, x, y, z);
        const uint32_t wall = b.parity ? wall_mask_of(b, cell) : 0;
        for (int i = 0; i < 19; ++i) f[pop_index<S>(b.g, b.parity, x, y, z, i, wall, cell)] = host[k * 19 + i];
    }
}

template<typename S>
__global__ void download_box_raw(const void* fv, BoxArgs b, uint16_t* host) {
    const uint16_t* f = reinterpret_cast<const uint16_t*>(fv);
    const uint64_t n = (uint64_t)b.lx * b.ly * b.lz;
    for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < n; k += (uint64_t)gridDim.x * blockDim.x) {
        const int lz = k % b.lz, ly = (k / b.lz) % b.ly, lx = k / ((uint64_t)b.lz * b.ly);
        const int x = b.x0 + lx, y = b.y0 + ly, z = b.z0 + lz;
        const uint64_t cell = lin(b.g, x, y, z);
        const uint32_t wall = b.parity ? wall_mask_of(b, cell) : 0;
        for (int i = 0; i < 19; ++i) host[k * 19 + i] = f[pop_index<S>(b.g, b.parity, x, y, z, i, wall, cell)];
    }
}

template<typename S, typename H>
__global__ void download_box(const void* fv, BoxArgs b, H* host) {
    const S* f = reinterpret_cast<const S*>(fv);
    const uint64_t n = (uint64_t)b.lx * b.ly * b.lz;
    for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < n; k += (uint64_t)gridDim.x * blockDim.x) {
        const int lz = k % b.lz, ly = (k / b.lz) % b.ly, lx = k / ((uint64_t)b.lz * b.ly);
        const int x = b.x0 + lx, y = b.y0 + ly, z = b.z0 + lz;
        const uint64_t cell = lin(b.g, x, y, z);
        const uint32_t wall = b.parity ? wall_mask_of(b, cell) : 0;
        for (int i = 0; i < 19; ++i)
            host[k * 19 + i] = (H)Store<S>::load(f + pop_index<S>(b.g, b.parity, x, y, z, i, wall, cell));
    }
}

// Completed snapshot of wrapped-factory plane cells inside a downloaded box:
// the populations the wrapped post-stream processor leaves in Palabos's lattice.
template<typename S, typename H>
__global__ void complete_download(const void* fv, BoxArgs b, const double* values, H* host, int raw) {
    using T = typename Store<S>::T;
    const S* f = reinterpret_cast<const S*>(fv);
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < b.nSpecial; k += gridDim.x * blockDim.x) {
        const SpecialCell sc = b.special[k];
        if ((sc.info & 3u) != SK_PLANE || !((sc.info >> 7) & 1u)) continue;
        const int x = sc.x, y = sc.y, z = sc.z;
        if (x < b.x0 || x >= b.x0 + b.lx || y < b.y0 || y >= b.y0 + b.ly || z < b.z0 || z >= b.z0 + b.lz) continue;
        T fc[19];
        for (int i = 0; i < 19; ++i)
            fc[i] = Store<S>::load(f + pop_index<S>(b.g, b.parity, x, y, z, i, b.parity ? sc.wall : 0u, sc.cell));
        complete_plane(fc, sc.info, values + sc.value);
        const uint64_t o = (((uint64_t)(x - b.x0) * b.ly + (y - b.y0)) * b.lz + (z - b.z0)) * 19;
        for (int i = 0; i < 19; ++i) {
            if (raw) reinterpret_cast<uint16_t*>(host)[o + i] = Store<S>::bits(fc[i]);
            else host[o + i] = (H)Store<S>::round(fc[i]);
        }
    }
}

// Equilibrium initialisation (core/cell.hh:124-132 iniCellAtEquilibrium) in the
// compute precision of the lattice.
template<typename S>
__global__ void equilibrium_box(void* fv, BoxArgs b, const double* rho, const double* u) {
    using T = typename Store<S>::T;
    S* f = reinterpret_cast<S*>(fv);
    const uint64_t n = (uint64_t)b.lx * b.ly * b.lz;
    for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < n; k += (uint64_t)gridDim.x * blockDim.x) {
        const int lz = k % b.lz, ly = (k / b.lz) % b.ly, lx = k / ((uint64_t)b.lz * b.ly);
        const int x = b.x0 + lx, y = b.y0 + ly, z = b.z0 + lz;
        const uint64_t cell = lin(b.g, x, y, z);
        const uint32_t wall = b.parity ? wall_mask_of(b, cell) : 0;
        const T density = (T)rho[k];
        const T j[3] = {(T)u[3 * k] * density, (T)u[3 * k + 1] * density, (T)u[3 * k + 2] * density};
        const T jSqr = j[0] * j[0] + j[1] * j[1] + j[2] * j[2];
        const T rhoBar = density - (T)1;
        const T ir = inv_rho(rhoBar);
        for (int i = 0; i < 19; ++i)
            Store<S>::store(f + pop_index<S>(b.g, b.parity, x, y, z, i, wall, cell), feq(i, rhoBar, ir, j, jSqr));
    }
}

// Per-cell density and velocity as Palabos's analysis functions see them
// (Dynamics::computeRhoBar / computeVelocity of the cell's dynamics):
//   bulk            rhoBar = plain sum (momentTemplates3D.h:375-382), u = j / rho
//                   (compute_uLb, :425-432), plus F/2 with a body force
//                   (externalForceDynamics.hh:59-74)
//   velocity plane  rhoBar from the Dirichlet formula, u = imposed velocity
//                   (boundaryDynamics.hh:643-671, :324-330)
//   pressure plane  rhoBar imposed, u = j(rhoBar) / rho (boundaryDynamics.hh:776-835)
//   solid nodes     rhoBar = 0 (rho = 1), u = 0 (core/dynamics.hh:824-829,1004-1058;
//                   bounceBackModels.hh:170-231)
template<typename T, int DIR, int ORI>
__device__ __forceinline__ void plane_macroscopic(const T* f, int bckind, const double* v, T& rhoBar, T u[3]) {
    T j[3];
    if (bckind == PLBAMD_VELOCITY) {
        const T uu[3] = {(T)v[0], (T)v[1], (T)v[2]};
        velocity_moments<T, DIR, ORI>(f, uu, rhoBar, j);
        u[0] = uu[0]; u[1] = uu[1]; u[2] = uu[2];
    } else {
        rhoBar = (T)v[0] - (T)1;
        pressure_moments<T, DIR, ORI>(f, rhoBar, j);
        const T ir = inv_rho(rhoBar);
        u[0] = j[0] * ir; u[1] = j[1] * ir; u[2] = j[2] * ir;
    }
}

template<typename T>
__device__ void plane_macroscopic_any(const T* f, uint32_t info, const double* v, T& rhoBar, T u[3]) {
    const int bckind = (info >> 3) & 1, dir = (info >> 4) & 3;
    const bool pos = (info >> 6) & 1;
    switch (dir * 2 + (pos ? 1 : 0)) {
    case 0: plane_macroscopic<T, 0, -1>(f, bckind, v, rhoBar, u); break;
    case 1: plane_macroscopic<T, 0, 1>(f, bckind, v, rhoBar, u); break;
    case 2: plane_macroscopic<T, 1, -1>(f, bckind, v, rhoBar, u); break;
    case 3: plane_macroscopic<T, 1, 1>(f, bckind, v, rhoBar, u); break;
    case 4: plane_macroscopic<T, 2, -1>(f, bckind, v, rhoBar, u); break;
    default: plane_macroscopic<T, 2, 1>(f, bckind, v, rhoBar, u); break;
    }
}

struct MomentsArgs {
    const uint8_t* tags;
    const double* values;
    double force[3];
};

template<typename S>
__global__ void __launch_bounds__(256) moments_box(const void* fv, BoxArgs b, MomentsArgs m, double* rhoOut,
                                                   double* uOut, double* partials) {
    using T = typename Store<S>::T;
    const S* f = reinterpret_cast<const S*>(fv);
    const uint64_t n = (uint64_t)b.lx * b.ly * b.lz;
    double st[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < n; k += (uint64_t)gridDim.x * blockDim.x) {
        const int lz = k % b.lz, ly = (k / b.lz) % b.ly, lx = k / ((uint64_t)b.lz * b.ly);
        const int x = b.x0 + lx, y = b.y0 + ly, z = b.z0 + lz;
        const uint64_t cell = lin(b.g, x, y, z);
        const uint8_t tag = m.tags ? m.tags[cell] : (uint8_t)BT_BULK;
        // locate the special record of the cell, if any
        int32_t sp = -1;
        if (tag == BT_SKIP) {
            uint32_t lo = 0, hi = b.nSpecial;
            while (lo < hi) {
                uint32_t mid = (lo + hi) / 2;
                if (b.special[mid].cell < cell) lo = mid + 1; else hi = mid;
            }
            if (lo < b.nSpecial && b.special[lo].cell == cell) sp = (int32_t)lo;
        }
        const uint32_t wall = (b.parity && sp >= 0) ? b.special[sp].wall : 0u;
        T rhoBar = T(), u[3] = {T(), T(), T()};
        const bool solid = tag == BT_BOUNCE_BACK || tag == BT_NO_DYNAMICS ||
                           (tag == BT_SKIP && (sp < 0 || (b.special[sp].info & 3u) == SK_MOMENTUM_EXCHANGE));
        if (!solid) {
            T fc[19];
            for (int i = 0; i < 19; ++i) fc[i] = Store<S>::load(f + pop_index<S>(b.g, b.parity, x, y, z, i, wall, cell));
            if (sp >= 0 && (b.special[sp].info & 3u) == SK_PLANE) {
                plane_macroscopic_any(fc, b.special[sp].info, m.values + b.special[sp].value, rhoBar, u);
            } else {
                T rb, j[3];
                rhobar_j(fc, rb, j);
                rhoBar = rhobar_sum(fc);
                const T ir = inv_rho(rhoBar);
                for (int c = 0; c < 3; ++c) u[c] = j[c] * ir + (T)m.force[c] / (T)2;
            }
        }
        if (rhoOut) rhoOut[k] = (double)(rhoBar + (T)1);
        if (uOut) { uOut[3 * k] = (double)u[0]; uOut[3 * k + 1] = (double)u[1]; uOut[3 * k + 2] = (double)u[2]; }
        const T uSqr = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
        st[0] += (double)rhoBar;
        st[1] += (double)uSqr;
        st[3] += (double)u[0];
        st[4] += (double)u[1];
        st[5] += (double)u[2];
    }
    if (partials) block_reduce_store(st, partials + 8 * (uint64_t)blockIdx.x);
}

__global__ void sum_records(const double* p, uint32_t n, double* out) {
    double r[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (uint32_t b = threadIdx.x; b < n; b += blockDim.x)
        for (int k = 0; k < 8; ++k) if (k != 2) r[k] += p[8 * b + k];
    block_reduce_store(r, out);
}

// ---------------------------------------------------------------------------
// Host runtime

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

struct Plane {
    int family, kind, dir, ori;
    bool wrapped;
    plbamd_box box;
    std::vector<double> values;
    bool valuesSet = false;
};

struct Lattice {
    plbamd_lattice_desc desc;
    Geometry g;
    uint64_t cells = 0;
    size_t elemSize = 0;
    void* f = nullptr;
    int64_t time = 0;
    bool statsOn = true;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;

    std::vector<uint8_t> hostTags;    // PLBAMD_TAG_* or empty
    std::vector<Plane> planes;
    std::vector<double> hostAlpha;

    bool dirty = true;                // device tag / special structures need a rebuild
    bool anyWrapped = false;          // some plane stands for a wrapped factory
    uint8_t* dTags = nullptr;
    SpecialCell* dSpecial = nullptr;
    uint32_t nSpecial = 0;
    double* dValues = nullptr;
    double* dAlpha = nullptr;
    double* dBulkPartials = nullptr;
    uint64_t nBulkPartials = 0;
    double* dSpecialPartials = nullptr;
    uint64_t nSpecialPartials = 0;
    double* dStats = nullptr;         // 16 doubles
    bool statsValid = false;

    dim3 bulkBlock, bulkGrid;
    int zChunk = 0, nzChunks = 1;

    ~Lattice() {
        hipFree(f); hipFree(dTags); hipFree(dSpecial); hipFree(dValues); hipFree(dAlpha);
        hipFree(dBulkPartials); hipFree(dSpecialPartials); hipFree(dStats);
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
           b.x1 < L.g.nx && b.y1 < L.g.ny && b.z1 < L.g.nz;
}

uint64_t boxCells(const plbamd_box& b) {
    return (uint64_t)(b.x1 - b.x0 + 1) * (b.y1 - b.y0 + 1) * (b.z1 - b.z0 + 1);
}

// Builds the device tag array and the special-cell list from the host-side
// description (tags, planes, walls).
int rebuild(Lattice& L) {
    if (!L.dirty) return PLBAMD_OK;
    const int nx = L.g.nx, ny = L.g.ny, nz = L.g.nz;
    auto at = [&](int x, int y, int z) { return ((uint64_t)x * ny + y) * nz + z; };
    std::vector<int> planeOf;
    std::vector<uint32_t> boxIndex;
    bool anyPlane = !L.planes.empty();
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
    auto tagOf = [&](int x, int y, int z) -> uint8_t { return haveTags ? L.hostTags[at(x, y, z)] : 0; };
    auto wrap = [&](int& v, int n, int per) {
        if (v < 0 || v >= n) { if (!per) return false; v = (v + n) % n; }
        return true;
    };
    std::vector<uint8_t> dev(L.cells, BT_BULK);
    std::vector<SpecialCell> special;
    std::vector<uint32_t> planeValueBase(L.planes.size(), 0);
    std::vector<double> values;
    for (size_t p = 0; p < L.planes.size(); ++p) {
        planeValueBase[p] = (uint32_t)values.size();
        values.insert(values.end(), L.planes[p].values.begin(), L.planes[p].values.end());
    }
    bool anyNonBulk = anyPlane;
    L.anyWrapped = false;
    for (int x = 0; x < nx; ++x)
        for (int y = 0; y < ny; ++y)
            for (int z = 0; z < nz; ++z) {
                const uint64_t c = at(x, y, z);
                const uint8_t t = tagOf(x, y, z);
                const int p = anyPlane ? planeOf[c] : -1;
                if (t == PLBAMD_TAG_WALL_LINK) { dev[c] = BT_SKIP; anyNonBulk = true; continue; }
                uint32_t wall = 0, fluid = 0;
                if (haveTags) {
                    for (int i = 1; i < 19; ++i) {
                        int ax = x + cx(i), ay = y + cy(i), az = z + cz(i);
                        bool in = wrap(ax, nx, L.g.px) && wrap(ay, ny, L.g.py) && wrap(az, nz, L.g.pz);
                        uint8_t nt = in ? tagOf(ax, ay, az) : PLBAMD_TAG_BULK;
                        if (in && nt == PLBAMD_TAG_WALL_LINK) wall |= 1u << i;
                        // bounceBackModels3D.hh:48-65: fluid unless BB / MEBB / NoDynamics;
                        // beyond a non-periodic edge the envelope holds the background dynamics.
                        if (nt != PLBAMD_TAG_MOMENTUM_EXCHANGE && nt != PLBAMD_TAG_BOUNCE_BACK &&
                            nt != PLBAMD_TAG_NO_DYNAMICS && nt != PLBAMD_TAG_WALL_LINK)
                            fluid |= 1u << i;
                    }
                }
                SpecialCell s{};
                s.cell = (uint32_t)c;
                s.x = (uint16_t)x; s.y = (uint16_t)y; s.z = (uint16_t)z;
                s.wall = wall;
                if (p >= 0) {
                    const Plane& P = L.planes[p];
                    s.info = SK_PLANE | (P.family << 2) | (P.kind << 3) | (P.dir << 4) | ((P.ori > 0 ? 1 : 0) << 6) |
                             ((P.wrapped ? 1 : 0) << 7);
                    if (P.wrapped) L.anyWrapped = true;
                    s.value = planeValueBase[p] + boxIndex[c] * (P.kind == PLBAMD_VELOCITY ? 3u : 1u);
                    dev[c] = BT_SKIP;
                    special.push_back(s);
                } else if (t == PLBAMD_TAG_MOMENTUM_EXCHANGE) {
                    s.info = SK_MOMENTUM_EXCHANGE;
                    s.fluid = fluid;
                    dev[c] = BT_SKIP;
                    special.push_back(s);
                    anyNonBulk = true;
                } else if (t == PLBAMD_TAG_BOUNCE_BACK || t == PLBAMD_TAG_NO_DYNAMICS) {
                    dev[c] = (t == PLBAMD_TAG_BOUNCE_BACK) ? BT_BOUNCE_BACK : BT_NO_DYNAMICS;
                    anyNonBulk = true;
                    if (wall) {  // reflecting node next to a link wall: handled as near-wall bulk is wrong;
                                 // such cells keep node semantics and ignore the wall
                    }
                } else if (wall) {
                    s.info = SK_NEAR_WALL;
                    dev[c] = BT_SKIP;
                    special.push_back(s);
                }
            }
    hipFree(L.dTags); L.dTags = nullptr;
    hipFree(L.dSpecial); L.dSpecial = nullptr;
    hipFree(L.dValues); L.dValues = nullptr;
    hipFree(L.dSpecialPartials); L.dSpecialPartials = nullptr;
    L.nSpecial = (uint32_t)special.size();
    if (anyNonBulk) {
        PLBAMD_HIP(hipMalloc(&L.dTags, L.cells));
        PLBAMD_HIP(hipMemcpy(L.dTags, dev.data(), L.cells, hipMemcpyHostToDevice));
    }
    if (!special.empty()) {
        PLBAMD_HIP(hipMalloc(&L.dSpecial, special.size() * sizeof(SpecialCell)));
        PLBAMD_HIP(hipMemcpy(L.dSpecial, special.data(), special.size() * sizeof(SpecialCell), hipMemcpyHostToDevice));
        L.nSpecialPartials = (special.size() + 255) / 256;
        PLBAMD_HIP(hipMalloc(&L.dSpecialPartials, L.nSpecialPartials * 8 * sizeof(double)));
    }
    if (!values.empty()) {
        PLBAMD_HIP(hipMalloc(&L.dValues, values.size() * sizeof(double)));
        PLBAMD_HIP(hipMemcpy(L.dValues, values.data(), values.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    L.dirty = false;
    return PLBAMD_OK;
}

template<typename S, int DYN>
int launchStepTyped(Lattice& L, bool odd, bool stats) {
    StepArgs a{};
    a.f = L.f;
    a.g = L.g;
    a.omega = L.desc.omega;
    for (int c = 0; c < 3; ++c) a.force[c] = L.desc.force[c];
    a.tags = L.dTags;
    a.alpha = L.dAlpha;
    a.partials = L.dBulkPartials;
    a.zChunk = L.zChunk;
    a.nzChunks = L.nzChunks;
    const bool forced = L.dAlpha || L.desc.force[0] != 0 || L.desc.force[1] != 0 || L.desc.force[2] != 0;
    const bool tags = L.dTags != nullptr;
    // 2 x 2 x 2 x 2 dispatch over (odd, tags, stats, forced)
#define PLBAMD_BULK(O, TG, ST, FO) \
    hipLaunchKernelGGL((bulk_step<S, DYN, O, TG, ST, FO>), L.bulkGrid, L.bulkBlock, 0, L.stream, a)
#define PLBAMD_BULK_F(O, TG, ST) if (forced) PLBAMD_BULK(O, TG, ST, true); else PLBAMD_BULK(O, TG, ST, false)
#define PLBAMD_BULK_S(O, TG) if (stats) { PLBAMD_BULK_F(O, TG, true); } else { PLBAMD_BULK_F(O, TG, false); }
#define PLBAMD_BULK_T(O) if (tags) { PLBAMD_BULK_S(O, true) } else { PLBAMD_BULK_S(O, false) }
    if (odd) { PLBAMD_BULK_T(true) } else { PLBAMD_BULK_T(false) }
#undef PLBAMD_BULK_T
#undef PLBAMD_BULK_S
#undef PLBAMD_BULK_F
#undef PLBAMD_BULK
    PLBAMD_HIP(hipGetLastError());
    if (L.nSpecial) {
        SpecialArgs s{};
        s.f = L.f; s.g = L.g; s.omega = L.desc.omega;
        s.cells = L.dSpecial; s.count = L.nSpecial; s.values = L.dValues;
        s.partials = L.dSpecialPartials;
        dim3 grid((L.nSpecial + 255) / 256), block(256);
        if (odd) {
            if (stats) hipLaunchKernelGGL((special_step<S, DYN, true, true>), grid, block, 0, L.stream, s);
            else hipLaunchKernelGGL((special_step<S, DYN, true, false>), grid, block, 0, L.stream, s);
        } else {
            if (stats) hipLaunchKernelGGL((special_step<S, DYN, false, true>), grid, block, 0, L.stream, s);
            else hipLaunchKernelGGL((special_step<S, DYN, false, false>), grid, block, 0, L.stream, s);
        }
        PLBAMD_HIP(hipGetLastError());
    }
    if (stats) {
        hipLaunchKernelGGL(reduce_partials, dim3(1), dim3(256), 0, L.stream, L.dBulkPartials, L.nBulkPartials,
                           L.dSpecialPartials, L.nSpecial ? L.nSpecialPartials : 0, L.dStats);
        PLBAMD_HIP(hipGetLastError());
    }
    return PLBAMD_OK;
}

template<typename S>
int launchStep(Lattice& L, bool odd, bool stats) {
    return L.desc.dynamics == PLBAMD_BGK ? launchStepTyped<S, 0>(L, odd, stats)
                                         : launchStepTyped<S, 1>(L, odd, stats);
}

int step(Lattice& L, int nsteps) {
    int rc = rebuild(L);
    if (rc) return rc;
    for (int s = 0; s < nsteps; ++s) {
        const bool odd = (L.time & 1) != 0;
        const bool stats = L.statsOn && s == nsteps - 1;
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

BoxArgs boxArgs(const Lattice& L, const plbamd_box& b) {
    BoxArgs a{};
    a.g = L.g;
    a.parity = (int)(L.time & 1);
    a.x0 = b.x0; a.y0 = b.y0; a.z0 = b.z0;
    a.lx = b.x1 - b.x0 + 1; a.ly = b.y1 - b.y0 + 1; a.lz = b.z1 - b.z0 + 1;
    a.special = L.dSpecial;
    a.nSpecial = L.nSpecial;
    return a;
}

dim3 gridFor(uint64_t n) {
    uint64_t b = (n + 255) / 256;
    if (b > 65536) b = 65536;
    if (b == 0) b = 1;
    return dim3((unsigned)b);
}

template<typename S>
int transfer(Lattice& L, const plbamd_box& b, int elem, void* data, size_t bytes, bool up) {
    const uint64_t n = boxCells(b);
    const size_t hostElem = (elem == PLBAMD_F64) ? 8 : (elem == PLBAMD_F32 ? 4 : 2);
    if (bytes != n * 19 * hostElem) return fail(PLBAMD_EINVAL, "transfer: byte count does not match the box");
    if ((elem == PLBAMD_F16 || elem == PLBAMD_BF16) && elem != L.desc.storage)
        return fail(PLBAMD_EINVAL, "transfer: 16-bit host data must match the storage type");
    int rc = rebuild(L);
    if (rc) return rc;
    void* staging = nullptr;
    PLBAMD_HIP(hipMalloc(&staging, bytes));
    BoxArgs a = boxArgs(L, b);
    dim3 grid = gridFor(n), block(256);
    hipError_t err = hipSuccess;
    if (up) {
        err = hipMemcpyAsync(staging, data, bytes, hipMemcpyHostToDevice, L.stream);
        if (err == hipSuccess) {
            if (elem == PLBAMD_F64) hipLaunchKernelGGL((upload_box<S, double>), grid, block, 0, L.stream, L.f, a, (const double*)staging);
            else if (elem == PLBAMD_F32) hipLaunchKernelGGL((upload_box<S, float>), grid, block, 0, L.stream, L.f, a, (const float*)staging);
            else hipLaunchKernelGGL((upload_box_raw<S>), grid, block, 0, L.stream, L.f, a, (const uint16_t*)staging);
            err = hipGetLastError();
        }
    } else {
        if (elem == PLBAMD_F64) hipLaunchKernelGGL((download_box<S, double>), grid, block, 0, L.stream, L.f, a, (double*)staging);
        else if (elem == PLBAMD_F32) hipLaunchKernelGGL((download_box<S, float>), grid, block, 0, L.stream, L.f, a, (float*)staging);
        else hipLaunchKernelGGL((download_box_raw<S>), grid, block, 0, L.stream, L.f, a, (uint16_t*)staging);
        err = hipGetLastError();
        if (err == hipSuccess && L.anyWrapped && L.nSpecial) {
            dim3 g2 = gridFor(L.nSpecial);
            if (elem == PLBAMD_F64)
                hipLaunchKernelGGL((complete_download<S, double>), g2, block, 0, L.stream, L.f, a, L.dValues, (double*)staging, 0);
            else if (elem == PLBAMD_F32)
                hipLaunchKernelGGL((complete_download<S, float>), g2, block, 0, L.stream, L.f, a, L.dValues, (float*)staging, 0);
            else
                hipLaunchKernelGGL((complete_download<S, float>), g2, block, 0, L.stream, L.f, a, L.dValues, (float*)staging, 1);
            err = hipGetLastError();
        }
        if (err == hipSuccess) err = hipMemcpyAsync(data, staging, bytes, hipMemcpyDeviceToHost, L.stream);
    }
    if (err == hipSuccess) err = hipStreamSynchronize(L.stream);
    hipFree(staging);
    if (err != hipSuccess) return fail(PLBAMD_EHIP, std::string("transfer: ") + hipGetErrorString(err));
    return PLBAMD_OK;
}

template<typename S>
int equilibrium(Lattice& L, const plbamd_box& b, const double* rho, const double* u) {
    const uint64_t n = boxCells(b);
    int rc = rebuild(L);
    if (rc) return rc;
    double* d = nullptr;
    PLBAMD_HIP(hipMalloc(&d, n * 4 * sizeof(double)));
    hipError_t err = hipMemcpyAsync(d, rho, n * sizeof(double), hipMemcpyHostToDevice, L.stream);
    if (err == hipSuccess) err = hipMemcpyAsync(d + n, u, 3 * n * sizeof(double), hipMemcpyHostToDevice, L.stream);
    if (err == hipSuccess) {
        hipLaunchKernelGGL((equilibrium_box<S>), gridFor(n), dim3(256), 0, L.stream, L.f, boxArgs(L, b), d, d + n);
        err = hipGetLastError();
    }
    if (err == hipSuccess) err = hipStreamSynchronize(L.stream);
    hipFree(d);
    if (err != hipSuccess) return fail(PLBAMD_EHIP, std::string("equilibrium: ") + hipGetErrorString(err));
    return PLBAMD_OK;
}

template<typename S>
int moments(Lattice& L, const plbamd_box& b, double* rho, double* u, double* sums) {
    const uint64_t n = boxCells(b);
    int rc = rebuild(L);
    if (rc) return rc;
    dim3 grid = gridFor(n);
    double* d = nullptr;
    const size_t fieldBytes = (rho || u) ? n * 4 * sizeof(double) : 0;
    const size_t partBytes = (size_t)grid.x * 8 * sizeof(double) + 8 * sizeof(double);
    PLBAMD_HIP(hipMalloc(&d, fieldBytes + partBytes));
    double* dRho = rho ? d : nullptr;
    double* dU = u ? d + n : nullptr;
    double* dPart = d + fieldBytes / sizeof(double);
    double* dSum = dPart + (size_t)grid.x * 8;
    MomentsArgs m{};
    m.tags = L.dTags;
    m.values = L.dValues;
    for (int c = 0; c < 3; ++c) m.force[c] = L.desc.force[c];
    hipLaunchKernelGGL((moments_box<S>), grid, dim3(256), 0, L.stream, L.f, boxArgs(L, b), m, dRho, dU, dPart);
    hipError_t err = hipGetLastError();
    if (err == hipSuccess) {
        hipLaunchKernelGGL(sum_records, dim3(1), dim3(256), 0, L.stream, dPart, grid.x, dSum);
        err = hipGetLastError();
    }
    double host[8];
    if (err == hipSuccess && rho) err = hipMemcpyAsync(rho, dRho, n * sizeof(double), hipMemcpyDeviceToHost, L.stream);
    if (err == hipSuccess && u) err = hipMemcpyAsync(u, dU, 3 * n * sizeof(double), hipMemcpyDeviceToHost, L.stream);
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

size_t storageSize(int storage) {
    return storage == PLBAMD_F64 ? 8 : (storage == PLBAMD_F32 ? 4 : 2);
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
    if (desc->nx <= 0 || desc->ny <= 0 || desc->nz <= 0 || desc->nx > 65535 || desc->ny > 65535 || desc->nz > 65535)
        return fail(PLBAMD_EINVAL, "lattice_create: extents must be in [1, 65535]");
    if (desc->storage < PLBAMD_F64 || desc->storage > PLBAMD_BF16)
        return fail(PLBAMD_EINVAL, "lattice_create: unknown storage type");
    if (desc->dynamics != PLBAMD_BGK && desc->dynamics != PLBAMD_RBGK)
        return fail(PLBAMD_EINVAL, "lattice_create: unknown dynamics");
    if (!(desc->omega > 0.0 && desc->omega < 2.0)) return fail(PLBAMD_EINVAL, "lattice_create: omega must lie in (0, 2)");
    const uint64_t cells = (uint64_t)desc->nx * desc->ny * desc->nz;
    if (cells >= (1ull << 32)) return fail(PLBAMD_EINVAL, "lattice_create: at most 2^32 - 1 cells");
    auto L = std::make_unique<Lattice>();
    L->desc = *desc;
    L->g.nx = desc->nx; L->g.ny = desc->ny; L->g.nz = desc->nz;
    L->g.px = desc->periodic[0] != 0; L->g.py = desc->periodic[1] != 0; L->g.pz = desc->periodic[2] != 0;
    L->cells = cells;
    L->g.slot = (cells + 63) / 64 * 64;
    L->elemSize = storageSize(desc->storage);
    const size_t bytes = L->g.slot * 19 * L->elemSize;
    if (hipMalloc(&L->f, bytes) != hipSuccess) {
        hipGetLastError();
        return fail(PLBAMD_ENOMEM, "lattice_create: cannot allocate " + std::to_string(bytes) + " bytes");
    }
    PLBAMD_HIP(hipMemset(L->f, 0, bytes));
    PLBAMD_HIP(hipStreamCreateWithFlags(&L->stream, hipStreamNonBlocking));
    PLBAMD_HIP(hipEventCreate(&L->ev0));
    PLBAMD_HIP(hipEventCreate(&L->ev1));
    // Bulk launch shape: TX threads along z (a multiple of the 64-lane wavefront),
    // each covering CPT cells of a row segment; 256/TX... rows in y per workgroup.
    int txMax = 128, cpt = 1;
    if (const char* e = getenv("PLBAMD_BULK_TX")) txMax = std::max(64, std::min(256, atoi(e) / 64 * 64));
    if (const char* e = getenv("PLBAMD_BULK_CPT")) cpt = std::max(1, std::min(16, atoi(e)));
    int tx = (int)std::min<int64_t>(txMax, ((int64_t)desc->nz + 63) / 64 * 64);
    int ty = std::max(1, txMax / tx);
    L->zChunk = tx * cpt;
    L->nzChunks = (desc->nz + L->zChunk - 1) / L->zChunk;
    L->bulkBlock = dim3(tx, ty);
    L->bulkGrid = dim3((unsigned)(L->nzChunks * ((desc->ny + ty - 1) / ty)), desc->nx);
    L->nBulkPartials = (uint64_t)L->bulkGrid.x * L->bulkGrid.y;
    PLBAMD_HIP(hipMalloc(&L->dBulkPartials, L->nBulkPartials * 8 * sizeof(double)));
    PLBAMD_HIP(hipMalloc(&L->dStats, 16 * sizeof(double)));
    PLBAMD_HIP(hipMemset(L->dStats, 0, 16 * sizeof(double)));
    std::lock_guard<std::mutex> lock(g_mutex);
    *handle = g_nextHandle++;
    g_lattices[*handle] = std::move(L);
    return PLBAMD_OK;
}

int plbamd_lattice_destroy(int handle) {
    std::lock_guard<std::mutex> lock(g_mutex);
    auto it = g_lattices.find(handle);
    if (it == g_lattices.end()) return fail(PLBAMD_EINVAL, "lattice_destroy: unknown handle");
    hipStreamSynchronize(it->second->stream);
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
    L->dirty = true;
    return PLBAMD_OK;
}

int plbamd_set_porosity(int handle, const double* alpha) {
    Lattice* L = find(handle);
    if (!L) return fail(PLBAMD_EINVAL, "set_porosity: unknown handle");
    hipFree(L->dAlpha);
    L->dAlpha = nullptr;
    if (!alpha) return PLBAMD_OK;
    PLBAMD_HIP(hipMalloc(&L->dAlpha, L->cells * sizeof(double)));
    PLBAMD_HIP(hipMemcpy(L->dAlpha, alpha, L->cells * sizeof(double), hipMemcpyHostToDevice));
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
    int rc = step(*L, nsteps);
    if (rc) return rc;
    PLBAMD_HIP(hipStreamSynchronize(L->stream));
    return PLBAMD_OK;
}

int plbamd_collide_and_stream_timed(int handle, int nsteps, double* ms) {
    Lattice* L = find(handle);
    if (!L) return fail(PLBAMD_EINVAL, "collide_and_stream_timed: unknown handle");
    if (nsteps < 0 || !ms) return fail(PLBAMD_EINVAL, "collide_and_stream_timed: bad argument");
    int rc = rebuild(*L);
    if (rc) return rc;
    PLBAMD_HIP(hipEventRecord(L->ev0, L->stream));
    rc = step(*L, nsteps);
    if (rc) return rc;
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

static int dispatchMoments(int handle, const plbamd_box* box, double* rho, double* u, double* sums) {
    Lattice* L = find(handle);
    if (!L) return fail(PLBAMD_EINVAL, "moments: unknown handle");
    if (!box || !boxValid(*L, *box)) return fail(PLBAMD_EINVAL, "moments: box outside the lattice");
    switch (L->desc.storage) {
    case PLBAMD_F64: return moments<double>(*L, *box, rho, u, sums);
    case PLBAMD_F32: return moments<float>(*L, *box, rho, u, sums);
    case PLBAMD_F16: return moments<f16s>(*L, *box, rho, u, sums);
    default: return moments<bf16s>(*L, *box, rho, u, sums);
    }
}

int plbamd_compute_macroscopic(int handle, const plbamd_box* box, double* rho, double* u) {
    if (!rho && !u) return fail(PLBAMD_EINVAL, "compute_macroscopic: no output");
    return dispatchMoments(handle, box, rho, u, nullptr);
}

int plbamd_reduce_moments(int handle, const plbamd_box* box, double* out) {
    if (!out) return fail(PLBAMD_EINVAL, "reduce_moments: null output");
    return dispatchMoments(handle, box, nullptr, nullptr, out);
}

int plbamd_get_time(int handle, int64_t* t) {
    Lattice* L = find(handle);
    if (!L || !t) return fail(PLBAMD_EINVAL, "get_time: bad argument");
    *t = L->time;
    return PLBAMD_OK;
}

}  // extern "C"
