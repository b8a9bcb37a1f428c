// Discrete adjoint of the collide-and-stream step (BGK or regularized BGK, an
// optional Guo body force, a Brinkman porosity term), for gradients of a
// time-averaged objective with respect to the porosity field and the scale of
// each velocity plane. No Palabos counterpart.
//
// Forward step t (state f^t: post-stream populations, natural layout):
//   f*(y)      = C(f^t(y); alpha(y), planes)      collision (plane cells:
//                                                 completion, then BGK)
//   f^{t+1}_i(x) = f*_i(x - c_i)  or  f*_opp(i)(x) when x - c_i is outside
// Objective: J = 1/(T - t0) sum_{t = t0+1..T} j(f^t), with
//   j(f) = w_dp cs2 (<rho>_{x = xin} - <rho>_{x = xout}) + w_drag omega sum_x alpha(x) j_x(x)
//          + w_flux <j_x>                  (<.> over the lattice's cells)
// (the drag term is the momentum the porosity removes in the next collision).
//
// Reverse step (lambda^t = dJ/df^t):
//   mu_i(y)    = lambda^{t+1}_i(y + c_i)  or  lambda^{t+1}_opp(i)(y) when y + c_i is outside
//   lambda^t(y) = (dC/df)^T mu(y) + dj/df(f^t)(y) / (T - t0)          (t > t0)
//   g_alpha(y) += mu(y) . dC/dalpha + w_drag omega j_x(f^t(y)) / (T - t0)
//   g_s(plane) += mu(y) . dC/ds        (s scales the plane's imposed velocity)
//
// Bulk cells, closed form. With rho = 1 + rhoBar, u_d = (1 - alpha)(j / rho + F/2),
// J = rho u_d, feq_i(rhoBar, J) = t_i (rhoBar + 3 c_i.J + (4.5 (c_i.J)^2 - 1.5 J^2) / rho),
// Pi = sum c c f - rhoBar/3 I - j j^T / rho, fneq_i(Pi) = 4.5 t_i (c_i c_i^T - I/3) : Pi and
// the Guo source G_i(u_d) = (1 - w/2) t_i (3 (c_i - u_d).F + 9 (c_i.u_d)(c_i.F)):
//   BGK:  f' = (1 - w) f + w feq(rhoBar, J) + G(u_d)
//   RBGK: f' = feq(rhoBar, J) + (1 - w) fneq(Pi) + G(u_d)
// and the contractions of mu
//   A = sum_i mu_i t_i (1 - (4.5 (c_i.J)^2 - 1.5 J^2) / rho^2)      (d feq / d rhoBar)
//   B = sum_i mu_i t_i (3 c_i + (9 (c_i.J) c_i - 3 J) / rho)       (d feq / d J)
//   M = sum_i mu_i 4.5 t_i (c_i c_i^T - I/3)                       (d fneq / d Pi)
//   g = (1 - w/2) sum_i mu_i t_i (-3 F + 9 (c_i.F) c_i)            (d G / d u_d)
// the reverse collision is lambda_k = s + v.c_k (+ (1 - w) c_k^T M c_k for RBGK) with
//   BGK:  lambda_k += (1 - w) mu_k,  s = w (A + (1-alpha) B.F/2) - (1-alpha) g.j / rho^2,
//         v = w (1-alpha) B + (1-alpha) g / rho
//   RBGK: s = A + (1-alpha) B.F/2 + (1-w)(j^T M j / rho^2 - tr M / 3) - (1-alpha) g.j / rho^2,
//         v = (1-alpha) B - 2 (1-w) M j / rho + (1-alpha) g / rho
//   mu . dC/dalpha = -(w_B B.(j + rho F/2) + g.(j / rho + F/2)), w_B = w (BGK) or 1 (RBGK).
// Plane cells use the device completion and collision routines evaluated on
// dual numbers, one thread per (cell, direction).
//
// The reverse sweep needs of f^t only rhoBar and j for bulk cells (32 bytes
// per cell instead of 19 populations): the recompute runs the forward step
// kernels with moment recording on, and keeps the full f^t of the plane cells
// only. Multi-block lattices run the reverse step block by block and refresh
// the envelopes of lambda from their owners after every reverse step.
#pragma once

#include "kernels.h"

namespace plbamd {

// Forward-mode dual number: value and one directional derivative.
struct Dual {
    double v, d;
    __host__ __device__ constexpr Dual() : v(0), d(0) {}
    __host__ __device__ constexpr Dual(double x) : v(x), d(0) {}
    __host__ __device__ constexpr Dual(double x, double dx) : v(x), d(dx) {}
};
__host__ __device__ constexpr Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.d + b.d); }
__host__ __device__ constexpr Dual operator-(Dual a, Dual b) { return Dual(a.v - b.v, a.d - b.d); }
__host__ __device__ constexpr Dual operator-(Dual a) { return Dual(-a.v, -a.d); }
__host__ __device__ constexpr Dual operator*(Dual a, Dual b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
__host__ __device__ constexpr Dual operator/(Dual a, Dual b) {
    return Dual(a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v));
}
__host__ __device__ constexpr Dual operator*(int a, Dual b) { return Dual(a * b.v, a * b.d); }
__host__ __device__ constexpr Dual operator*(Dual a, int b) { return Dual(a.v * b, a.d * b); }
__host__ __device__ constexpr Dual operator-(int a, Dual b) { return Dual(a - b.v, -b.d); }
__host__ __device__ inline Dual& operator+=(Dual& a, Dual b) { a = a + b; return a; }
__host__ __device__ inline Dual& operator-=(Dual& a, Dual b) { a = a - b; return a; }
__host__ __device__ inline Dual& operator*=(Dual& a, Dual b) { a = a * b; return a; }

struct AdjointArgs {
    const void* mom;          // moments of f^t: rhoBar, jx, jy, jz, [4][momStride] by cellOffset + cell, type LT
    uint64_t momStride;
    const double* planeF;     // f^t of the special (plane) records, 19 per record
    const void* lamNext;      // lambda^{t+1} of this block (slot stride g.slot), FP64 or FP32
    void* lam;                // lambda^t of this block
    const void* alpha;        // porosity by cellOffset + cell, type LT (the forward's arithmetic type), or null
    void* gAlpha;             // gradient accumulator by cellOffset + cell, type LT, or null
    const uint8_t* tags;      // BulkTag by cellOffset + cell (BT_SKIP: plane cells, or link-wall solid), or null
    const uint32_t* walls;    // link-wall mask by cellOffset + cell (bit i: wall at x + c_i), or null
    int solidSkip;            // 1: the BT_SKIP cells are link-wall solid (no planes): no objective there
    BlockGeom g;
    double omega;
    double force[3];
    double invT;              // 1 / (T - t0)
    double wdp, wdrag;        // objective weights
    double wflux;             // w_flux / (number of lattice cells)
    double inW, outW;         // cs2 / |plane| of the inlet and outlet measuring planes
    int xin, xout;            // global x of the measuring planes
    int explicitTerm;         // 1 when j(f^t) is part of J (t > t0)
    // every block in one launch: block b takes workgroups [wgBegin[b], wgBegin[b+1])
    // of 64 (z) x 4 (y) threads over its (ceil(nz/64), ceil(ny/4), nx) tiles
    const BlockGeom* blocks;
    const uint16_t* wgBlock;  // block of every workgroup
    const uint32_t* wgBegin;  // first workgroup of every block
    double* acc;              // per workgroup: J accumulated over the reverse sweep
};

// Cell of this thread (one cell per thread) and the per-workgroup J
// accumulation (fixed order: each workgroup adds its own record once per step).
struct AdjointCell {
    BlockGeom g;
    int x, y, z;
};
__device__ __forceinline__ AdjointCell adjoint_cell_of(const AdjointArgs& a) {
    AdjointCell c;
    const uint32_t wg = blockIdx.x;
    const int b = (int)a.wgBlock[wg];
    c.g = a.blocks[b];
    const uint32_t local = wg - a.wgBegin[b];
    const uint32_t gz = (uint32_t)(c.g.nz + 63) / 64, gy = (uint32_t)(c.g.ny + 3) / 4;
    c.z = (int)(local % gz) * 64 + (int)threadIdx.x;
    const uint32_t rest = local / gz;
    c.y = (int)(rest % gy) * 4 + (int)threadIdx.y;
    c.x = (int)(rest / gy);
    return c;
}
// The shuffle tree and wavefront order of block_reduce_store, for the one value.
__device__ __forceinline__ void accumulate_J(double J, double* acc) {
    __shared__ double red[16];
    const int t = threadIdx.x + threadIdx.y * blockDim.x;
    const int n = blockDim.x * blockDim.y;
    const int lane = t & 63, wave = t >> 6, nWaves = (n + 63) >> 6;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_down(J, off, 64);
        if (lane + off < 64 && t + off < n) J += o;
    }
    if (lane == 0) red[wave] = J;
    __syncthreads();
    if (t == 0) {
        double r = red[0];
        for (int w = 1; w < nWaves; ++w) r += red[w];
        acc[blockIdx.x] += r;
    }
}

// dj/df_k(y) for the objective (plane-mean pressure difference, drag, mean flux).
__device__ __forceinline__ void objective_gradient(const AdjointArgs& a, int gx, double alphaY, double djdf[19]) {
    const double p = (gx == a.xin ? a.inW : 0.0) - (gx == a.xout ? a.outW : 0.0);
    for (int k = 0; k < 19; ++k) djdf[k] = a.wdp * p + (a.wdrag * a.omega * alphaY + a.wflux) * (double)cx(k);
}

// mu(y) gathered from lambda^{t+1} (transpose of streaming with edge
// reflection: mu_i(y) = lambda_i(y + c_i), or lambda_opp(i)(y) where y + c_i
// leaves the block), by the lanes of a wavefront that hold consecutive cells
// z0 .. z0 + 63 of one row (lane = z - z0; the row, and so every x/y edge test,
// is the wavefront's). Every direction is loaded at the lane's own z in its
// source row, aligned; the z-moving ones are then passed one lane along by a
// cross-lane read (ds_bpermute) instead of a load straddling one more cache line
// per wavefront. A lane whose source lies beyond the wavefront's cells, wraps
// around a periodic row or leaves the block loads its own value. Every lane of
// the row with z < nz takes part (their loads are what the others read).
// WALLS, wall: the cell's link-wall mask (bit i: wall at x + c_i): those
// directions take lambda_opp(i) of the cell itself, loaded beside the row loads.
// A lattice without link walls takes the kernel without them: no conditional
// load between the row loads, which then go out back to back.
template<typename LT, typename C, bool WALLS>
__device__ __forceinline__ void gather_mu_wave(const BlockGeom& g, const LT* lamNext, int x, int y, int z, int z0,
                                               uint64_t cell, C mu[19], uint32_t wall) {
    const int lane = z - z0;
    mu[0] = (C)lamNext[cell];
#pragma unroll
    for (int i = 1; i < 19; ++i) {
        int dx, dy;
        const LT* refl = lamNext + (uint64_t)opp(i) * g.slot + cell;
        if (!(shift(x, cx(i), g.nx, g.px, dx) && shift(y, cy(i), g.ny, g.py, dy))) {
            mu[i] = (C)*refl;
            continue;
        }
        const bool wr = WALLS && ((wall >> i) & 1u);
        const LT* row = lamNext + (uint64_t)i * g.slot + ((uint64_t)dx * g.ny + dy) * g.nz;
        const LT v = row[z];
        const LT w = wr ? *refl : LT(0);
        if (cz(i) == 0) {
            mu[i] = wr ? (C)w : (C)v;
            continue;
        }
        const LT nb = __shfl(v, lane + cz(i), 64);
        const int zs = z + cz(i);
        if (wr) {
            mu[i] = (C)w;
        } else if (zs >= z0 && zs < z0 + 64 && zs < g.nz) {
            mu[i] = (C)nb;
        } else {
            int dz;
            mu[i] = shift(z, cz(i), g.nz, g.pz, dz) ? (C)row[dz] : (C)*refl;
        }
    }
}

// c_x of a direction known only at run time, from bit masks (a lookup into a
// local table would live in scratch memory).
__device__ __forceinline__ int cx_of(uint32_t i) {
    constexpr uint32_t neg = (1u << 1) | (1u << 4) | (1u << 5) | (1u << 6) | (1u << 7);
    constexpr uint32_t pos = (1u << 10) | (1u << 13) | (1u << 14) | (1u << 15) | (1u << 16);
    return (int)((pos >> i) & 1u) - (int)((neg >> i) & 1u);
}

// Reverse step of one bulk (non-plane) cell; plane cells are left to adjoint_planes.
// The arithmetic type C is the adjoint state's: FP64 for an FP64 forward, FP32
// for a 16-bit one (whose forward arithmetic is FP32), which is also the type
// of the porosity and of the porosity-gradient accumulator (one segment's sum
// for a 16-bit forward, added into an FP64 total); the objective accumulates
// in FP64.
template<typename LT, int DYN, bool FORCED, bool WALLS>
__device__ __forceinline__ void adjoint_cell(const AdjointArgs& a, const BlockGeom& g, int x, int y, int z, int z0,
                                             double& J) {
    using C = LT;
    const uint64_t cell = ((uint64_t)x * g.ny + y) * g.nz + z;
    const uint64_t idx = g.cellOffset + cell;
    const LT* lamNext = reinterpret_cast<const LT*>(a.lamNext) + g.offset;
    LT* lam = reinterpret_cast<LT*>(a.lam) + g.offset;
    // the gather first, by every cell of the row (plane cells included: their
    // loads feed their neighbours' lanes), then the skip test
    // a link wall at x + c_i sends f_i back into the cell's own slot opp(i) (the
    // forward's reflection): mu_i = lambda_opp(i) of the cell itself. The mask
    // is loaded first, so that those loads go out with the gather's.
    const uint32_t wall = WALLS ? a.walls[idx] : 0u;
    C mu[19];
    gather_mu_wave<LT, C, WALLS>(g, lamNext, x, y, z, z0, cell, mu, wall);
    if (!(a.tags && a.tags[idx] == BT_SKIP)) {
        const C al = a.alpha ? reinterpret_cast<const LT*>(a.alpha)[idx] : C(0);
        const double al64 = (double)al;
        const C dm = C(1) - al;
        const LT* mom = reinterpret_cast<const LT*>(a.mom);
        const C rhoBar = (C)mom[idx];
        const C j[3] = {(C)mom[a.momStride + idx], (C)mom[2 * a.momStride + idx], (C)mom[3 * a.momStride + idx]};
        const C rho = C(1) + rhoBar;
        const C F[3] = {FORCED ? (C)a.force[0] : C(0), FORCED ? (C)a.force[1] : C(0), FORCED ? (C)a.force[2] : C(0)};
        const C Jv[3] = {dm * (j[0] + rho * F[0] * C(0.5)), dm * (j[1] + rho * F[1] * C(0.5)),
                         dm * (j[2] + rho * F[2] * C(0.5))};
        const C J2 = Jv[0] * Jv[0] + Jv[1] * Jv[1] + Jv[2] * Jv[2];
        const C w = (C)a.omega;
        // Every term below is linear in mu through its weighted moments
        //   m0 = sum_i t_i mu_i, m1_c = sum_i t_i c_ic mu_i, m2_ce = sum_i t_i c_ic c_ie mu_i,
        // accumulated with the velocity components as compile-time signs (a
        // product by a zero component is not dropped under IEEE rules, so the
        // zero terms are left out explicitly).
        C m0 = C(0), m1[3] = {C(0), C(0), C(0)};
        C m2[3][3] = {{C(0), C(0), C(0)}, {C(0), C(0), C(0)}, {C(0), C(0), C(0)}};
#pragma unroll
        for (int i = 0; i < 19; ++i) {
            const C mt = mu[i] * weight<C>(i);
            m0 += mt;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                if (ci(i, c) == 0) continue;
                m1[c] += ci(i, c) > 0 ? mt : -mt;
#pragma unroll
                for (int e = c; e < 3; ++e)
                    if (ci(i, e) != 0) m2[c][e] += ci(i, c) * ci(i, e) > 0 ? mt : -mt;
            }
        }
        m2[1][0] = m2[0][1];
        m2[2][0] = m2[0][2];
        m2[2][1] = m2[1][2];
        const C ir = C(1) / rho, ir2 = ir * ir;
        // sum_i t_i mu_i (c_i.Jv) c_ic = (m2 Jv)_c, sum_i t_i mu_i (c_i.Jv)^2 = Jv.m2.Jv
        C Jm[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) Jm[c] = m2[c][0] * Jv[0] + m2[c][1] * Jv[1] + m2[c][2] * Jv[2];
        const C JmJ = Jv[0] * Jm[0] + Jv[1] * Jm[1] + Jv[2] * Jm[2];
        const C A = m0 - (C(4.5) * JmJ - C(1.5) * J2 * m0) * ir2;
        C B[3], gG[3] = {C(0), C(0), C(0)};
#pragma unroll
        for (int c = 0; c < 3; ++c) B[c] = C(3) * m1[c] + (C(9) * Jm[c] - C(3) * Jv[c] * m0) * ir;
        if constexpr (FORCED) {
            const C h = C(1) - w * C(0.5);
#pragma unroll
            for (int c = 0; c < 3; ++c)
                gG[c] = (C(-3) * F[c] * m0 + C(9) * (m2[c][0] * F[0] + m2[c][1] * F[1] + m2[c][2] * F[2])) * h;
        }
        // M_ce = 4.5 sum_i t_i mu_i (c_ic c_ie - delta_ce / 3)
        C M[3][3];
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int e = 0; e < 3; ++e) M[c][e] = C(4.5) * (c == e ? m2[c][e] - m0 * (C(1) / C(3)) : m2[c][e]);
        const C BF = B[0] * F[0] + B[1] * F[1] + B[2] * F[2];
        const C gj = gG[0] * j[0] + gG[1] * j[1] + gG[2] * j[2];
        C s, v[3];
        if constexpr (DYN == 0) {
            s = w * (A + dm * BF * C(0.5)) - dm * gj * ir2;
#pragma unroll
            for (int c = 0; c < 3; ++c) v[c] = w * dm * B[c] + dm * gG[c] * ir;
        } else {
            C Mj[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) Mj[c] = M[c][0] * j[0] + M[c][1] * j[1] + M[c][2] * j[2];
            const C jMj = j[0] * Mj[0] + j[1] * Mj[1] + j[2] * Mj[2];
            const C trM = M[0][0] + M[1][1] + M[2][2];
            s = A + dm * BF * C(0.5) + (C(1) - w) * (jMj * ir2 - trM * (C(1) / C(3))) - dm * gj * ir2;
#pragma unroll
            for (int c = 0; c < 3; ++c) v[c] = dm * B[c] - C(2) * (C(1) - w) * Mj[c] * ir + dm * gG[c] * ir;
        }
        const int gx = g.gx0 + x;
        // dj/df of the objective (objective_gradient), in C
        const C p = (C)((gx == a.xin ? a.inW : 0.0) - (gx == a.xout ? a.outW : 0.0));
        const C djBase = (C)a.wdp * p, djX = (C)(a.wdrag * a.omega * al64 + a.wflux);
        const C invT = (C)a.invT;
#pragma unroll
        for (int k = 0; k < 19; ++k) {
            C val = s;
#pragma unroll
            for (int c = 0; c < 3; ++c)
                if (ci(k, c) != 0) val += ci(k, c) > 0 ? v[c] : -v[c];
            if constexpr (DYN == 0) {
                val += (C(1) - w) * mu[k];
            } else {
                // c_k.M.c_k
                C cMc = C(0);
#pragma unroll
                for (int c = 0; c < 3; ++c)
#pragma unroll
                    for (int e = 0; e < 3; ++e)
                        if (ci(k, c) != 0 && ci(k, e) != 0) cMc += ci(k, c) * ci(k, e) > 0 ? M[c][e] : -M[c][e];
                val += (C(1) - w) * cMc;
            }
            if (a.explicitTerm) val += invT * (cx(k) == 0 ? djBase : (cx(k) > 0 ? djBase + djX : djBase - djX));
            lam[(uint64_t)k * g.slot + cell] = (LT)val;
        }
        if (in_bulk(g, x, y, z)) {
            const C wB = DYN == 0 ? w : C(1);
            const C ga = -(wB * (B[0] * (j[0] + rho * F[0] * C(0.5)) + B[1] * (j[1] + rho * F[1] * C(0.5)) +
                                 B[2] * (j[2] + rho * F[2] * C(0.5))) +
                           gG[0] * (j[0] * ir + F[0] * C(0.5)) + gG[1] * (j[1] * ir + F[1] * C(0.5)) +
                           gG[2] * (j[2] * ir + F[2] * C(0.5)));
            double ga64 = (double)ga;
            if (a.explicitTerm) {
                ga64 += a.invT * a.wdrag * a.omega * (double)j[0];
                const double p64 = (gx == a.xin ? a.inW : 0.0) - (gx == a.xout ? a.outW : 0.0);
                J += a.invT * (a.wdp * p64 * (double)rho + a.wdrag * a.omega * al64 * (double)j[0] +
                               a.wflux * (double)j[0]);
            }
            if (a.gAlpha) reinterpret_cast<LT*>(a.gAlpha)[idx] += (LT)ga64;
        }
    }
}

// Reverse step of every bulk cell of every block (64 x 4 cells per workgroup).
// WALLS: the lattice has link-wall masks (a.walls not null).
template<typename LT, int DYN, bool FORCED, bool WALLS>
__global__ void __launch_bounds__(256) adjoint_bulk(AdjointArgs a) {
    const AdjointCell c = adjoint_cell_of(a);
    double J = 0.0;
    if (c.z < c.g.nz && c.y < c.g.ny)
        adjoint_cell<LT, DYN, FORCED, WALLS>(a, c.g, c.x, c.y, c.z, c.z - (int)threadIdx.x, J);
    if (a.explicitTerm) accumulate_J(J, a.acc);   // J is zero before the objective's window
}

struct AdjointSpecialArgs {
    AdjointArgs base;
    const SpecialCell* cells;
    uint32_t count;
    const double* values;
    double* partials;           // per workgroup: J accumulated over the reverse sweep
    double* gScale;             // per special cell: d/ds contribution (summed per plane on the host)
};

// Reverse step for plane cells: the collision Jacobian by dual numbers. One
// thread per (record, direction): direction k < 19 seeds df_k and yields
// lambda_k = mu . dC/df_k; direction 19 seeds the plane's velocity scale.
template<typename LT>
__global__ void __launch_bounds__(256) adjoint_planes(AdjointSpecialArgs s) {
    const AdjointArgs& a = s.base;
    const BlockGeom g = a.g;
    double J = 0.0;
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t k = t % s.count, dirn = t / s.count;
    if (t < 20 * s.count) {
        const SpecialCell sc = s.cells[k];
        const bool velocity = ((sc.info >> 3) & 1u) == PLBAMD_VELOCITY;
        if (dirn < 19 || velocity) {
            const int x = sc.x, y = sc.y, z = sc.z;
            const uint64_t cell = sc.cell;
            const double* f0 = a.planeF + 19 * (uint64_t)k;
            Dual fd[19];
#pragma unroll
            for (int i = 0; i < 19; ++i) fd[i] = Dual(f0[i], (uint32_t)i == dirn ? 1.0 : 0.0);
            const double* v0 = s.values + sc.value;
            const double vals[3] = {v0[0], velocity ? v0[1] : 0.0, velocity ? v0[2] : 0.0};
            // the velocity-scale direction: d(s u)/ds = u
            const double sd = dirn == 19 ? 1.0 : 0.0;
            const Dual vd[3] = {Dual(vals[0], sd * vals[0]), Dual(vals[1], sd * vals[1]), Dual(vals[2], sd * vals[2])};
            complete_plane<Dual, Dual>(fd, sc.info, vd);
            Dual rhoBar, drag[3];
            const Dual F0[3] = {Dual(0.0), Dual(0.0), Dual(0.0)};
            collide_bulk<Dual, 0, false>(fd, Dual(a.omega), F0, Dual(0.0), rhoBar, drag);
            // mu . d f' / d(direction), mu gathered from lambda^{t+1} as it is consumed
            const LT* lamNext = reinterpret_cast<const LT*>(a.lamNext);
            double dot = (double)lamNext[cell] * fd[0].d;
#pragma unroll
            for (int i = 1; i < 19; ++i) {
                int dx, dy, dz;
                const bool ok = shift(x, cx(i), g.nx, g.px, dx) && shift(y, cy(i), g.ny, g.py, dy) &&
                                shift(z, cz(i), g.nz, g.pz, dz);
                const double m = (double)(ok ? lamNext[(uint64_t)i * g.slot + lin(g, dx, dy, dz)]
                                             : lamNext[(uint64_t)opp(i) * g.slot + cell]);
                dot += m * fd[i].d;
            }
            const int gx = g.gx0 + x;
            const double p = (gx == a.xin ? a.inW : 0.0) - (gx == a.xout ? a.outW : 0.0);
            if (dirn < 19) {
                // dj/df_dirn of a plane cell (no porosity there), without a local array
                if (a.explicitTerm) dot += a.invT * (a.wdp * p + a.wflux * (double)cx_of(dirn));
                reinterpret_cast<LT*>(a.lam)[(uint64_t)dirn * g.slot + cell] = (LT)dot;
                if (dirn == 0 && a.explicitTerm) {
                    double rhoBar0 = 0, jx0 = 0;
#pragma unroll
                    for (int i = 0; i < 19; ++i) {
                        rhoBar0 += f0[i];
                        jx0 += cx(i) * f0[i];
                    }
                    J += a.invT * (a.wdp * p * (1.0 + rhoBar0) + a.wflux * jx0);
                }
            } else {
                s.gScale[k] += dot;
            }
        }
    }
    accumulate_J(J, s.partials);
}

// lambda^T = dj/df(f^T) / (T - t0); J += j(f^T) / (T - t0);
// g_alpha += w_drag omega j_x(f^T) / (T - t0).
template<typename LT>
__global__ void __launch_bounds__(256) adjoint_final(AdjointArgs a) {
    const AdjointCell c = adjoint_cell_of(a);
    const BlockGeom& g = c.g;
    const int x = c.x, y = c.y, z = c.z;
    double J = 0.0;
    if (z < g.nz && y < g.ny) {
        LT* lam = reinterpret_cast<LT*>(a.lam) + g.offset;
        const uint64_t cell = ((uint64_t)x * g.ny + y) * g.nz + z;
        const uint64_t idx = g.cellOffset + cell;
        const bool plane = a.tags && a.tags[idx] == BT_SKIP;
        const double al = (a.alpha && !plane) ? (double)reinterpret_cast<const LT*>(a.alpha)[idx] : 0.0;
        const LT* mom = reinterpret_cast<const LT*>(a.mom);
        const double rhoBar = (double)mom[idx];
        const double jx = (double)mom[a.momStride + idx];
        const int gx = g.gx0 + x;
        double djdf[19];
        objective_gradient(a, gx, al, djdf);
#pragma unroll
        for (int k = 0; k < 19; ++k) lam[(uint64_t)k * g.slot + cell] = (LT)(a.invT * djdf[k]);
        if (in_bulk(g, x, y, z) && !(plane && a.solidSkip)) {
            if (a.gAlpha && !plane) reinterpret_cast<LT*>(a.gAlpha)[idx] += (LT)(a.invT * a.wdrag * a.omega * jx);
            const double p = (gx == a.xin ? a.inW : 0.0) - (gx == a.xout ? a.outW : 0.0);
            J += a.invT * (a.wdp * p * (1.0 + rhoBar) + a.wdrag * a.omega * al * jx + a.wflux * jx);
        }
    }
    accumulate_J(J, a.acc);
}

// The FP32 gradient accumulator of a 16-bit adjoint (one segment's sum) added
// into the FP64 gradient, and cleared for the next segment.
static __global__ void __launch_bounds__(256) flush_gradient(float* seg, double* total, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        total[i] += (double)seg[i];
        seg[i] = 0.0f;
    }
}

// Sum of the per-workgroup accumulators of the bulk and plane kernels (one
// workgroup, fixed order).
static __global__ void __launch_bounds__(256) sum_values(const double* a, uint64_t na, const double* b, uint64_t nb,
                                                  double* out) {
    double v = 0.0;
    for (uint64_t i = threadIdx.x; i < na; i += blockDim.x) v += a[i];
    for (uint64_t i = threadIdx.x; i < nb; i += blockDim.x) v += b[i];
    double r[8] = {v, 0, 0, 0, 0, 0, 0, 0};
    block_reduce_store(r, out);
}

}  // namespace plbamd
