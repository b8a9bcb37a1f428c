// Discrete adjoint of the FP64 BGK collide-and-stream with a Brinkman porosity
// term, for gradients of a time-averaged objective with respect to the
// porosity field and the scale of each velocity plane. No Palabos counterpart.
//
// Forward step t (state f^t: post-stream populations, natural layout):
//   f*(y)      = C(f^t(y); alpha(y), planes)      collision (plane cells:
//                                                 completion, then BGK)
//   f^{t+1}_i(x) = f*_i(x - c_i)  or  f*_opp(i)(x) when x - c_i is outside
// Objective: J = (1/T) sum_{t=1..T} j(f^t), with
//   j(f) = w_dp cs2 (<rho>_{x = xin} - <rho>_{x = xout}) + w_drag omega sum_x alpha(x) j_x(x)
// (the second term is the momentum the porosity removes in the next collision).
//
// Reverse step (lambda^t = dJ/df^t):
//   mu_i(y)    = lambda^{t+1}_i(y + c_i)  or  lambda^{t+1}_opp(i)(y) when y + c_i is outside
//   lambda^t(y) = (dC/df)^T mu(y) + (1/T) dj/df(f^t)(y)          (t >= 1)
//   g_alpha(y) += mu(y) . dC/dalpha + (1/T) w_drag omega j_x(f^t(y))
//   g_s(plane) += mu(y) . dC/ds        (s scales the plane's imposed velocity)
// Bulk cells, closed form (f' = (1-w) f + w feq(rhoBar, J), J = (1-alpha) j,
// rho = 1 + rhoBar, feq_i = t_i (rhoBar + 3 c_i.J + (4.5 (c_i.J)^2 - 1.5 J^2) / rho)):
//   A = sum_i mu_i t_i (1 - (4.5 (c_i.J)^2 - 1.5 J^2) / rho^2)
//   B = sum_i mu_i t_i (3 c_i + (9 (c_i.J) c_i - 3 J) / rho)
//   (dC/df)^T mu = (1-w) mu + w (A + (1-alpha) B.c_k),   mu . dC/dalpha = -w B.j
// Plane cells use the device completion and collision routines evaluated on
// dual numbers, one direction per pass (the Jacobian-transpose product column
// by column).
//
// The reverse sweep needs of f^t only rhoBar and j for bulk cells (32 bytes
// per cell instead of 152): the recompute runs the forward step kernels with
// moment recording on, and keeps the full f^t of the plane cells only.
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
    const double* mom;        // moments of f^t: rhoBar, jx, jy, jz, [4][momStride]
    uint64_t momStride;
    const double* planeF;     // f^t of the special (plane) records, 19 per record
    const double* lamNext;    // lambda^{t+1}
    double* lam;              // lambda^t
    const double* alpha;      // porosity per cell
    double* gAlpha;           // gradient accumulator per cell
    const uint8_t* tags;      // BulkTag per cell (BT_SKIP: plane cells)
    BlockGeom g;
    double omega;
    double invT;              // 1 / T
    double wdp, wdrag;        // objective weights
    double inW, outW;         // cs2 / |plane| of the inlet and outlet measuring planes
    int xin, xout;
    int explicitTerm;         // 1 when j(f^t) is part of J (t >= 1)
    double* partials;         // per workgroup: J contribution (slot 0)
};

// dj/df_k(y) for the objective (plane-mean pressure difference and drag).
__device__ __forceinline__ void objective_gradient(const AdjointArgs& a, int x, double alphaY, double djdf[19]) {
    const double p = (x == a.xin ? a.inW : 0.0) - (x == a.xout ? a.outW : 0.0);
    for (int k = 0; k < 19; ++k) djdf[k] = a.wdp * p + a.wdrag * a.omega * alphaY * (double)cx(k);
}

// mu(y) gathered from lambda^{t+1} (transpose of streaming with edge reflection).
__device__ __forceinline__ void gather_mu(const AdjointArgs& a, int x, int y, int z, uint64_t cell, double mu[19]) {
    const BlockGeom& g = a.g;
    mu[0] = a.lamNext[cell];
#pragma unroll
    for (int i = 1; i < 19; ++i) {
        int dx, dy, dz;
        const bool ok = shift(x, cx(i), g.nx, g.px, dx) && shift(y, cy(i), g.ny, g.py, dy) &&
                        shift(z, cz(i), g.nz, g.pz, dz);
        mu[i] = ok ? a.lamNext[(uint64_t)i * g.slot + lin(g, dx, dy, dz)]
                   : a.lamNext[(uint64_t)opp(i) * g.slot + cell];
    }
}

// Launch shape of the cell-wise adjoint kernels: z along threadIdx.x (64),
// four y rows per workgroup, one x plane per blockIdx.z.
__device__ __forceinline__ uint64_t adjoint_wg() {
    return ((uint64_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
}

// Reverse step for bulk (non-plane) cells; plane cells are left to adjoint_planes.
__global__ void __launch_bounds__(256) adjoint_bulk(AdjointArgs a) {
    const BlockGeom g = a.g;
    const int z = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    const int x = blockIdx.z;
    double J = 0.0;
    const uint64_t cell = ((uint64_t)x * g.ny + y) * g.nz + z;
    if (z < g.nz && y < g.ny && !(a.tags && a.tags[cell] == BT_SKIP)) {
        double mu[19];
        gather_mu(a, x, y, z, cell, mu);
        const double al = a.alpha ? a.alpha[cell] : 0.0;
        const double rhoBar = a.mom[cell];
        const double j[3] = {a.mom[a.momStride + cell], a.mom[2 * a.momStride + cell], a.mom[3 * a.momStride + cell]};
        const double rho = 1.0 + rhoBar;
        const double Jv[3] = {(1 - al) * j[0], (1 - al) * j[1], (1 - al) * j[2]};
        const double J2 = Jv[0] * Jv[0] + Jv[1] * Jv[1] + Jv[2] * Jv[2];
        double A = 0, B[3] = {0, 0, 0};
#pragma unroll
        for (int i = 0; i < 19; ++i) {
            const double t = weight<double>(i);
            const double cJ = cx(i) * Jv[0] + cy(i) * Jv[1] + cz(i) * Jv[2];
            A += mu[i] * t * (1.0 - (4.5 * cJ * cJ - 1.5 * J2) / (rho * rho));
#pragma unroll
            for (int c = 0; c < 3; ++c) B[c] += mu[i] * t * (3.0 * ci(i, c) + (9.0 * cJ * ci(i, c) - 3.0 * Jv[c]) / rho);
        }
        const double w = a.omega;
        double djdf[19];
        objective_gradient(a, x, al, djdf);
#pragma unroll
        for (int k = 0; k < 19; ++k) {
            double v = (1 - w) * mu[k] + w * (A + (1 - al) * (B[0] * cx(k) + B[1] * cy(k) + B[2] * cz(k)));
            if (a.explicitTerm) v += a.invT * djdf[k];
            a.lam[(uint64_t)k * g.slot + cell] = v;
        }
        double ga = -w * (B[0] * j[0] + B[1] * j[1] + B[2] * j[2]);
        if (a.explicitTerm) {
            ga += a.invT * a.wdrag * w * j[0];
            const double p = (x == a.xin ? a.inW : 0.0) - (x == a.xout ? a.outW : 0.0);
            J += a.invT * (a.wdp * p * rho + a.wdrag * w * al * j[0]);
        }
        if (a.gAlpha) a.gAlpha[cell] += ga;
    }
    double r[8] = {J, 0, 0, 0, 0, 0, 0, 0};
    block_reduce_store(r, a.partials + 8 * adjoint_wg());
}

struct AdjointSpecialArgs {
    AdjointArgs base;
    const SpecialCell* cells;
    uint32_t count;
    const double* values;
    double* partials;           // per workgroup: slot 0 J
    double* gScale;             // per special cell: d/ds contribution (summed per plane on the host)
};

// Reverse step for plane cells: the collision Jacobian by dual numbers. One
// thread per (record, direction): direction k < 19 seeds df_k and yields
// lambda_k = mu . dC/df_k; direction 19 seeds the plane's velocity scale.
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
            double mu[19];
            gather_mu(a, x, y, z, cell, mu);
            Dual fd[19];
            double rhoBar0 = 0;
            for (int i = 0; i < 19; ++i) {
                const double fi = a.planeF[19 * (uint64_t)k + i];
                fd[i] = Dual(fi, (uint32_t)i == dirn ? 1.0 : 0.0);
                rhoBar0 += fi;
            }
            const double* v0 = s.values + sc.value;
            const double vals[3] = {v0[0], velocity ? v0[1] : 0.0, velocity ? v0[2] : 0.0};
            // the velocity-scale direction: d(s u)/ds = u
            const double sd = dirn == 19 ? 1.0 : 0.0;
            const Dual vd[3] = {Dual(vals[0], sd * vals[0]), Dual(vals[1], sd * vals[1]), Dual(vals[2], sd * vals[2])};
            complete_plane<Dual, Dual>(fd, sc.info, vd);
            Dual rhoBar, drag[3];
            const Dual F0[3] = {Dual(0.0), Dual(0.0), Dual(0.0)};
            collide_bulk<Dual, 0, false>(fd, Dual(a.omega), F0, Dual(0.0), rhoBar, drag);
            double dot = 0;
            for (int i = 0; i < 19; ++i) dot += mu[i] * fd[i].d;
            const double p = (x == a.xin ? a.inW : 0.0) - (x == a.xout ? a.outW : 0.0);
            if (dirn < 19) {
                double djdf[19];
                objective_gradient(a, x, 0.0, djdf);
                if (a.explicitTerm) dot += a.invT * djdf[dirn];
                a.lam[(uint64_t)dirn * g.slot + cell] = dot;
                if (dirn == 0 && a.explicitTerm) J += a.invT * a.wdp * p * (1.0 + rhoBar0);
            } else {
                s.gScale[k] += dot;
            }
        }
    }
    double r[8] = {J, 0, 0, 0, 0, 0, 0, 0};
    block_reduce_store(r, s.partials + 8 * (uint64_t)blockIdx.x);
}

// lambda^T = (1/T) dj/df(f^T); J += (1/T) j(f^T); g_alpha += (1/T) w_drag omega j_x(f^T).
__global__ void __launch_bounds__(256) adjoint_final(AdjointArgs a) {
    const BlockGeom g = a.g;
    const int z = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    const int x = blockIdx.z;
    double J = 0.0;
    const uint64_t cell = ((uint64_t)x * g.ny + y) * g.nz + z;
    if (z < g.nz && y < g.ny) {
        const bool plane = a.tags && a.tags[cell] == BT_SKIP;
        const double al = (a.alpha && !plane) ? a.alpha[cell] : 0.0;
        const double rhoBar = a.mom[cell];
        const double jx = a.mom[a.momStride + cell];
        double djdf[19];
        objective_gradient(a, x, al, djdf);
        for (int k = 0; k < 19; ++k) a.lam[(uint64_t)k * g.slot + cell] = a.invT * djdf[k];
        if (a.gAlpha && !plane) a.gAlpha[cell] += a.invT * a.wdrag * a.omega * jx;
        const double p = (x == a.xin ? a.inW : 0.0) - (x == a.xout ? a.outW : 0.0);
        J += a.invT * (a.wdp * p * (1.0 + rhoBar) + a.wdrag * a.omega * al * jx);
    }
    double r[8] = {J, 0, 0, 0, 0, 0, 0, 0};
    block_reduce_store(r, a.partials + 8 * adjoint_wg());
}

}  // namespace plbamd
