// D3Q19 lattice arithmetic for the gfx950 kernels.
//
// Every routine keeps the reference's floating-point operation order (cited
// per routine), and the library is compiled with -ffp-contract=off, so FP64
// and FP32 results are the reference's bit for bit. The populations are the
// shifted f-bar = f - t_i of DefaultRoundOffPolicy
// (latticeBoltzmann/roundOffPolicy.h:73-90).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace plbamd {

// Velocity set: rest, nine "negative-first" velocities, their opposites at i+9
// (latticeBoltzmann/nearestNeighborLattices3D.hh:166-178).
__host__ __device__ constexpr int cx(int i) {
    constexpr int v[19] = {0, -1, 0, 0, -1, -1, -1, -1, 0, 0, 1, 0, 0, 1, 1, 1, 1, 0, 0};
    return v[i];
}
__host__ __device__ constexpr int cy(int i) {
    constexpr int v[19] = {0, 0, -1, 0, -1, 1, 0, 0, -1, -1, 0, 1, 0, 1, -1, 0, 0, 1, 1};
    return v[i];
}
__host__ __device__ constexpr int cz(int i) {
    constexpr int v[19] = {0, 0, 0, -1, 0, 0, -1, 1, -1, 1, 0, 0, 1, 0, 0, 1, -1, 1, -1};
    return v[i];
}
__host__ __device__ constexpr int ci(int i, int a) { return a == 0 ? cx(i) : (a == 1 ? cy(i) : cz(i)); }
__host__ __device__ constexpr int opp(int i) { return i == 0 ? 0 : (i <= 9 ? i + 9 : i - 9); }

template<typename T> __host__ __device__ constexpr T w0() { return (T)1 / (T)3; }
template<typename T> __host__ __device__ constexpr T w1() { return (T)1 / (T)18; }
template<typename T> __host__ __device__ constexpr T w2() { return (T)1 / (T)36; }
template<typename T> __host__ __device__ constexpr T weight(int i) {
    return i == 0 ? w0<T>() : ((cx(i) * cx(i) + cy(i) * cy(i) + cz(i) * cz(i)) == 1 ? w1<T>() : w2<T>());
}
template<typename T> __host__ __device__ constexpr T cs2() { return (T)1 / (T)3; }

template<typename T> __device__ __forceinline__ T inv_rho(T rhoBar) { return (T)1 / (rhoBar + (T)1); }

// rhoBar and j from directional partial sums
// (momentTemplates3D.h:351-373 partial_rho, :411-423 get_rhoBar_j).
template<typename T>
struct Surfaces {
    T xm, x0, xp, ym, yp, zm, zp;
    __device__ __forceinline__ explicit Surfaces(const T* f) {
        xm = f[1] + f[4] + f[5] + f[6] + f[7];
        xp = f[10] + f[13] + f[14] + f[15] + f[16];
        ym = f[2] + f[4] + f[8] + f[9] + f[14];
        yp = f[5] + f[11] + f[13] + f[17] + f[18];
        zm = f[3] + f[6] + f[8] + f[16] + f[18];
        zp = f[7] + f[9] + f[12] + f[15] + f[17];
        x0 = f[0] + f[2] + f[3] + f[8] + f[9] + f[11] + f[12] + f[17] + f[18];
    }
    __device__ __forceinline__ T rhoBar() const { return xm + x0 + xp; }
};

template<typename T>
__device__ __forceinline__ void rhobar_j(const T* f, T& rhoBar, T j[3]) {
    Surfaces<T> s(f);
    rhoBar = s.rhoBar();
    j[0] = (s.xp - s.xm);
    j[1] = (s.yp - s.ym);
    j[2] = (s.zp - s.zm);
}

// Plain index-order sum (momentTemplates3D.h:375-382 get_rhoBar).
template<typename T>
__device__ __forceinline__ T rhobar_sum(const T* f) {
    return f[0] + f[1] + f[2] + f[3] + f[4] + f[5] + f[6] + f[7] + f[8]
         + f[9] + f[10] + f[11] + f[12] + f[13] + f[14] + f[15] + f[16] + f[17] + f[18];
}

// rhoBar, j and the off-equilibrium stress, order xx xy xz yy yz zz
// (momentTemplates3D.h:517-539 compute_rhoBar_j_PiNeq).
template<typename T>
__device__ __forceinline__ void rhobar_j_pineq(const T* f, T& rhoBar, T j[3], T pi[6]) {
    Surfaces<T> s(f);
    rhoBar = s.rhoBar();
    j[0] = (s.xp - s.xm);
    j[1] = (s.yp - s.ym);
    j[2] = (s.zp - s.zm);
    const T ir = inv_rho(rhoBar);
    const T c = cs2<T>();
    pi[0] = s.xp + s.xm - c * rhoBar - ir * j[0] * j[0];
    pi[3] = s.yp + s.ym - c * rhoBar - ir * j[1] * j[1];
    pi[5] = s.zp + s.zm - c * rhoBar - ir * j[2] * j[2];
    pi[1] = f[4] - f[5] + f[13] - f[14] - ir * j[0] * j[1];
    pi[2] = f[6] - f[7] + f[15] - f[16] - ir * j[0] * j[2];
    pi[4] = f[8] - f[9] + f[17] - f[18] - ir * j[1] * j[2];
}

// The off-equilibrium stress of the populations for a given rhoBar and j (the
// moments of an external field), order xx xy xz yy yz zz
// (momentTemplates3D.h:482-496 compute_PiNeq with invRho(rhoBar)).
template<typename T>
__device__ __forceinline__ void pineq_given(const T* f, T rhoBar, const T j[3], T pi[6]) {
    Surfaces<T> s(f);
    const T ir = inv_rho(rhoBar);
    const T c = cs2<T>();
    pi[0] = s.xp + s.xm - c * rhoBar - ir * j[0] * j[0];
    pi[3] = s.yp + s.ym - c * rhoBar - ir * j[1] * j[1];
    pi[5] = s.zp + s.zm - c * rhoBar - ir * j[2] * j[2];
    pi[1] = f[4] - f[5] + f[13] - f[14] - ir * j[0] * j[1];
    pi[2] = f[6] - f[7] + f[15] - f[16] - ir * j[0] * j[2];
    pi[4] = f[8] - f[9] + f[17] - f[18] - ir * j[1] * j[2];
}

// Second-order equilibrium of one population
// (dynamicsTemplates3D.h:1994-1997 bgk_ma2_equilibrium).
template<typename T>
__device__ __forceinline__ T feq(int i, T rhoBar, T invRho, const T j[3], T jSqr) {
    T c_j = cx(i) * j[0] + cy(i) * j[1] + cz(i) * j[2];
    return weight<T>(i) * (rhoBar + (T)3. * c_j + invRho * ((T)4.5 * c_j * c_j - (T)1.5 * jSqr));
}

// Constants of the BGK update f_i <- om1 f_i + a_k X_i (weight class k of i);
// `omega` is the relaxation rate they stand for.
template<typename T>
struct BgkRates {
    T om1, a0, a1, a2, omega;
};

// Palabos's constants (dynamicsTemplates3D.h:2291-2294): 1 - omega and t_k omega.
template<typename T>
__host__ __device__ inline BgkRates<T> palabos_rates(T omega) {
    return {(T)1 - omega, w0<T>() * omega, w1<T>() * omega, w2<T>() * omega, omega};
}

// The 16-bit storage modes' constants. With Palabos's rounded t_k omega and
// 1 - omega in FP32, the update multiplies rho-bar and j by
// om1 + sum_k a_k (moment weights) = 1 + O(1e-7) every step: a constant bias
// that grows linearly in time (8e-8 per step in the velocity of a smooth flow,
// the largest round-off term of an FP32 step). Here omega is moved to
// omega' = 36 w with w = rint(omega 2^25 / 36) 2^-25 (|omega' - omega| <= 5.4e-7
// for omega in (0, 2]): w, 2w, 12w, 36w and 1 - 36w are then exact in FP32
// (9 rint(...) < 2^24), a_0 = 12w, a_1 = 2w, a_2 = w are t_k omega' exactly, and
// the factor is om1 + 36w = 1 with no rounding. q is formed in FP64, so omega'
// is the grid point nearest to omega.
__host__ __device__ inline BgkRates<float> exact_rates(float omega) {
    const float q = (float)rint((double)omega * (33554432.0 / 36.0));
    const float w = q * 0x1p-25f;
    const float om = 36.0f * w;
    return {1.0f - om, 12.0f * w, 2.0f * w, w, om};
}

template<typename T, bool EXACT>
__device__ __forceinline__ BgkRates<T> bgk_rates(T omega) {
    if constexpr (EXACT) return exact_rates(omega);
    else return palabos_rates<T>(omega);
}

// BGK relaxation in i / i+9 pairs (dynamicsTemplates3D.h:2289-2368
// bgk_ma2_collision_base), with the rates `r`. Returns |u|^2.
template<typename T>
__device__ __forceinline__ T bgk(T* f, T rhoBar, const T j[3], const BgkRates<T>& r, T invRho) {
    const T om1 = r.om1;
    const T a0 = r.a0;
    const T a1 = r.a1;
    const T a2 = r.a2;
    const T jSqr = j[0] * j[0] + j[1] * j[1] + j[2] * j[2];
    const T kx = (T)3 * j[0];
    const T ky = (T)3 * j[1];
    const T kz = (T)3 * j[2];
    const T kxx = invRho / (T)2 * kx * kx;
    const T kyy = invRho / (T)2 * ky * ky;
    const T kzz = invRho / (T)2 * kz * kz;
    const T kxy = invRho * kx * ky;
    const T kxz = invRho * kx * kz;
    const T kyz = invRho * ky * kz;
    const T C1 = rhoBar + invRho * (T)3 * jSqr;
    T C2, C3;
    C3 = -kxx - kyy - kzz;
    f[0] *= om1; f[0] += a0 * (C1 + C3);
    // (lower index, C2, C3, weight class) for the nine pairs
#define PLBAMD_PAIR(I, C2E, C3E, A)                          \
    C2 = (C2E); C3 = (C3E);                                  \
    f[I] *= om1; f[I] += A * (C1 + C2 + C3);                 \
    f[I + 9] *= om1; f[I + 9] += A * (C1 - C2 + C3);
    PLBAMD_PAIR(1, -kx, -kyy - kzz, a1)
    PLBAMD_PAIR(2, -ky, -kxx - kzz, a1)
    PLBAMD_PAIR(3, -kz, -kxx - kyy, a1)
    PLBAMD_PAIR(4, -kx - ky, kxy - kzz, a2)
    PLBAMD_PAIR(5, -kx + ky, -kxy - kzz, a2)
    PLBAMD_PAIR(6, -kx - kz, kxz - kyy, a2)
    PLBAMD_PAIR(7, -kx + kz, -kxz - kyy, a2)
    PLBAMD_PAIR(8, -ky - kz, kyz - kxx, a2)
    PLBAMD_PAIR(9, -ky + kz, -kyz - kxx, a2)
#undef PLBAMD_PAIR
    return invRho * invRho * jSqr;
}

// Off-equilibrium populations from PiNeq, one value per i / i+9 pair
// (dynamicsTemplates3D.h:1902-1973 neqPiD3Q19).
template<typename T>
__device__ __forceinline__ void pi_to_fneq(const T pi[6], T p[10]) {
    const T xx = pi[0], xy = pi[1], xz = pi[2], yy = pi[3], yz = pi[4], zz = pi[5];
    const T t3 = (T)1. / (T)3, t23 = (T)2. / (T)3;
    p[0] = (T)3. / (T)2 * (-(T)1. / (T)3 * xx - t3 * yy - t3 * zz);
    p[1] = (T)1. / (T)4 * (t23 * xx - t3 * yy - t3 * zz);
    p[2] = (T)1. / (T)4 * (-(T)1. / (T)3 * xx + t23 * yy - t3 * zz);
    p[3] = (T)1. / (T)4 * (-(T)1. / (T)3 * xx - t3 * yy + t23 * zz);
    p[4] = (T)1. / (T)8 * (t23 * xx + t23 * yy - t3 * zz + (T)2 * xy);
    p[5] = (T)1. / (T)8 * (t23 * xx + t23 * yy - t3 * zz - (T)2 * xy);
    p[6] = (T)1. / (T)8 * (t23 * xx - t3 * yy + t23 * zz + (T)2 * xz);
    p[7] = (T)1. / (T)8 * (t23 * xx - t3 * yy + t23 * zz - (T)2 * xz);
    p[8] = (T)1. / (T)8 * (-(T)1. / (T)3 * xx + t23 * yy + t23 * zz + (T)2 * yz);
    p[9] = (T)1. / (T)8 * (-(T)1. / (T)3 * xx + t23 * yy + t23 * zz - (T)2 * yz);
}

// Regularized BGK (dynamicsTemplates3D.h:2500-2554 rlb_collision). Returns |u|^2.
template<typename T>
__device__ __forceinline__ T rlb(T* f, T rhoBar, T invRho, const T j[3], const T pi[6], T omega) {
    const T jSqr = j[0] * j[0] + j[1] * j[1] + j[2] * j[2];
    T p[10];
    pi_to_fneq(pi, p);
#pragma unroll
    for (int i = 0; i < 19; ++i) f[i] = feq(i, rhoBar, invRho, j, jSqr) + ((T)1 - omega) * p[i <= 9 ? i : i - 9];
    return jSqr * invRho * invRho;
}

// Generic Pi -> f_neq (offEquilibriumTemplates.h:91-113), used by regularize().
template<typename T>
__device__ __forceinline__ T pi_to_fneq_generic(int i, const T pi[6]) {
    T fNeq = T();
    int iPi = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        fNeq += pi[iPi] * (ci(i, a) * ci(i, a) - cs2<T>());
        ++iPi;
#pragma unroll
        for (int b = a + 1; b < 3; ++b) {
            fNeq += pi[iPi] * (T)2 * ci(i, a) * ci(i, b);
            ++iPi;
        }
    }
    fNeq *= weight<T>(i) * (T)3 * (T)3 / (T)2;
    return fNeq;
}

// Guo forcing term, amplitude 1 (externalForceTemplates3D.h:73-150 addGuoForce).
template<typename T>
__device__ __forceinline__ void add_guo_force(T* f, const T u[3], const T F[3], T omega) {
    const T mu = (T)1 * ((T)1 - omega / (T)2);
    const T o6 = (T)1 / (T)6;
    const T o12 = (T)1 / (T)12;
    f[0] += -mu * (F[0] * u[0] + F[1] * u[1] + F[2] * u[2]);
    f[1] += o6 * mu * (F[0] * (-(T)1 + 2 * u[0]) - F[1] * u[1] - F[2] * u[2]);
    f[2] += -o6 * mu * (F[0] * u[0] + F[1] * ((T)1 - 2 * u[1]) + F[2] * u[2]);
    f[3] += -o6 * mu * (F[0] * u[0] + F[1] * u[1] + F[2] * ((T)1 - 2 * u[2]));
    f[4] += o12 * mu * (F[0] * (-(T)1 + 2 * u[0] + 3 * u[1]) + F[1] * (-(T)1 + 2 * u[1] + 3 * u[0]) - F[2] * u[2]);
    f[5] += o12 * mu * (F[0] * (-(T)1 + 2 * u[0] - 3 * u[1]) + F[1] * ((T)1 + 2 * u[1] - 3 * u[0]) - F[2] * u[2]);
    f[6] += o12 * mu * (F[0] * (-(T)1 + 2 * u[0] + 3 * u[2]) - F[1] * u[1] + F[2] * (-(T)1 + 2 * u[2] + 3 * u[0]));
    f[7] += o12 * mu * (F[0] * (-(T)1 + 2 * u[0] - 3 * u[2]) - F[1] * u[1] + F[2] * ((T)1 + 2 * u[2] - 3 * u[0]));
    f[8] += -o12 * mu * (F[0] * u[0] + F[1] * ((T)1 - 2 * u[1] - 3 * u[2]) + F[2] * ((T)1 - 2 * u[2] - 3 * u[1]));
    f[9] += -o12 * mu * (F[0] * u[0] + F[1] * ((T)1 - 2 * u[1] + 3 * u[2]) + F[2] * (-(T)1 - 2 * u[2] + 3 * u[1]));
    f[10] += o6 * mu * (F[0] * ((T)1 + 2 * u[0]) - F[1] * u[1] - F[2] * u[2]);
    f[11] += -o6 * mu * (F[0] * u[0] + F[1] * (-(T)1 - 2 * u[1]) + F[2] * u[2]);
    f[12] += -o6 * mu * (F[0] * u[0] + F[1] * u[1] + F[2] * (-(T)1 - 2 * u[2]));
    f[13] += o12 * mu * (F[0] * ((T)1 + 2 * u[0] + 3 * u[1]) + F[1] * ((T)1 + 2 * u[1] + 3 * u[0]) - F[2] * u[2]);
    f[14] += o12 * mu * (F[0] * ((T)1 + 2 * u[0] - 3 * u[1]) + F[1] * (-(T)1 + 2 * u[1] - 3 * u[0]) - F[2] * u[2]);
    f[15] += o12 * mu * (F[0] * ((T)1 + 2 * u[0] + 3 * u[2]) - F[1] * u[1] + F[2] * ((T)1 + 2 * u[2] + 3 * u[0]));
    f[16] += o12 * mu * (F[0] * ((T)1 + 2 * u[0] - 3 * u[2]) - F[1] * u[1] + F[2] * (-(T)1 + 2 * u[2] - 3 * u[0]));
    f[17] += -o12 * mu * (F[0] * u[0] + F[1] * (-(T)1 - 2 * u[1] - 3 * u[2]) + F[2] * (-(T)1 - 2 * u[2] - 3 * u[1]));
    f[18] += -o12 * mu * (F[0] * u[0] + F[1] * (-(T)1 - 2 * u[1] + 3 * u[2]) + F[2] * ((T)1 - 2 * u[2] + 3 * u[1]));
}

// Bulk collision selector. DYN: 0 BGK, 1 RBGK. FORCED: Guo body force and/or
// Brinkman porosity active (a = alpha of the cell). Returns |u|^2 and rhoBar.
// EXACT (the 16-bit storage modes, T = float): BGK relaxes with exact_rates,
// and omega' replaces omega in the Guo term and the Brinkman drag as well.
//   BGK + Guo follows GuoExternalForceBGKdynamics::collide
//   (externalForceDynamics.hh:215-232; velocity :59-74). RBGK + Guo and the
//   Brinkman damping u <- (1 - alpha) u of the equilibrium velocity have no
//   Palabos counterpart and are defined by this library.
// `drag` receives the momentum the Brinkman term removes from the cell in the
// collision: omega * rho * (u - u_damped) for BGK, rho * (u - u_damped) for
// regularized BGK (which sets the first moment to j_eq), u the undamped
// half-force velocity.
template<typename T, int DYN, bool FORCED, bool EXACT = false>
__device__ __forceinline__ T collide_bulk(T* f, T omegaIn, const T F[3], T a, T& rhoBarOut, T drag[3]) {
    const BgkRates<T> rates = bgk_rates<T, EXACT && DYN == 0>(omegaIn);
    const T omega = rates.omega;
    if constexpr (!FORCED) {
        T rhoBar, j[3];
        T uSqr;
        if constexpr (DYN == 0) {
            rhobar_j(f, rhoBar, j);
            uSqr = bgk(f, rhoBar, j, rates, inv_rho(rhoBar));
        } else {
            T pi[6];
            rhobar_j_pineq(f, rhoBar, j, pi);
            uSqr = rlb(f, rhoBar, inv_rho(rhoBar), j, pi, omega);
        }
        rhoBarOut = rhoBar;
        return uSqr;
    } else {
        T rhoBar, rhoBarP, j[3], pi[6];
        if constexpr (DYN == 0) {
            rhoBar = rhobar_sum(f);
            rhobar_j(f, rhoBarP, j);
        } else {
            rhobar_j_pineq(f, rhoBarP, j, pi);
            rhoBar = rhoBarP;
        }
        const T irP = inv_rho(rhoBarP);
        T u[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) u[c] = j[c] * irP + F[c] / (T)2;
        const T damp = (T)1 - a;
        const T rho = rhoBar + (T)1;
        T jeq[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const T u0 = u[c];
            u[c] = u0 * damp;
            jeq[c] = rho * u[c];
            // momentum the damping removes in this collision: BGK relaxes by omega
            // towards j_eq, regularized BGK replaces the first moment by j_eq
            drag[c] = (DYN == 0) ? omega * (rho * u0 - jeq[c]) : (rho * u0 - jeq[c]);
        }
        T uSqr;
        if constexpr (DYN == 0) uSqr = bgk(f, rhoBar, jeq, rates, inv_rho(rhoBar));
        else uSqr = rlb(f, rhoBar, inv_rho(rhoBar), jeq, pi, omega);
        if (F[0] != (T)0 || F[1] != (T)0 || F[2] != (T)0) add_guo_force(f, u, F, omega);
        rhoBarOut = rhoBar;
        return uSqr;
    }
}

// Collision with the moments of an external field instead of the populations'
// own: BGKdynamics::collideExternal (isoThermalDynamics.hh:332-338,
// bgk_ma2_collision with the given rhoBar and j) and
// RegularizedBGKdynamics::collideExternal (isoThermalDynamics.hh:1604-1613,
// PiNeq of the populations about the given moments, then rlb_collision). No
// statistics, as there. EXACT: the 16-bit storage modes' BGK rates.
template<typename T, int DYN, bool EXACT>
__device__ __forceinline__ void collide_external(T* f, T omega, T rhoBar, const T j[3]) {
    if constexpr (DYN == 0) {
        (void)bgk(f, rhoBar, j, bgk_rates<T, EXACT>(omega), inv_rho(rhoBar));
    } else {
        T pi[6];
        pineq_given(f, rhoBar, j, pi);
        (void)rlb(f, rhoBar, inv_rho(rhoBar), j, pi, omega);
    }
}

}  // namespace plbamd
