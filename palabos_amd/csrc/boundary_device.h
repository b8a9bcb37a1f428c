// Boundary-plane arithmetic on one cell: Zou-He and regularized completion of
// velocity and pressure planes, and the density/velocity such a cell reports.
#pragma once

#include "../../include/palabos_amd.h"
#include "d3q19_device.h"

namespace plbamd {

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

template<typename T, int DIR, int ORI, typename V>
__device__ __forceinline__ void complete_plane_dir(T* f, int family, int bckind, const V* v) {
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

template<typename T, typename V = double>
__device__ __forceinline__ void complete_plane(T* f, uint32_t info, const V* v) {
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
__device__ __forceinline__ void plane_macroscopic_any(const T* f, uint32_t info, const double* v, T& rhoBar, T u[3]) {
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

}  // namespace plbamd
