// Data processors that run after the stream of every step (and once from
// MultiBlock3D::initialize, multiBlock3D.cpp:330-332), on the device:
//
//   * the Neumann-type copies of a boundary's stored value from the cell one
//     step inside it (CopyVelocityFunctional3D, CopyTangentialVelocityFunctional3D,
//     CopyNormalVelocityFunctional3D, CopyDensityFunctional3D;
//     boundaryCondition/neumannCondition3D.hh:116-252), installed for the
//     outflow, neumann, freeslip and normalOutflow boundary types
//     (boundaryInstantiator3D.h:1042-1063, 1097-1102, 1139-1157, 1216-1232);
//   * the finite-difference closures of external velocity edges and corners
//     (OuterVelocityEdgeFunctional3D, OuterVelocityCornerFunctional3D;
//     finiteDifferenceBoundaryProcessor3D.hh:130-329): density from the
//     neighbours inside, the strain rate from one-sided differences of the
//     neighbours' velocities (centred along an edge), and the cell rebuilt as
//     equilibrium plus the off-equilibrium part of that strain rate
//     (IsoThermalBulkDynamics::regularize, isoThermalDynamics.hh:55-69).
//
// Palabos runs the processors one after the other in the order the boundary
// objects were added. The device runs them in four launches (phases): the
// copies of faces and edges, the edge closures, the copies of corners, the
// corner closures. The host checks that no cell is read in the phase that
// writes it and that the phases keep the order of the objects that touch
// (engine.hip, rebuild), so the results are those of the sequential order.
//
// Every neighbour is resolved on the host into a reference: a special record
// (plane, edge or corner cell) or a plain cell of the block; the neighbour's
// density and velocity are what its dynamics reports (Cell::computeDensity /
// computeVelocity), from its populations at the current layout parity. A
// plane cell of a wrapped factory is read completed, as its own processor
// leaves it (wrappedLocalBoundaryProcessor3D.hh:41-62).
#pragma once

#include "kernels.h"

namespace plbamd {

enum ProcOp : uint8_t {
    PROC_COPY_VELOCITY = 1,    // CopyVelocityFunctional3D
    PROC_COPY_TANGENTIAL = 2,  // CopyTangentialVelocityFunctional3D: normal components zeroed
    PROC_COPY_NORMAL = 3,      // CopyNormalVelocityFunctional3D: tangential components zeroed
    PROC_COPY_DENSITY = 4,     // CopyDensityFunctional3D
    PROC_EDGE = 5,             // OuterVelocityEdgeFunctional3D
    PROC_CORNER = 6,           // OuterVelocityCornerFunctional3D
};

// Neighbour reference: a special record (flag set) or a block-local cell index.
constexpr uint32_t kNbSpecial = 0x80000000u;

struct ProcCell {
    uint32_t special;   // special record of the processed cell (position, block, stored value)
    uint8_t op;         // ProcOp
    int8_t n[3];        // outward normal of the boundary (zero along an edge's axis)
    uint32_t nb[4];     // copies: nb[0] = cell - n
                        // edges: nb[0] = cell - n1 e1, nb[1] = cell - n2 e2 (e1, e2 the axes after the
                        //        edge's axis), nb[2] / nb[3] = cell +- e along the edge
                        // corners: nb[0..2] = cell - n_x e_x, - n_y e_y, - n_z e_z
};
static_assert(sizeof(ProcCell) == 24, "processor record is 24 bytes");

struct ProcArgs {
    void* arena;
    const BlockGeom* blocks;
    const SpecialCell* special;
    double* values;          // stored boundary values (velocity: 3, density: 1), rewritten by the copies
    const uint8_t* tags;     // bulk tags or null
    const uint32_t* walls;   // link-wall masks or null
    const ProcCell* cells;
    uint32_t begin, end;     // the records of this phase
    int parity;              // layout parity of the state the processors act on
    double omega;
};

struct NbCell {
    bool special;
    SpecialCell sc;          // valid when special
    int x, y, z;
    uint64_t cell;
    uint32_t wall;
};

__device__ __forceinline__ NbCell resolve_nb(const ProcArgs& a, const BlockGeom& g, uint32_t ref) {
    NbCell c{};
    if (ref & kNbSpecial) {
        c.special = true;
        c.sc = a.special[ref & ~kNbSpecial];
        c.x = c.sc.x; c.y = c.sc.y; c.z = c.sc.z;
        c.cell = c.sc.cell;
        c.wall = c.sc.wall;
    } else {
        c.special = false;
        c.cell = ref;
        c.z = (int)(ref % (uint32_t)g.nz);
        c.y = (int)((ref / (uint32_t)g.nz) % (uint32_t)g.ny);
        c.x = (int)(ref / ((uint32_t)g.ny * (uint32_t)g.nz));
        const uint8_t tag = a.tags ? a.tags[g.cellOffset + ref] : (uint8_t)BT_BULK;
        c.wall = (tag == BT_WALL_NEAR && a.walls) ? a.walls[g.cellOffset + ref] : 0u;
    }
    return c;
}

template<typename S>
__device__ __forceinline__ void load_nb(const ProcArgs& a, const BlockGeom& g, const NbCell& c,
                                        typename Store<S>::T f[19]) {
    const S* base = reinterpret_cast<const S*>(a.arena) + g.offset;
#pragma unroll
    for (int i = 0; i < 19; ++i)
        f[i] = Store<S>::load(base + pop_index(g, a.parity, c.x, c.y, c.z, i, a.parity ? c.wall : 0u, c.cell));
}

// Cell::computeDensity of a neighbour: fullRho of the Dirichlet rhoBar for a
// plane cell (boundaryDynamics.hh:640-679, StoreDensityDynamics :229-233),
// fullRho(get_rhoBar) otherwise (core/dynamics.hh:278-281, boundaryDynamics.hh:82-87).
template<typename S>
__device__ typename Store<S>::T nb_density(const ProcArgs& a, const BlockGeom& g, uint32_t ref) {
    using T = typename Store<S>::T;
    const NbCell c = resolve_nb(a, g, ref);
    T f[19];
    load_nb<S>(a, g, c, f);
    if (c.special && (c.sc.info & 3u) == SK_PLANE) {
        const double* v = a.values + c.sc.value;
        if ((c.sc.info >> 7) & 1u) complete_plane(f, c.sc.info, v);
        T rhoBar, u[3];
        plane_macroscopic_any(f, c.sc.info, v, rhoBar, u);
        return rhoBar + (T)1;
    }
    return rhobar_sum(f) + (T)1;
}

// Cell::computeVelocity of a neighbour: the stored velocity of edge, corner and
// velocity-plane cells (boundaryDynamics.hh:324-330), j / rho of a pressure
// plane's closure (:776-835), j / rho of the populations otherwise (compute_uLb,
// momentTemplates3D.h:138-144, with the sequentially summed rhoBar).
template<typename S>
__device__ void nb_velocity(const ProcArgs& a, const BlockGeom& g, uint32_t ref, typename Store<S>::T u[3]) {
    using T = typename Store<S>::T;
    const NbCell c = resolve_nb(a, g, ref);
    if (c.special) {
        const uint32_t kind = c.sc.info & 3u;
        if (kind == SK_STORE_VELOCITY || (kind == SK_PLANE && ((c.sc.info >> 3) & 1u) == PLBAMD_VELOCITY)) {
            const double* v = a.values + c.sc.value;
            u[0] = (T)v[0]; u[1] = (T)v[1]; u[2] = (T)v[2];
            return;
        }
    }
    T f[19];
    load_nb<S>(a, g, c, f);
    if (c.special) {
        const double* v = a.values + c.sc.value;
        if ((c.sc.info >> 7) & 1u) complete_plane(f, c.sc.info, v);
        T rhoBar;
        plane_macroscopic_any(f, c.sc.info, v, rhoBar, u);
        return;
    }
    T rb, j[3];
    rhobar_j(f, rb, j);
    const T ir = inv_rho(rhobar_sum(f));
    u[0] = j[0] * ir; u[1] = j[1] * ir; u[2] = j[2] * ir;
}

// pi from the velocity gradient g[a][b] = d_a u_b and the density
// (finiteDifferenceBoundaryProcessor3D.hh:162-172, :288-296), order xx xy xz yy yz zz.
template<typename T>
__device__ __forceinline__ void strain_to_pi(const T g[3][3], T rho, T omega, T pi[6]) {
    const T sToPi = -rho / (T)3 / omega;
    pi[0] = (T)2 * g[0][0] * sToPi;
    pi[3] = (T)2 * g[1][1] * sToPi;
    pi[5] = (T)2 * g[2][2] * sToPi;
    pi[1] = (g[0][1] + g[1][0]) * sToPi;
    pi[2] = (g[0][2] + g[2][0]) * sToPi;
    pi[4] = (g[1][2] + g[2][1]) * sToPi;
}

// IsoThermalBulkDynamics::regularize (isoThermalDynamics.hh:55-69) with
// j = rho u and rhoBar = rho - 1, written to the cell's own populations.
template<typename S>
__device__ void regularize_store(const ProcArgs& a, const BlockGeom& g, const SpecialCell& sc,
                                 typename Store<S>::T rho, const typename Store<S>::T u[3],
                                 const typename Store<S>::T pi[6]) {
    using T = typename Store<S>::T;
    T j[3];
    for (int c = 0; c < 3; ++c) j[c] = rho * u[c];
    const T jSqr = j[0] * j[0] + j[1] * j[1] + j[2] * j[2];
    const T rhoBar = rho - (T)1;
    const T ir = inv_rho(rhoBar);
    T f[19];
    f[0] = feq(0, rhoBar, ir, j, jSqr) + pi_to_fneq_generic(0, pi);
#pragma unroll
    for (int i = 1; i <= 9; ++i) {
        f[i] = feq(i, rhoBar, ir, j, jSqr);
        f[i + 9] = feq(i + 9, rhoBar, ir, j, jSqr);
        const T fn = pi_to_fneq_generic(i, pi);
        f[i] += fn;
        f[i + 9] += fn;
    }
    S* base = reinterpret_cast<S*>(a.arena) + g.offset;
#pragma unroll
    for (int i = 0; i < 19; ++i)
        Store<S>::store(base + pop_index(g, a.parity, sc.x, sc.y, sc.z, i, a.parity ? sc.wall : 0u, sc.cell), f[i]);
}

template<typename S>
__global__ void __launch_bounds__(256) boundary_processors(ProcArgs a) {
    using T = typename Store<S>::T;
    const uint32_t k = a.begin + blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.end) return;
    const ProcCell p = a.cells[k];
    const SpecialCell sc = a.special[p.special];
    const BlockGeom g = a.blocks[sc.block];
    double* v = a.values + sc.value;
    const T omega = (T)a.omega;
    if (p.op == PROC_COPY_DENSITY) {
        v[0] = (double)nb_density<S>(a, g, p.nb[0]);
    } else if (p.op <= PROC_COPY_NORMAL) {
        T u[3];
        nb_velocity<S>(a, g, p.nb[0], u);
        for (int c = 0; c < 3; ++c) {
            if (p.op == PROC_COPY_TANGENTIAL && p.n[c] != 0) u[c] = T();
            if (p.op == PROC_COPY_NORMAL && p.n[c] == 0) u[c] = T();
        }
        v[0] = (double)u[0]; v[1] = (double)u[1]; v[2] = (double)u[2];
    } else if (p.op == PROC_EDGE) {
        const int plane = p.n[0] == 0 ? 0 : (p.n[1] == 0 ? 1 : 2);
        const int d1 = (plane + 1) % 3, d2 = (plane + 2) % 3;
        const T n1 = (T)(-p.n[d1]), n2 = (T)(-p.n[d2]);   // -orientation of the one-sided differences
        const T rho10 = nb_density<S>(a, g, p.nb[0]);
        const T rho01 = nb_density<S>(a, g, p.nb[1]);
        const T rho = (T)0.5 * (rho01 + rho10);
        const T u0[3] = {(T)v[0], (T)v[1], (T)v[2]};
        T u10[3], u01[3], up[3], um[3];
        nb_velocity<S>(a, g, p.nb[0], u10);
        nb_velocity<S>(a, g, p.nb[1], u01);
        nb_velocity<S>(a, g, p.nb[2], up);
        nb_velocity<S>(a, g, p.nb[3], um);
        T gr[3][3];
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            const T along = (up[b] - um[b]) / (T)2;       // fd::ctl_diff
            const T g1 = n1 * (-u0[b] + u10[b]);          // -orientation * fd::o1_fwd_diff
            const T g2 = n2 * (-u0[b] + u01[b]);
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                if (r == plane) gr[r][b] = along;
                else if (r == d1) gr[r][b] = g1;
                else gr[r][b] = g2;
            }
        }
        T pi[6];
        strain_to_pi(gr, rho, omega, pi);
        regularize_store<S>(a, g, sc, rho, u0, pi);
    } else {   // PROC_CORNER
        const T rho100 = nb_density<S>(a, g, p.nb[0]);
        const T rho010 = nb_density<S>(a, g, p.nb[1]);
        const T rho001 = nb_density<S>(a, g, p.nb[2]);
        const T rho = (T)1 / (T)3 * (rho001 + rho010 + rho100);
        const T u0[3] = {(T)v[0], (T)v[1], (T)v[2]};
        T un[3][3];
        nb_velocity<S>(a, g, p.nb[0], un[0]);
        nb_velocity<S>(a, g, p.nb[1], un[1]);
        nb_velocity<S>(a, g, p.nb[2], un[2]);
        T gr[3][3];
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int b = 0; b < 3; ++b) gr[r][b] = (T)(-p.n[r]) * (-u0[b] + un[r][b]);
        T pi[6];
        strain_to_pi(gr, rho, omega, pi);
        regularize_store<S>(a, g, sc, rho, u0, pi);
    }
}

}  // namespace plbamd
