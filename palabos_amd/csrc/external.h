// External-moment processors on the device:
//   BoxRhoBarJfunctional3D (dataProcessors/dataAnalysisFunctional3D.hh:528-561):
//     rhoBar and j of every cell of a box into a field, by each cell's
//     computeRhoBarJ: the partial sums of get_rhoBar_j for the bulk dynamics
//     (momentTemplates3D.h:411-423), rho - 1 and 0 for BounceBack, 0 and 0 for
//     NoDynamics (core/dynamics.hh:880-885, :1061-1066);
//   ExternalRhoJcollideAndStream3D (basicDynamics/dynamicsProcessor3D.hh:41-209):
//     one collide-and-stream step in which every cell collides with the field's
//     moments (collideExternal: the bulk dynamics with the given rhoBar and j,
//     BounceBack as its collide, NoDynamics nothing; no statistics), then streams.
// The field is [4][cells] in the storage's arithmetic type (rhoBar, jx, jy, jz),
// indexed by the block-local cell of a single-block lattice.
#pragma once

#include "kernels.h"

namespace plbamd {

template<typename S, int PAR>
__global__ void __launch_bounds__(256) rhobar_j_box(const void* arena, BoxArgs b, const uint8_t* tags,
                                                    double solidRhoBB, typename Store<S>::T* field,
                                                    uint64_t fstride) {
    using T = typename Store<S>::T;
    const S* f = reinterpret_cast<const S*>(arena) + b.g.offset;
    PLBAMD_BOX_LOOP(b) {
        const BoxCell c = box_cell(b, k);
        const uint8_t tag = tags ? tags[c.cell] : (uint8_t)BT_BULK;
        T rhoBar = T(), j[3] = {T(), T(), T()};
        if (tag == BT_BOUNCE_BACK) {
            rhoBar = (T)solidRhoBB - (T)1;
        } else if (tag != BT_NO_DYNAMICS) {
            T fc[19];
            for (int i = 0; i < 19; ++i) fc[i] = Store<S>::load(f + pop_index(b.g, PAR, c.x, c.y, c.z, i, 0u, c.cell));
            rhobar_j(fc, rhoBar, j);
        }
        field[c.cell] = rhoBar;
        field[fstride + c.cell] = j[0];
        field[2 * fstride + c.cell] = j[1];
        field[3 * fstride + c.cell] = j[2];
    }
}

// Field values of a box to / from host-order staging (rhoBar: 1, j: 3 per cell).
template<typename T>
__global__ void __launch_bounds__(256) field_get(BoxArgs b, const T* field, uint64_t fstride, double* rhoBar,
                                                 double* j) {
    PLBAMD_BOX_LOOP(b) {
        const BoxCell c = box_cell(b, k);
        rhoBar[c.staged] = (double)field[c.cell];
        for (int a = 0; a < 3; ++a) j[3 * c.staged + a] = (double)field[(uint64_t)(a + 1) * fstride + c.cell];
    }
}

template<typename T>
__global__ void __launch_bounds__(256) field_set(BoxArgs b, T* field, uint64_t fstride, const double* rhoBar,
                                                 const double* j) {
    PLBAMD_BOX_LOOP(b) {
        const BoxCell c = box_cell(b, k);
        field[c.cell] = (T)rhoBar[c.staged];
        for (int a = 0; a < 3; ++a) field[(uint64_t)(a + 1) * fstride + c.cell] = (T)j[3 * c.staged + a];
    }
}

// One step of every cell of a single-block lattice with the field's moments, in
// the engine's AA order: even steps read slot i of the cell and write slot
// opp(i); odd steps gather slot opp(i) from x - c_i and scatter to slot i of
// x + c_i, with the same-step reflection at non-periodic edges of the bulk
// kernels. One cell per thread.
template<typename S, int DYN, bool ODD>
__global__ void __launch_bounds__(256) external_step(void* arena, BlockGeom g, const uint8_t* tags,
                                                     const typename Store<S>::T* field, double omega) {
    using T = typename Store<S>::T;
    const uint64_t cells = (uint64_t)g.nx * g.ny * g.nz;
    const uint64_t cell = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    if (cell >= cells) return;
    S* f = reinterpret_cast<S*>(arena) + g.offset;
    const uint32_t plane = (uint32_t)g.ny * (uint32_t)g.nz;
    const uint32_t cc = (uint32_t)cell;
    const int x = (int)(cc / plane), rem = (int)(cc % plane);
    const int y = rem / g.nz, z = rem % g.nz;
    const uint8_t tag = tags ? tags[cell] : (uint8_t)BT_BULK;
    int xs[3], ys[3], zs[3];
    bool vx[3], vy[3], vz[3];
    vx[0] = shift(x, -1, g.nx, g.px, xs[0]); vx[1] = true; xs[1] = x; vx[2] = shift(x, 1, g.nx, g.px, xs[2]);
    vy[0] = shift(y, -1, g.ny, g.py, ys[0]); vy[1] = true; ys[1] = y; vy[2] = shift(y, 1, g.ny, g.py, ys[2]);
    vz[0] = shift(z, -1, g.nz, g.pz, zs[0]); vz[1] = true; zs[1] = z; vz[2] = shift(z, 1, g.nz, g.pz, zs[2]);
    T fc[19];
    fc[0] = Store<S>::load(f + cell);
#pragma unroll
    for (int i = 1; i < 19; ++i) {
        uint64_t src = (uint64_t)i * g.slot + cell;
        if constexpr (ODD) {
            const int ax = 1 - cx(i), ay = 1 - cy(i), az = 1 - cz(i);
            if (vx[ax] && vy[ay] && vz[az])
                src = (uint64_t)opp(i) * g.slot + (((uint64_t)xs[ax] * g.ny + ys[ay]) * g.nz + zs[az]);
        }
        fc[i] = Store<S>::load(f + src);
    }
    if (tag == BT_BOUNCE_BACK) {
#pragma unroll
        for (int i = 1; i <= 9; ++i) { T t = fc[i]; fc[i] = fc[i + 9]; fc[i + 9] = t; }
    } else if (tag != BT_NO_DYNAMICS) {
        const T j[3] = {field[cells + cell], field[2 * cells + cell], field[3 * cells + cell]};
        collide_external<T, DYN, sizeof(S) == 2>(fc, (T)omega, field[cell], j);
    }
    Store<S>::store(f + cell, fc[0]);
#pragma unroll
    for (int i = 1; i < 19; ++i) {
        uint64_t dst = (uint64_t)opp(i) * g.slot + cell;
        if constexpr (ODD) {
            const int ax = 1 + cx(i), ay = 1 + cy(i), az = 1 + cz(i);
            if (vx[ax] && vy[ay] && vz[az])
                dst = (uint64_t)i * g.slot + (((uint64_t)xs[ax] * g.ny + ys[ay]) * g.nz + zs[az]);
        }
        Store<S>::store(f + dst, fc[i]);
    }
}

}  // namespace plbamd
