"""The benchmark configurations as runnable studies on the engine.

    python -m palabos_amd.studies taylor_green [--n 512 --steps 2000]
    python -m palabos_amd.studies sphere_drag  [--steps 5000 --storage f64 --ramp 0 --average-last 2000]
    python -m palabos_amd.studies porous       [--steps 20000 --storage f16]
    python -m palabos_amd.studies porous_adjoint [--steps 2000 --storage f16]
    python -m palabos_amd.studies porous_walls_adjoint [--steps 2000 --storage f16]

Each prints one JSON object. Everything runs on the device through the
native library; the host only builds the geometry, reads back reductions and
velocity slabs, and does the FP64 bookkeeping of the error norms.
"""
import argparse
import os
import sys
import json
import math
import time

import numpy as np

from . import lattice as _lat
from . import setups


def _slabs(n, cells_per_slab=1 << 22, ny=None, nz=None):
    ny = ny or n
    nz = nz or n
    step = max(1, min(n, cells_per_slab // (ny * nz)))
    for x0 in range(0, n, step):
        yield x0, min(n, x0 + step)


# ---------------------------------------------------------------------------
# Taylor-Green vortex: precision study

def taylor_green_2d_velocity(n, u0, x0, x1, nz):
    """u = u0 (sin kx cos ky, -cos kx sin ky, 0), k = 2 pi / n: an exact
    Navier-Stokes solution whose velocity decays as exp(-2 nu k^2 t)."""
    k = 2.0 * np.pi / n
    c = np.arange(n) * k
    s, co = np.sin(c), np.cos(c)
    u = np.zeros((x1 - x0, n, nz, 3))
    u[..., 0] = u0 * s[x0:x1, None, None] * co[None, :, None]
    u[..., 1] = -u0 * co[x0:x1, None, None] * s[None, :, None]
    return u


def _tg_lattice(n, storage, omega, u0, mode):
    lat = _lat.Block3D(n, n, n, _lat.D3Q19, _lat.BGK(omega), storage=storage, periodic=True)
    for x0, x1 in _slabs(n):
        u = (setups.taylor_green_velocity(n, u0, x0, x1) if mode == "3d"
             else taylor_green_2d_velocity(n, u0, x0, x1, n))
        lat.initializeAtEquilibrium(1.0, u, box=(x0, x1 - 1, 0, n - 1, 0, n - 1))
    lat.toggleInternalStatistics(False)
    return lat


def taylor_green_precision(n=512, steps=2000, omega=1.8, u0=0.04, storages=("f32", "f16", "bf16"),
                           modes=("3d", "2d")):
    """Runs the periodic Taylor-Green vortex in FP64 and in each reduced storage;
    reports, per storage, the relative velocity L2 error (and the absolute RMS
    velocity error) against the FP64 run, split into its component along the
    FP64 field (signed: a systematic amplitude drift) and the rest (noise),
    and (all storages) against the analytic decay: the exact 2D-mode solution
    u(t) = u(0) exp(-2 nu k^2 t), and for the 3D vortex the early-time kinetic
    energy E0 exp(-6 nu k^2 t) (an approximation: the 3D vortex has no closed form)."""
    nu = (1.0 / omega - 0.5) / 3.0
    k = 2.0 * np.pi / n
    out = {"lattice": [n, n, n], "steps": steps, "omega": omega, "u0": u0, "nu": nu, "modes": {}}
    for mode in modes:
        ref = _tg_lattice(n, "f64", omega, u0, mode)
        res = {}
        try:
            e0 = ref.averageEnergy()
            t0 = time.perf_counter()
            ref.collideAndStream(steps)
            e64 = ref.averageEnergy()
            res["f64"] = {"seconds": time.perf_counter() - t0, "energy_ratio": e64 / e0}
            decay_u = math.exp(-2.0 * nu * k * k * steps)
            decay_e = math.exp((-4.0 if mode == "2d" else -6.0) * nu * k * k * steps)
            for storage in ("f64",) + tuple(storages):
                lat = ref if storage == "f64" else _tg_lattice(n, storage, omega, u0, mode)
                try:
                    if storage != "f64":
                        lat.collideAndStream(steps)
                        res[storage] = {"energy_ratio": lat.averageEnergy() / e0}
                    d_ref = s_ref = d_an = s_an = p_ref = 0.0
                    for x0, x1 in _slabs(n):
                        box = (x0, x1 - 1, 0, n - 1, 0, n - 1)
                        u = lat.velocity(box).reshape(-1, 3)
                        if storage != "f64":
                            u64 = ref.velocity(box).reshape(-1, 3)
                            d_ref += float(((u - u64) ** 2).sum())
                            s_ref += float((u64 ** 2).sum())
                            p_ref += float(((u - u64) * u64).sum())
                        if mode == "2d":
                            ua = taylor_green_2d_velocity(n, u0, x0, x1, n).reshape(-1, 3) * decay_u
                            d_an += float(((u - ua) ** 2).sum())
                            s_an += float((ua ** 2).sum())
                    r = res[storage]
                    if storage != "f64":
                        r["velocity_l2_vs_f64"] = math.sqrt(d_ref / s_ref)          # relative
                        r["velocity_rms_error_vs_f64"] = math.sqrt(d_ref / n ** 3)  # absolute, lattice units
                        amp = p_ref / s_ref
                        r["velocity_amplitude_error_vs_f64"] = amp
                        r["velocity_l2_orthogonal_vs_f64"] = math.sqrt(max(0.0, d_ref / s_ref - amp * amp))
                    if mode == "2d":
                        r["velocity_l2_vs_analytic"] = math.sqrt(d_an / s_an)
                    r["energy_error_vs_analytic"] = r["energy_ratio"] / decay_e - 1.0
                    print(f"taylor_green {mode} {storage}: {json.dumps(r)}", file=sys.stderr, flush=True)
                finally:
                    if storage != "f64":
                        lat.close()
            res["analytic_energy_ratio"] = decay_e
        finally:
            ref.close()
        out["modes"][mode] = res
    return out


# ---------------------------------------------------------------------------
# Channel flow past a sphere: drag coefficient

def sphere_drag(nx=256, ny=128, nz=128, radius=16.0, center=(64.0, 64.0, 64.0), u_max=0.05, re=100.0,
                steps=5000, storage="f64", average_last=2000, sample_every=10, ramp=0):
    """Zou-He Poiseuille inlet (x = 0, u_max), Zou-He pressure outlet (rho = 1),
    half-way bounce-back side walls (the domain edges), a momentum-exchange
    sphere. Re = u_max 2 r / nu. ramp = 0 is the impulsive start (the full
    profile from the first step); otherwise the inlet velocity rises as sin^2
    over the first `ramp` steps. An impulsive start fills the channel with
    pressure waves that the velocity and pressure planes reflect (a quarter-wave
    resonator of period 4 nx / c_s steps, damped only by the viscosity), so
    the force oscillates about its mean for far longer than 5000 steps.
    C_d = 2 F_x / (rho u^2 pi r^2) with u = u_max and with the mean inlet
    velocity; F_x sampled every `sample_every` steps over the last
    `average_last` steps. Reported with the band of the samples (min, max,
    standard deviation) and the mean over the whole oscillation periods the
    window holds (upward crossings of the mean)."""
    nu = u_max * 2.0 * radius / re
    omega = 1.0 / (3.0 * nu + 0.5)
    lat = _lat.Block3D(nx, ny, nz, _lat.D3Q19, _lat.BGK(omega), storage=storage)
    try:
        tags = np.zeros((nx, ny, nz), np.uint8)
        tags[setups.sphere_mask((nx, ny, nz), center, radius)] = _lat.native.TAG_MOMENTUM_EXCHANGE
        lat.setTags(tags)
        prof = setups.poiseuille_profile(ny, nz, u_max)
        vel = np.zeros(((ny - 2) * (nz - 2), 3))
        vel[:, 0] = prof.reshape(-1)
        p_in = lat.addVelocityBoundary(0, -1, (0, 0, 1, ny - 2, 1, nz - 2))
        lat.setBoundaryVelocity(p_in, vel)
        p_out = lat.addPressureBoundary(0, 1, (nx - 1, nx - 1, 1, ny - 2, 1, nz - 2))
        lat.setBoundaryDensity(p_out, np.ones((ny - 2) * (nz - 2)))
        lat.toggleInternalStatistics(False)
        main = steps - average_last
        ms = 0.0
        done = 0
        while done < min(ramp, main):
            m = min(10, ramp - done)
            lat.setBoundaryVelocity(p_in, vel * math.sin(0.5 * math.pi * (done + m) / ramp) ** 2)
            ms += lat.collideAndStreamTimed(m)
            done += m
        lat.setBoundaryVelocity(p_in, vel)
        ms += lat.collideAndStreamTimed(main - done)
        lat.toggleInternalStatistics(True)
        fx = []
        done = main
        while done < steps:
            m = min(sample_every, steps - done)
            lat.collideAndStream(m)
            done += m
            fx.append(lat.getStatistics()["force"][0])
        fx = np.array(fx)
        u_mean = float(prof.sum()) / ((ny - 2) * (nz - 2))
        area = math.pi * radius * radius
        cd = 2.0 / (u_max ** 2 * area)
        up = np.nonzero((fx[:-1] < fx.mean()) & (fx[1:] >= fx.mean()))[0] + 1
        periods = len(up) - 1
        whole = fx[up[0]:up[-1]] if periods > 0 else fx
        return {"lattice": [nx, ny, nz], "storage": storage, "steps": steps, "re": re, "omega": omega, "ramp": ramp,
                "sphere_cells": int((tags != 0).sum()), "fx_mean": float(fx.mean()), "fx_std": float(fx.std()),
                "cd_u_max": cd * fx.mean(), "cd_u_max_min": cd * float(fx.min()), "cd_u_max_max": cd * float(fx.max()),
                "cd_u_max_std": cd * float(fx.std()),
                "swing_pct": 50.0 * float(fx.max() - fx.min()) / abs(float(fx.mean())),
                "oscillation_periods": periods,
                "period_steps": sample_every * float(up[-1] - up[0]) / periods if periods > 0 else None,
                "cd_u_max_whole_periods": cd * float(whole.mean()),
                "cd_u_mean": 2.0 * fx.mean() / (u_mean ** 2 * area),
                "mlups": nx * ny * nz * main / (ms * 1e-3) / 1e6, "timed_steps": main}
    finally:
        lat.close()


# ---------------------------------------------------------------------------
# Porous medium of random spheres, 60 uneven blocks: permeability

def porous_geometry(shape=(384, 256, 256), count=400, r_range=(6.0, 14.0), seed=5, scale=1.0):
    centres, radii = setups.random_spheres(shape, count, r_range, seed, periodic=(True, False, False))
    return setups.spheres_mask(shape, centres, radii, scale=scale, periodic=(True, False, False))


def porous_permeability(shape=(384, 256, 256), count=400, steps=20000, storage="f16", omega=1.7, g=1e-6,
                        scale=1.0, blocks=None, block_seed=11, check_every=500, keep_blocks=False):
    """Regularized BGK driven by a Guo body force g along x through spheres made
    of link-wise walls (half-way bounce-back links), periodic in x, walls on the
    y and z faces, 60 uneven blocks (5 x 4 x 3 pieces of 32-128 cells). Darcy
    permeability k = nu <u_x> / g with <u_x> the superficial velocity (mean over
    all cells, solid cells counting zero). The device runs the tiling as one
    block unless keep_blocks (then the 60 blocks exchange their envelopes)."""
    nu = (1.0 / omega - 0.5) / 3.0
    solid = porous_geometry(shape, count, scale=scale)
    if blocks is None:
        blocks = setups.uneven_blocks(shape, (5, 4, 3), 32, 128, block_seed)
    lat = _lat.Block3D(*shape, _lat.D3Q19, _lat.RegularizedBGK(omega), storage=storage,
                       periodic=(True, False, False), force=(g, 0.0, 0.0), blocks=blocks, keep_blocks=keep_blocks)
    try:
        tags = np.zeros(shape, np.uint8)
        tags[solid] = _lat.native.TAG_WALL_LINK
        lat.setTags(tags)
        lat.toggleInternalStatistics(False)
        history = []
        done = 0
        ms = 0.0
        while done < steps:
            m = min(check_every, steps - done)
            ms += lat.collideAndStreamTimed(m)
            done += m
            mom = lat.reduceMoments()
            history.append(float(mom["sum_u"][0]) / mom["cells"])
        ux = history[-1]
        return {"lattice": list(shape), "storage": storage, "blocks": len(blocks), "device_blocks": lat.blockCount(),
                "steps": steps,
                "porosity": float(1.0 - solid.mean()), "omega": omega, "g": g, "radius_scale": scale,
                "superficial_ux": ux, "permeability": nu * ux / g,
                "ux_history": history, "mlups": float(np.prod(shape)) * steps / (ms * 1e-3) / 1e6}
    finally:
        lat.close()


def brinkman_spheres(shape, centres, radii, scale=1.0, width=1.0, periodic=(True, False, False)):
    """Smooth porosity of a set of spheres: alpha = 1 - prod_b (1 - sigma((s r_b - d_b(x)) / w)),
    sigma the logistic function, and its derivative d alpha / d s =
    (1 - alpha) sum_b sigma_b r_b / w (each sphere evaluated on the box where
    sigma exceeds ~1e-6)."""
    log_open = np.zeros(shape)      # sum_b log(1 - sigma_b) = -sum_b softplus(z_b)
    dsum = np.zeros(shape)          # sum_b sigma_b r_b / w
    margin = 14.0 * width
    for c, r in zip(centres, radii):
        rs = r * scale
        lo = np.floor(c - rs - margin).astype(int)
        hi = np.ceil(c + rs + margin).astype(int)
        idx = []
        for a in range(3):
            v = np.arange(lo[a], hi[a] + 1)
            if periodic[a]:
                d = v - c[a]
                v = v % shape[a]
            else:
                keep = (v >= 0) & (v < shape[a])
                v = v[keep]
                d = v - c[a]
            idx.append((v, d))
        (vx, dx), (vy, dy), (vz, dz) = idx
        dist = np.sqrt(dx[:, None, None] ** 2 + dy[None, :, None] ** 2 + dz[None, None, :] ** 2)
        z = (rs - dist) / width
        sig = 1.0 / (1.0 + np.exp(-z))
        ix = np.ix_(vx, vy, vz)
        log_open[ix] -= np.logaddexp(0.0, z)
        dsum[ix] += sig * r / width
    open_frac = np.exp(log_open)
    return 1.0 - open_frac, open_frac * dsum


def porous_adjoint(shape=(384, 256, 256), count=400, r_range=(6.0, 14.0), seed=5, steps=2000, checkpoint_every=50,
                   window=1000, storage="f16", omega=1.7, g=1e-5, width=1.0, blocks=None, block_seed=11, fd_eps=1e-2,
                   check_fd=True, keep_blocks=False, record_bytes=-1, compare_full_recompute=True):
    """Darcy permeability of the porous medium with the spheres as a smooth
    Brinkman porosity, and its gradient with respect to a global sphere-radius
    scale s by the adjoint: K = nu <j_x> / g averaged over the last `window` of
    `steps` steps from rest, dK/ds = nu / g sum_x dJ/dalpha(x) dalpha/ds(x);
    checked against central finite differences in s (K(s +- eps) from two more
    runs). Regularized BGK with a Guo force, periodic in x, walls on the y and z
    faces, 60 uneven blocks, checkpoints every `checkpoint_every` steps.
    record_bytes: the device memory for segments recorded during the forward
    sweep (-1: the free memory beyond a reserve; 0: every segment recomputed);
    with compare_full_recompute the gradient run is repeated with every segment
    recomputed, for its cost and to confirm the same gradient."""
    nu = (1.0 / omega - 0.5) / 3.0
    centres, radii = setups.random_spheres(shape, count, r_range, seed, periodic=(True, False, False))
    if blocks is None:
        blocks = setups.uneven_blocks(shape, (5, 4, 3), 32, 128, block_seed)

    def run(scale, want_gradient, budget=record_bytes):
        alpha, dads = brinkman_spheres(shape, centres, radii, scale=scale, width=width)
        lat = _lat.Block3D(*shape, _lat.D3Q19, _lat.RegularizedBGK(omega), storage=storage,
                           periodic=(True, False, False), force=(g, 0.0, 0.0), blocks=blocks,
                           keep_blocks=keep_blocks)
        try:
            lat.setPorosity(alpha)
            lat.setAdjointRecordBytes(budget)
            J, ga, _, (ms_f, ms_b) = lat.adjointGradient(steps, checkpoint_every, 0, 0, 0.0, 0.0,
                                                         weight_flux=1.0, first_step=steps - window)
            kept = lat.adjointRecordInfo()
        finally:
            lat.close()
        dK = nu / g * float((ga * dads).sum()) if want_gradient else None
        print(f"[porous_adjoint] scale {scale} record budget {budget}: forward {ms_f:.1f} ms, reverse {ms_b:.1f} ms, "
              f"{kept[0]} segments recorded forward", file=sys.stderr, flush=True)
        return nu * J / g, dK, ms_f, ms_b, float(1.0 - alpha.mean()), kept

    K, dK, ms_f, ms_b, porosity, kept = run(1.0, True)
    out = {"lattice": list(shape), "storage": storage, "blocks": len(blocks), "steps": steps,
           "checkpoint_every": checkpoint_every, "window": window, "omega": omega, "g": g, "width": width,
           "porosity": porosity, "permeability": K, "dK_dscale_adjoint": dK,
           "forward_ms": ms_f, "reverse_ms": ms_b, "reverse_over_forward": ms_b / ms_f,
           "segments_recorded_forward": kept[0], "segments": -(-steps // checkpoint_every),
           "recorded_bytes": kept[1]}
    if compare_full_recompute and kept[0] > 0:
        K0, dK0, f0, b0, _, _ = run(1.0, True, budget=0)
        out.update({"forward_ms_full_recompute": f0, "reverse_ms_full_recompute": b0,
                    "reverse_over_forward_full_recompute": b0 / f0,
                    "same_result_full_recompute": bool(K0 == K and dK0 == dK)})
    if check_fd:
        Kp = run(1.0 + fd_eps, False)[0]
        Km = run(1.0 - fd_eps, False)[0]
        fd = (Kp - Km) / (2 * fd_eps)
        out.update({"dK_dscale_fd": fd, "fd_eps": fd_eps, "relative_difference": abs(dK - fd) / abs(fd)})
    return out


def porous_walls_adjoint(shape=(384, 256, 256), count=400, steps=2000, checkpoint_every=50, window=1000,
                         storage="f16", omega=1.7, g=1e-5, s0=0.1, blocks=None, block_seed=11, fd_eps=1e-2,
                         check_fd=True, record_bytes=-1):
    """The permeability of the link-wall sphere medium (the forward's medium of
    `porous_permeability`: half-way bounce-back links, regularized BGK, Guo
    force, periodic in x, 60 uneven blocks) with a smooth Brinkman porosity
    alpha = s alpha0 laid over its fluid (alpha0: 32 seeded Gaussian blobs
    through a sigmoid, zero in the solid), and dK/ds at s = s0 by the adjoint
    through the walls: dK/ds = nu / g sum_x dJ/dalpha(x) alpha0(x), checked
    against central finite differences in s."""
    nu = (1.0 / omega - 0.5) / 3.0
    solid = porous_geometry(shape, count)
    alpha0 = setups.brinkman_blobs(shape, n_blobs=32, seed=4)
    alpha0[solid] = 0.0
    if blocks is None:
        blocks = setups.uneven_blocks(shape, (5, 4, 3), 32, 128, block_seed)
    tags = np.zeros(shape, np.uint8)
    tags[solid] = _lat.native.TAG_WALL_LINK

    def run(s, want_gradient):
        lat = _lat.Block3D(*shape, _lat.D3Q19, _lat.RegularizedBGK(omega), storage=storage,
                           periodic=(True, False, False), force=(g, 0.0, 0.0), blocks=blocks)
        try:
            lat.setTags(tags)
            lat.setPorosity(s * alpha0)
            lat.setAdjointRecordBytes(record_bytes)
            J, ga, _, (ms_f, ms_b) = lat.adjointGradient(steps, checkpoint_every, 0, 0, 0.0, 0.0,
                                                         weight_flux=1.0, first_step=steps - window)
            kept = lat.adjointRecordInfo()[0]
        finally:
            lat.close()
        dK = nu / g * float((ga * alpha0).sum()) if want_gradient else None
        print(f"[porous_walls_adjoint] s {s}: forward {ms_f:.1f} ms, reverse {ms_b:.1f} ms, {kept} segments recorded "
              f"forward", file=sys.stderr, flush=True)
        return nu * J / g, dK, ms_f, ms_b, kept

    K, dK, ms_f, ms_b, kept = run(s0, True)
    out = {"lattice": list(shape), "storage": storage, "blocks": len(blocks), "steps": steps,
           "checkpoint_every": checkpoint_every, "window": window, "omega": omega, "g": g, "s0": s0,
           "porosity": float(1.0 - solid.mean()), "permeability": K, "dK_ds_adjoint": dK,
           "forward_ms": ms_f, "reverse_ms": ms_b, "reverse_over_forward": ms_b / ms_f,
           "segments_recorded_forward": kept, "segments": -(-steps // checkpoint_every)}
    if check_fd:
        Kp = run(s0 + fd_eps, False)[0]
        Km = run(s0 - fd_eps, False)[0]
        fd = (Kp - Km) / (2 * fd_eps)
        out.update({"dK_ds_fd": fd, "fd_eps": fd_eps, "relative_difference": abs(dK - fd) / abs(fd)})
    return out


def checkpoint_throughput(n=256, storage="f32", T="f32", directory=None):
    """Saves an n^3 lattice as a Palabos checkpoint file and loads it back into
    a second lattice (wall-clock times of both calls, file size, and that the
    populations came back unchanged). The device kernels' own durations come
    from a kernel trace of this run."""
    import tempfile
    rng = np.random.RandomState(2)
    shape = (n, n, n)
    a = _lat.Block3D(*shape, _lat.D3Q19, _lat.BGK(1.7), storage=storage, periodic=True)
    b = _lat.Block3D(*shape, _lat.D3Q19, _lat.BGK(1.7), storage=storage, periodic=True)
    try:
        for x0 in range(0, n, 32):   # seeded state, uploaded in slabs
            box = (x0, min(n, x0 + 32) - 1, 0, n - 1, 0, n - 1)
            a.setPopulations((rng.standard_normal((box[1] - x0 + 1, n, n, 19)) * 0.02).astype(np.float32), box)
        with tempfile.TemporaryDirectory(dir=directory) as tmp:
            path = os.path.join(tmp, "lattice.dat")
            t0 = time.perf_counter()
            a.saveBinaryBlock(path, T)
            t1 = time.perf_counter()
            b.loadBinaryBlock(path, T)
            t2 = time.perf_counter()
            size = os.path.getsize(path)
        same = all(np.array_equal(a.getPopulations((x0, min(n, x0 + 32) - 1, 0, n - 1, 0, n - 1)),
                                  b.getPopulations((x0, min(n, x0 + 32) - 1, 0, n - 1, 0, n - 1)))
                   for x0 in range(0, n, 32))
    finally:
        a.close()
        b.close()
    data = n ** 3 * 19 * (8 if T == "f64" else 4)
    return {"shape": list(shape), "storage": storage, "file_element": T, "cell_bytes": data, "file_bytes": size,
            "save_s": t1 - t0, "load_s": t2 - t1, "identical_after_load": bool(same)}


def fields_throughput(n=256, storage="f32", repeats=3):
    """Analysis fields (density, rhoBar, velocity, kinetic energy, velocity
    norm) of an n^3 lattice after a few steps: wall-clock time of each call
    (including the 56 bytes per cell copied to the host) and a consistency
    check; the device kernel's own duration comes from a kernel trace."""
    shape = (n, n, n)
    lat = _lat.Block3D(*shape, _lat.D3Q19, _lat.BGK(1.7), storage=storage, periodic=True)
    try:
        x = np.arange(n) * 2 * np.pi / n
        u = np.zeros(shape + (3,))
        u[..., 0] = 0.03 * np.sin(x)[:, None, None] * np.cos(x)[None, :, None]
        u[..., 1] = -0.03 * np.cos(x)[:, None, None] * np.sin(x)[None, :, None]
        lat.initializeAtEquilibrium(rho=1.0, u=u)
        lat.collideAndStream(4)
        times = []
        for _ in range(repeats):
            t0 = time.perf_counter()
            f = lat.analysisFields()
            times.append(time.perf_counter() - t0)
        mean_energy = float(f[..., 5].mean())
        avg_energy = lat.averageEnergy()
    finally:
        lat.close()
    return {"shape": list(shape), "storage": storage, "call_s": times, "mean_kinetic_energy": mean_energy,
            "average_energy": avg_energy}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("study", choices=["taylor_green", "sphere_drag", "porous", "porous_adjoint",
                                      "porous_walls_adjoint", "checkpoint", "fields"])
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--storage", default=None)
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--keep-blocks", action="store_true", help="porous studies: keep the 60 blocks on the device")
    ap.add_argument("--fd-eps", type=float, default=1e-2, help="adjoint studies: finite-difference step in the scale")
    ap.add_argument("--record-bytes", type=int, default=-1,
                    help="porous_adjoint: device bytes for forward-recorded segments (-1 auto, 0 none)")
    ap.add_argument("--ramp", type=int, default=0, help="sphere_drag: inlet ramp in steps (0: impulsive start)")
    ap.add_argument("--average-last", type=int, default=2000, help="sphere_drag: force-averaging window in steps")
    args = ap.parse_args()
    from . import native
    native.init(0)
    if args.study == "taylor_green":
        r = taylor_green_precision(n=args.n, steps=args.steps or 2000)
    elif args.study == "sphere_drag":
        r = sphere_drag(steps=args.steps or 5000, storage=args.storage or "f64", ramp=args.ramp,
                        average_last=args.average_last)
    elif args.study == "porous":
        r = porous_permeability(steps=args.steps or 20000, storage=args.storage or "f16", scale=args.scale,
                                keep_blocks=args.keep_blocks)
    elif args.study == "porous_walls_adjoint":
        r = porous_walls_adjoint(steps=args.steps or 2000, storage=args.storage or "f16", record_bytes=args.record_bytes,
                                 fd_eps=args.fd_eps)
    elif args.study == "fields":
        r = fields_throughput(n=args.n if args.n != 512 else 256, storage=args.storage or "f32")
    elif args.study == "checkpoint":
        r = checkpoint_throughput(n=args.n if args.n != 512 else 256, storage=args.storage or "f32")
    else:
        r = porous_adjoint(steps=args.steps or 2000, storage=args.storage or "f16", keep_blocks=args.keep_blocks,
                           record_bytes=args.record_bytes, fd_eps=args.fd_eps)
    print(json.dumps(r))


if __name__ == "__main__":
    main()
