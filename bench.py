"""D3Q19 collide-and-stream benchmark on one MI355X.

Workload (the metric's headline configuration): periodic Taylor-Green vortex
on a 512^3 D3Q19 lattice, BGK, omega = 1.8, u0 = 0.04, synthetic initial state
f = f_eq(rho=1, u_TG) built on the device. A step is one collide-and-stream
pass over the whole lattice (all 134,217,728 cells), statistics off (as
Palabos's own timing runs toggle them off). The primary line is FP64; the FP32
and FP16-storage runs of the same workload are reported under "variants";
"adjoint" gives the reverse-sweep / forward time ratio of the hand-written
adjoint on the Brinkman porosity channel (128 x 64 x 64, 500 steps, checkpoint
every 50, FP64).

value        MLUPS = cells * steps / wall seconds / 1e6 over the timed steps,
             populations resident in HBM before the timed region starts.
roofline     the collide-and-stream launches: algorithmic bytes per launch
             (read + write 19 populations per cell) / average launch duration,
             measured with HIP events on the stream the kernels run on; one
             launch per step (even and odd steps alternate two kernels).
             "traffic" is the HBM bytes per launch from rocprofv3 counters
             (profiles/bulk_step_tg512_counters.txt).
cpu_baseline the compiled Palabos reference (oracle/_ref/ref_driver, serial,
             1 core) on a bounded sample (periodic BGK 128^3), or the oracle
             restatement when the reference binary is absent.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 512]
                       [--precision f64] [--no-variants] [--no-cpu] [--no-adjoint]
"""
import argparse
import json
import os
import re
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES_PER_CELL = {"f64": 19 * 2 * 8, "f32": 19 * 2 * 4, "f16": 19 * 2 * 2, "bf16": 19 * 2 * 2}
# The launches that make one step at this workload (even / odd steps alternate;
# the roofline figure is their average launch duration).
KERNELS = {"f64": "plbamd::bulk_step (even) / plbamd::bulk_odd_staged (odd)",
           "f32": "plbamd::bulk_step (even) / plbamd::bulk_odd_staged (odd)",
           "f16": "plbamd::bulk_pairs (even and odd)", "bf16": "plbamd::bulk_pairs (even and odd)"}


def taylor_green_init(lat, n, u0):
    """f = f_eq(rho = 1, u) with u = u0 (sin kx cos ky cos kz, -cos kx sin ky cos kz, 0), in x slabs."""
    k = 2.0 * np.pi / n
    c = np.arange(n) * k
    sx, cxv = np.sin(c), np.cos(c)
    slab = max(1, min(n, (1 << 22) // (n * n)))
    for x0 in range(0, n, slab):
        x1 = min(n, x0 + slab)
        xs = slice(x0, x1)
        ux = u0 * sx[xs, None, None] * cxv[None, :, None] * cxv[None, None, :]
        uy = -u0 * cxv[xs, None, None] * sx[None, :, None] * cxv[None, None, :]
        u = np.zeros((x1 - x0, n, n, 3))
        u[..., 0] = ux
        u[..., 1] = uy
        lat.initializeAtEquilibrium(1.0, u, box=(x0, x1 - 1, 0, n - 1, 0, n - 1))


def run_precision(precision, n, steps, warmup, omega=1.8, u0=0.04):
    import palabos_amd as plb
    lat = plb.Block3D(n, n, n, plb.D3Q19, plb.BGK(omega), storage=precision, periodic=True)
    try:
        taylor_green_init(lat, n, u0)
        lat.toggleInternalStatistics(False)
        if warmup:
            lat.collideAndStream(warmup)
        t0 = time.perf_counter()
        ms_events = lat.collideAndStreamTimed(steps)  # blocks until the last step has finished
        wall = time.perf_counter() - t0
        cells = n ** 3
        kernel_ms = ms_events / steps
        achieved = BYTES_PER_CELL[precision] * cells / (kernel_ms * 1e-3) / 1e9
        return {
            "dtype": precision,
            "value": cells * steps / wall / 1e6,
            "ms_per_step": wall * 1e3 / steps,
            "event_ms_per_step": kernel_ms,
            "achieved_gbs": achieved,
            "frac": achieved / HBM_PEAK_GBS,
        }
    finally:
        lat.close()


def adjoint_ratio(repeats=3):
    """Brinkman porosity channel 128 x 64 x 64 (Zou-He Poiseuille inlet u_max = 0.04,
    pressure outlet, omega = 1.6, porosity from 32 seeded Gaussian blobs), 500
    steps, checkpoint every 50: forward (with checkpoints) and reverse-sweep
    (recompute + adjoint) times from HIP events, best of `repeats`."""
    import palabos_amd as plb
    from palabos_amd import setups
    nx, ny, nz, u_max = 128, 64, 64, 0.04
    alpha = setups.brinkman_blobs((nx, ny, nz), n_blobs=32, seed=4)
    alpha[0] = 0.0
    alpha[-1] = 0.0
    prof = setups.poiseuille_profile(ny, nz, u_max)
    vel = np.zeros(((ny - 2) * (nz - 2), 3))
    vel[:, 0] = prof.reshape(-1)
    best = None
    for _ in range(repeats):
        lat = plb.Block3D(nx, ny, nz, plb.D3Q19, plb.BGK(1.6), storage="f64")
        try:
            p_in = lat.addVelocityBoundary(0, -1, (0, 0, 1, ny - 2, 1, nz - 2))
            lat.setBoundaryVelocity(p_in, vel)
            p_out = lat.addPressureBoundary(0, 1, (nx - 1, nx - 1, 1, ny - 2, 1, nz - 2))
            lat.setBoundaryDensity(p_out, np.ones((ny - 2) * (nz - 2)))
            lat.setPorosity(alpha)
            lat.setPopulations(np.zeros((nx, ny, nz, 19)))
            J, _, _, (ms_f, ms_b) = lat.adjointGradient(500, 50, 1, nx - 2, 1.0, 1.0)
        finally:
            lat.close()
        if best is None or ms_f + ms_b < best[0] + best[1]:
            best = (ms_f, ms_b, J)
    ms_f, ms_b, J = best
    cells_steps = nx * ny * nz * 500
    return {"workload": "brinkman_channel_128x64x64_500steps_checkpoint50_f64", "forward_ms": ms_f,
            "reverse_ms": ms_b, "ratio": ms_b / ms_f, "forward_mlups": cells_steps / ms_f / 1e3,
            "objective": J}


def cpu_baseline():
    ref = os.path.join(ROOT, "oracle", "_ref", "ref_driver")
    n, steps = 128, 200   # about 10-15 s of serial CPU work
    if os.path.exists(ref):
        out = subprocess.run([ref, "--bench", "double", str(n), str(steps)], capture_output=True,
                             text=True, timeout=600, check=True).stdout
        r = json.loads(out.strip().splitlines()[-1])
        return {"value": r["mlups"], "unit": "MLUPS", "cores": 1, "kind": "reference",
                "sample": f"compiled Palabos (serial MultiBlockLattice3D, BGK, FP64) periodic "
                          f"{n}^3, {steps} steps after 1 warm-up step, {r['seconds']:.1f} s"}
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from cases import Case
    from oracle import oracle
    n, steps = 64, 30
    case = Case(n, n, n, omega=1.8, periodic=(True, True, True))
    f = np.zeros((n, n, n, 19))
    t0 = time.perf_counter()
    oracle.run(case, f, steps)
    sec = time.perf_counter() - t0
    return {"value": n ** 3 * steps / sec / 1e6, "unit": "MLUPS", "cores": 1, "kind": "port",
            "sample": f"oracle restatement, periodic BGK FP64 {n}^3, {steps} steps, {sec:.1f} s"}


def traffic_from_profile(precision, n):
    """HBM bytes per launch of the bulk kernel from the committed rocprofv3 counter
    summary (profiles/bulk_step_tg512_counters.txt), or None."""
    path = os.path.join(ROOT, "profiles", "bulk_step_tg512_counters.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        m = re.match(r"\s*traffic_bytes_per_launch\s+(\S+)\s+n=(\d+)\s+([0-9.eE+]+)", line)
        if m and m.group(1) == precision and int(m.group(2)) == n:
            return float(m.group(3))
    return None


def copy_calibration(n_doubles=1 << 29, repeats=3):
    """A plain device-to-device copy of known size (torch's copy kernel): the
    reference that scales rocprofv3's FETCH_SIZE / WRITE_SIZE counters to bytes."""
    import torch
    src = torch.ones(n_doubles, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(repeats):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_read": n_doubles * 8, "copy_calibration_bytes_written": n_doubles * 8,
                      "launches": repeats}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--precision", default="f64", choices=["f64", "f32", "f16", "bf16"])
    ap.add_argument("--no-variants", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--adjoint-only", action="store_true", help="print only the adjoint timing object")
    ap.add_argument("--copy-calibration", action="store_true", help="run only the counter-calibration copy")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the engine has no multi-GPU path "
              "(the lattice would shard into x-slabs with per-face halo exchange over RCCL)",
              file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration()
        return
    if args.adjoint_only:
        print(json.dumps({"adjoint": adjoint_ratio()}))
        return

    main_run = run_precision(args.precision, args.n, args.steps, args.warmup)
    variants = []
    if not args.no_variants:
        for p in ("f64", "f32", "f16"):
            if p != args.precision:
                variants.append(run_precision(p, args.n, args.steps, args.warmup))
    for v in variants:
        v["kernel"] = KERNELS[v["dtype"]]
        v["traffic"] = traffic_from_profile(v["dtype"], args.n)
    traffic = traffic_from_profile(args.precision, args.n)
    line = {
        "metric": "D3Q19 MLUPS (FP64/FP32/FP16-store) & % HBM peak; adjoint/forward ratio",
        "value": main_run["value"],
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": main_run["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.precision,
        "data": "synthetic (Taylor-Green initial state built on the device)",
        "config": {"workload": f"taylor_green_{args.n}^3_periodic_bgk", "lattice": [args.n] * 3,
                   "omega": 1.8, "u0": 0.04, "precision": args.precision, "statistics": False},
        "roofline": {"bound": "hbm", "achieved": main_run["achieved_gbs"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": main_run["frac"], "traffic": traffic,
                     "kernel": KERNELS[args.precision], "algorithmic_bytes_per_cell": BYTES_PER_CELL[args.precision],
                     "event_ms_per_launch": main_run["event_ms_per_step"]},
        "variants": variants,
    }
    if not args.no_adjoint:
        line["adjoint"] = adjoint_ratio()
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline()
    print(json.dumps(line))


if __name__ == "__main__":
    main()
