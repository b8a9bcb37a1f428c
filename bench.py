"""D3Q19 collide-and-stream benchmark on one MI355X.

Workload (the metric's headline configuration): periodic Taylor-Green vortex
on a 512^3 D3Q19 lattice, BGK, omega = 1.8, u0 = 0.04, synthetic initial state
f = f_eq(rho=1, u_TG) built on the device. A step is one collide-and-stream
pass over the whole lattice (all 134,217,728 cells), statistics off (as
Palabos's own timing runs toggle them off). The primary line is FP64; the FP32
and FP16-storage runs of the same workload are reported under "variants";
"adjoint" gives the reverse-sweep / forward time ratio of the hand-written
adjoint on the Brinkman porosity channel (128 x 64 x 64, 500 steps, checkpoint
every 50, FP64), with the default recording budget (segments recorded during
the forward sweep when device memory holds them) and with every segment
recomputed.

value        MLUPS = cells * steps / wall seconds / 1e6 over the timed steps,
             populations resident in HBM before the timed region starts.
roofline     the collide-and-stream launches: algorithmic bytes per launch
             (read + write 19 populations per cell) / average launch duration,
             measured with HIP events on the stream the kernels run on; one
             launch per step (even and odd steps alternate two kernels).
             "traffic" is the HBM bytes per launch from rocprofv3 counters
             (profiles/bulk_step_tg512_counters.txt).
cpu_baseline the compiled Palabos reference (oracle/_ref/ref_driver, serial) on a
             bounded sample (periodic BGK FP64 128^3): "value" is the aggregate
             of one independent serial process per host core of this job's
             share (at most 16), each on its own lattice; "single_core" is one
             process alone; the host's CPU count and model are stated. The
             oracle restatement (1 core) stands in when the reference binary
             is absent.

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
           "f16": "plbamd::bulk_even_linear (even) / plbamd::bulk_pairs (odd)",
           "bf16": "plbamd::bulk_even_linear (even) / plbamd::bulk_pairs (odd)"}


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
    steps, checkpoint every 50: forward (with checkpoints, or records) and
    reverse-sweep (recompute + adjoint) times from HIP events, best of `repeats`,
    with the default recording budget and with none (every segment recomputed)."""
    import palabos_amd as plb
    from palabos_amd import setups
    nx, ny, nz, u_max = 128, 64, 64, 0.04
    alpha = setups.brinkman_blobs((nx, ny, nz), n_blobs=32, seed=4)
    alpha[0] = 0.0
    alpha[-1] = 0.0
    prof = setups.poiseuille_profile(ny, nz, u_max)
    vel = np.zeros(((ny - 2) * (nz - 2), 3))
    vel[:, 0] = prof.reshape(-1)
    def best_of(budget):
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
                lat.setAdjointRecordBytes(budget)
                J, _, _, (ms_f, ms_b) = lat.adjointGradient(500, 50, 1, nx - 2, 1.0, 1.0)
                kept = lat.adjointRecordInfo()[0]
            finally:
                lat.close()
            if best is None or ms_f + ms_b < best[0] + best[1]:
                best = (ms_f, ms_b, J, kept)
        return best

    ms_f, ms_b, J, kept = best_of(-1)
    ms_f0, ms_b0, J0, _ = best_of(0)
    cells_steps = nx * ny * nz * 500
    return {"workload": "brinkman_channel_128x64x64_500steps_checkpoint50_f64", "forward_ms": ms_f,
            "reverse_ms": ms_b, "ratio": ms_b / ms_f, "forward_mlups": cells_steps / ms_f / 1e3,
            "objective": J, "segments": 10, "segments_recorded_forward": kept,
            "forward_ms_full_recompute": ms_f0, "reverse_ms_full_recompute": ms_b0,
            "ratio_full_recompute": ms_b0 / ms_f0, "same_objective_full_recompute": J0 == J}


def host_cpu():
    """CPU count of this job (affinity) and of the machine, and the CPU model."""
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return {"nproc": usable, "machine_cpus": os.cpu_count(), "cpu_model": model}


def cpu_baseline():
    ref = os.path.join(ROOT, "oracle", "_ref", "ref_driver")
    n, steps = 128, 200   # about 10-15 s of serial CPU work per process
    host = host_cpu()
    if os.path.exists(ref):
        cmd = [ref, "--bench", "double", str(n), str(steps)]
        single = json.loads(subprocess.run(cmd, capture_output=True, text=True, timeout=600,
                                           check=True).stdout.strip().splitlines()[-1])
        # all cores of this job's share: independent serial processes, each its own lattice
        procs_n = max(1, min(16, host["nproc"]))
        procs = [subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, text=True)
                 for _ in range(procs_n)]
        each = [json.loads(p.communicate(timeout=900)[0].strip().splitlines()[-1]) for p in procs]
        total = sum(r["mlups"] for r in each)
        return {"value": total, "unit": "MLUPS", "cores": procs_n, "kind": "reference",
                "sample": f"compiled Palabos (serial MultiBlockLattice3D, BGK, FP64) periodic {n}^3, {steps} steps "
                          f"after 1 warm-up step: {procs_n} independent processes at once, one per core "
                          f"(slowest {max(r['seconds'] for r in each):.1f} s)",
                "per_process_mlups": [r["mlups"] for r in each],
                "single_core": {"value": single["mlups"], "cores": 1, "seconds": single["seconds"]},
                "host": host}
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
            "sample": f"oracle restatement, periodic BGK FP64 {n}^3, {steps} steps, {sec:.1f} s", "host": host}


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
    """A plain device-to-device copy of known size (torch's copy kernel, 4 GiB
    read + 4 GiB written): the reference that scales rocprofv3's FETCH_SIZE /
    WRITE_SIZE counters to bytes, and the practical streaming ceiling (its
    read + write bytes over its duration, from HIP events on torch's stream,
    where the copy runs)."""
    import torch
    src = torch.ones(n_doubles, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)   # untimed first touch
    torch.cuda.synchronize()
    ms = []
    for _ in range(repeats):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        dst.copy_(src)
        e1.record()
        e1.synchronize()
        ms.append(e0.elapsed_time(e1))
    best = min(ms)
    del src, dst
    torch.cuda.empty_cache()
    # the library's probes: a 16-byte nontemporal copy of the same size, and the
    # even step's data movement over the 512^3 FP64 lattice's bytes (19 slots in
    # place) in 8-, 4- and 16-byte elements per thread
    from palabos_amd import native
    probes = {}
    for pattern, nbytes in (("copy", 2 * n_doubles * 8), ("slots", 304 * 512 ** 3), ("slots4", 304 * 512 ** 3),
                            ("slots16", 304 * 512 ** 3)):
        p_ms, moved, gbs = native.bandwidth_probe(pattern, nbytes, repeats)
        probes[pattern] = {"bytes_per_launch": moved, "ms_best": p_ms, "gbs": gbs, "frac_of_peak": gbs / HBM_PEAK_GBS}
    print(json.dumps({"copy_calibration_bytes_read": n_doubles * 8, "copy_calibration_bytes_written": n_doubles * 8,
                      "launches": repeats + 1, "event_ms": ms,
                      "copy_gbs_best": 2 * n_doubles * 8 / (best * 1e-3) / 1e9,
                      "copy_frac_of_peak": 2 * n_doubles * 8 / (best * 1e-3) / 1e9 / HBM_PEAK_GBS,
                      "probes": probes}))


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
