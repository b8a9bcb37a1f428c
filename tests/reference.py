"""Runs a Case through the compiled Palabos reference (oracle/_ref/ref_driver).

Test infrastructure: used by tests/golden/make_golden.py and by the CPU tests
that compare the oracle with live Palabos when oracle/_ref is built.
"""
import os
import subprocess
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "oracle", "_ref", "ref_driver")


def available():
    return os.path.exists(DRIVER)


def run(case, init_f64, dumps, bcstyle="dynamics", blocks=None, timeout=1800, fields=False, block_bcs=None,
        family=None, external=None):
    """Returns ({step: populations in Palabos's T}, {step: stats row}); with
    fields, also {step: [nx,ny,nz,7] analysis fields} (computeDensity,
    computeRhoBar, computeVelocity x y z, computeKineticEnergy,
    computeVelocityNorm) as a third item. block_bcs: [(kind 'vel' | 'press',
    bc type name, per-cell values of the whole lattice)] installed with
    set{Velocity,Pressure}ConditionOnBlockBoundaries; family: the boundary
    factory when the case has no planes of its own. external: the coupling
    amplitude a of the external-moment cycle (BoxRhoBarJfunctional3D, j_x +=
    a (1 + rhoBar), ExternalRhoJcollideAndStream3D) that replaces
    collideAndStream; then {step: [nx,ny,nz,4] rhoBar jx jy jz field} is
    returned as the last item."""
    T = {"f64": "double", "f32": "float"}[case.storage]
    with tempfile.TemporaryDirectory() as tmp:
        lines = [f"T {T}", f"size {case.nx} {case.ny} {case.nz}", f"dyn {case.dynamics}",
                 f"omega {case.omega!r}",
                 "periodic " + " ".join(str(int(bool(p))) for p in case.periodic),
                 f"bcstyle {bcstyle}"]
        fams = {p.family for p in case.planes if getattr(p, "shape", "plane") == "plane"}
        if family:
            fams |= {family}
        assert len(fams) <= 1, "one boundary family per case"
        if fams:
            lines.append(f"bcfamily {fams.pop()}")
        init_path = os.path.join(tmp, "init.bin")
        np.ascontiguousarray(init_f64, dtype=np.float64).tofile(init_path)
        lines.append(f"init {init_path}")
        if case.tags is not None:
            tp = os.path.join(tmp, "tags.bin")
            np.ascontiguousarray(case.tags, dtype=np.uint8).tofile(tp)
            lines.append(f"tags {tp}")
        bct = {"dirichlet": 0, "neumann": 1, "freeslip": 2, "density": 3, "outflow": 4, "normalOutflow": 5}
        for k, p in enumerate(case.planes):
            vp = os.path.join(tmp, f"bc{k}.bin")
            np.ascontiguousarray(p.values, dtype=np.float64).tofile(vp)
            t = bct[getattr(p, "bc_type", "dirichlet")]
            shape = getattr(p, "shape", "plane")
            if shape == "edge":
                d1, d2 = (p.dir + 1) % 3, (p.dir + 2) % 3
                lines.append(f"edge {p.dir} {p.normal[d1]} {p.normal[d2]} " + " ".join(map(str, p.box)) +
                             f" {vp} {t}")
            elif shape == "corner":
                x0, _, y0, _, z0, _ = p.box
                lines.append("corner " + " ".join(map(str, p.normal)) + f" {x0} {y0} {z0} {vp} {t}")
            else:
                kind = "vel" if p.kind == "velocity" else "press"
                lines.append(f"bc {kind} {p.dir} {p.ori} " + " ".join(map(str, p.box)) + f" {vp} {t}")
        for k, (kind, t, vals) in enumerate(block_bcs or []):
            vp = os.path.join(tmp, f"blockbc{k}.bin")
            np.ascontiguousarray(vals, dtype=np.float64).tofile(vp)
            lines.append(f"blockbc {kind} {bct[t]} {vp}")
        if any(case.force):
            lines.append("force " + " ".join(repr(float(v)) for v in case.force))
        for b in blocks or []:
            lines.append("block " + " ".join(map(str, b)))
        if fields:
            lines.append("fields 1")
        if external is not None:
            lines.append(f"external {float(external)!r}")
        lines.append("dump " + " ".join(map(str, dumps)))
        out = os.path.join(tmp, "out")
        lines.append(f"out {out}")
        cfile = os.path.join(tmp, "case.txt")
        with open(cfile, "w") as fh:
            fh.write("\n".join(lines) + "\n")
        subprocess.run([DRIVER, cfile], check=True, timeout=timeout, capture_output=True)
        dt = np.float64 if T == "double" else np.float32
        pops = {s: np.fromfile(f"{out}_s{s}.bin", dtype=dt).reshape(case.shape + (19,)) for s in dumps}
        stats = {}
        for line in open(f"{out}_stats.txt"):
            v = line.split()
            stats[int(v[0])] = {"stored_avg_rho": float(v[1]), "stored_avg_energy": float(v[2]),
                                "avg_rho": float(v[3]), "avg_energy": float(v[4]),
                                "force": np.array([float(v[5]), float(v[6]), float(v[7])])}
        extra = []
        if fields:
            extra.append({s: np.fromfile(f"{out}_fields_s{s}.bin", dtype=dt).reshape(case.shape + (7,))
                          for s in dumps})
        if external is not None:
            extra.append({s: np.fromfile(f"{out}_rhobarj_s{s}.bin", dtype=dt).reshape(case.shape + (4,))
                          for s in dumps if s > 0})
        return (pops, stats, *extra)


def checkpoint(mode, T, forced, shape, src, dst, enforce_uint=False, timeout=300):
    """Runs the reference's saveBinaryBlock ("save": cell records -> file) or
    loadBinaryBlock ("load": file -> cell records) through ref_driver."""
    nx, ny, nz = shape
    cmd = [DRIVER, "--checkpoint", mode, T, "forced" if forced else "plain", str(nx), str(ny), str(nz), src, dst,
           "uint" if enforce_uint else "size"]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout)
    if r.returncode != 0:
        raise RuntimeError(f"ref_driver --checkpoint failed ({r.returncode}): {r.stderr[-2000:]}")
