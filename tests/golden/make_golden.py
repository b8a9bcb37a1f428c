"""Writes the golden vectors of tests/golden/ by running the compiled Palabos
reference (oracle/_ref/ref_driver, built by oracle/Makefile.ref from the
reference's own sources) on small seeded cases.

Each fixture <name>.npz holds the case description (JSON), the plane values
and tags it needs, the SHA-256 of the FP64 initial state (regenerated from its
seed by tests/cases.py:random_state), and Palabos's populations and
statistics after each dumped step. cavity64_f64.npz (the 64^3 lid-driven
cavity, 1000 steps) keeps a summary only: statistics, per-population sums,
the mid-plane velocity, and the SHA-256 of the full population array.

Run from the repository root:  python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import reference  # noqa: E402
from golden_cases import GOLDEN, build  # noqa: E402


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main(only=None):
    if not reference.available():
        raise SystemExit("oracle/_ref/ref_driver is missing: build it with make -f oracle/Makefile.ref")
    for name, spec in GOLDEN.items():
        if only and name not in only:
            continue
        case, init, meta = build(name)
        res = reference.run(case, init, spec["dumps"], bcstyle=spec.get("bcstyle", "dynamics"),
                            blocks=spec.get("blocks"), external=spec.get("external"))
        pops, stats = res[0], res[1]
        out = {"case_json": np.array(json.dumps(meta)), "init_sha256": np.array(sha(init))}
        for s in spec["dumps"]:
            st = stats[s]
            out[f"stats_s{s}"] = np.array([st["stored_avg_rho"], st["stored_avg_energy"], st["avg_rho"],
                                           st["avg_energy"], *st["force"]], dtype=np.float64)
            if spec.get("summary"):
                f = pops[s].astype(np.float64)
                out[f"sha256_s{s}"] = np.array(sha(pops[s]))
                out[f"popsum_s{s}"] = f.reshape(-1, 19).sum(axis=0)
                mid = case.ny // 2
                fm = f[:, mid, :, :]
                sXM = fm[..., [1, 4, 5, 6, 7]].sum(-1)
                sXP = fm[..., [10, 13, 14, 15, 16]].sum(-1)
                sYM = fm[..., [2, 4, 8, 9, 14]].sum(-1)
                sYP = fm[..., [5, 11, 13, 17, 18]].sum(-1)
                sZM = fm[..., [3, 6, 8, 16, 18]].sum(-1)
                sZP = fm[..., [7, 9, 12, 15, 17]].sum(-1)
                rho = fm.sum(-1) + 1.0
                out[f"midplane_u_s{s}"] = np.stack([(sXP - sXM) / rho, (sYP - sYM) / rho, (sZP - sZM) / rho], -1)
            else:
                out[f"s{s}"] = pops[s]
            if spec.get("external") is not None:
                out[f"rhobarj_s{s}"] = res[2][s]
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, **out)
        print(f"{name}: {os.path.getsize(path) / 1024:.0f} KiB")


if __name__ == "__main__":
    main(set(sys.argv[1:]) or None)
