"""Writes the analysis-field golden vectors of tests/golden/ by running the
compiled Palabos reference (oracle/_ref/ref_driver with "fields 1") on golden
cases: after the dumped step, computeDensity, computeRhoBar, computeVelocity,
computeKineticEnergy and computeVelocityNorm (dataAnalysisWrapper3D.h:695-777)
of the whole lattice, in the lattice's T.

Each fixture fields_<case>.npz holds the golden case name (its state is
rebuilt by tests/golden_cases.py:build), the step, and the fields
[nx][ny][nz][7] in the order above.

Run from the repository root:  python tests/golden/make_fields_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import reference  # noqa: E402
from golden_cases import GOLDEN, build  # noqa: E402

# golden cases (every cell kind: bulk BGK/RBGK, BounceBack, NoDynamics,
# momentum-exchange nodes, Zou-He and regularized planes dynamics-based and
# wrapped, Guo force, multi-block) and the step after which the fields are taken
CASES = {
    "periodic_rbgk_f32": 20,
    "walls_bb_nodyn_f64": 15,
    "cavity_zouhe_f64": 50,
    "cavity_zouhe_wrapped_f64": 50,
    "channel_zouhe_mebb_f64": 30,
    "channel_regularized_bb_rbgk_f64": 30,
    "faces_regularized_swapped_f64": 12,
    "faces_zouhe_f64": 12,
    "guo_f64": 11,
    "guo_f32": 11,
    "multiblock_mebb_f64": 10,
}


def main():
    if not reference.available():
        raise SystemExit("oracle/_ref/ref_driver is missing: build it with make -f oracle/Makefile.ref")
    for name, step in CASES.items():
        spec = GOLDEN[name]
        case, init, meta = build(name)
        _, _, fields = reference.run(case, init, [step], bcstyle=spec.get("bcstyle", "dynamics"),
                                     blocks=spec.get("blocks"), fields=True)
        np.savez_compressed(os.path.join(HERE, f"fields_{name}.npz"), case=name, step=step, fields=fields[step])
        print(name, step, fields[step].shape)


if __name__ == "__main__":
    main()
