"""Writes tests/golden/units.json: the IncomprFlowParam quantities
(core/units.h:39-95) that the compiled Palabos reference (oracle/_ref/ref_driver
--units) prints for a few parameter sets, both constructors.

Run from the repository root:  python tests/golden/make_units_golden.py
"""
import json
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
DRIVER = os.path.join(ROOT, "oracle", "_ref", "ref_driver")

# physU latU Re physLen N lx ly lz t  (physU = 0: the six-argument constructor)
PARAMS = [
    [0, 0.02, 100, 0, 128, 1, 1, 1, 2.5],
    [0, 0.05, 1000, 0, 100, 2, 0.5, 0.5, 10.0],
    [0, 0.01, 10, 0, 63, 1, 1, 0, 0.3],
    [1.0, 0.02, 100, 2.0, 100, 1, 1, 1, 1.0],
    [2.5, 0.04, 400, 0.1, 40, 5, 2, 2, 0.77],
    [0.3, 0.05, 37.5, 3.0, 51, 0.5, 0.25, 1.0, 12.0],
]


def main():
    rows = []
    for p in PARAMS:
        out = subprocess.run([DRIVER, "--units"] + [repr(float(v)) if not isinstance(v, int) or k in (0,) else str(v)
                                                    for k, v in enumerate(p)],
                             capture_output=True, text=True, check=True).stdout.strip()
        rows.append({"params": p, "line": out})
    with open(os.path.join(HERE, "units.json"), "w") as f:
        json.dump(rows, f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main()
