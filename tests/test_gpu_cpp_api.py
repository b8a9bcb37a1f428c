"""The Palabos-compatible C++ API (palabos_amd/cpp/palabos_amd.hh) and the
co-processor plug-in, driven by C++ programs written like Palabos examples."""
import hashlib
import json
import os
import subprocess

import numpy as np
import pytest

from cases import Case, random_state
from golden_cases import load
from oracle import oracle

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "palabos_amd", "cpp", "bin")
GOLDEN_DIR = os.path.join(ROOT, "tests", "golden")


def _run(args, timeout=300):
    res = subprocess.run([os.path.join(BIN, args[0])] + [str(a) for a in args[1:]], capture_output=True,
                         text=True, timeout=timeout)
    assert res.returncode == 0, res.stderr
    return json.loads(res.stdout.strip().splitlines()[-1])


@pytest.mark.parametrize("name,precision,wrapped", [
    ("cavity_zouhe_f64", "double", False),
    ("cavity_zouhe_f32", "float", False),
    ("cavity_zouhe_wrapped_f64", "double", True),
])
def test_cpp_cavity_matches_palabos(gpu, tmp_path, name, precision, wrapped):
    """The C++ cavity program (dynamics-based and wrapped Zou-He factories) gives
    compiled Palabos's populations bit for bit."""
    case, _, data = load(name, GOLDEN_DIR)
    out = tmp_path / "f.bin"
    args = ["cavity3d", case.nx, 50, out, precision] + (["wrapped"] if wrapped else [])
    r = _run(args)
    dt = np.float64 if precision == "double" else np.float32
    got = np.fromfile(out, dtype=dt).reshape(case.shape + (19,))
    assert np.array_equal(got, data["s50"])
    fields_path = os.path.join(GOLDEN_DIR, f"fields_{name}.npz")
    if os.path.exists(fields_path):
        # computeKineticEnergy / computeVelocityNorm / computeVelocity of the facade
        # against compiled Palabos's fields (tests/golden/make_fields_golden.py)
        fl = np.load(fields_path, allow_pickle=False)["fields"]
        c, top = case.nx // 2, case.nz - 2
        assert r["energy_below_lid"] == float(fl[c, c, top, 5])
        assert r["norm_below_lid"] == float(fl[c, c, top, 6])
        assert r["ux_below_lid"] == float(fl[c, c, top, 2])


def test_cpp_cavity64_1000_steps(gpu, tmp_path):
    """Config 1 through the C++ API: 64^3, 1000 steps, bit-identical to Palabos,
    and Palabos's computeAverageDensity/Energy and stored statistics."""
    _, _, data = load("cavity64_f64", GOLDEN_DIR)
    out = tmp_path / "f.bin"
    r = _run(["cavity3d", 64, 1000, out, "double"])
    got = np.fromfile(out, dtype=np.float64)
    assert hashlib.sha256(got.tobytes()).hexdigest() == str(data["sha256_s1000"])
    stored_rho, stored_energy, avg_rho, avg_energy = data["stats_s1000"][:4]
    assert r["avg_density"] == pytest.approx(avg_rho, rel=1e-13)
    assert r["avg_energy"] == pytest.approx(avg_energy, rel=1e-11)
    assert r["stored_density"] == pytest.approx(stored_rho, rel=1e-13)
    assert r["stored_energy"] == pytest.approx(stored_energy, rel=1e-11)


@pytest.mark.parametrize("precision", ["double", "float"])
def test_coprocessor_plugin_roundtrip(gpu, tmp_path, precision):
    """addDomain / send / collideAndStream / receive of the HIP co-processor equal
    the oracle on the same block (BGK, edge reflection at the block's outer layer)."""
    storage = "f64" if precision == "double" else "f32"
    case = Case(9, 10, 11, storage=storage, omega=1.45)
    f0 = oracle.to_storage(random_state(case, 41), storage)
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    f0.tofile(fin)
    _run(["coprocessor", 9, 10, 11, 1.45, 7, fin, fout, precision])
    got = np.fromfile(fout, dtype=f0.dtype).reshape(f0.shape)
    want, _ = oracle.run(case, f0, 7)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("size,precision", [("20x20x20", "double"), ("20x31x42", "double"), ("20x31x42", "float")])
def test_cpp_doctest_cavity_matches_palabos(gpu, tmp_path, size, precision):
    """The jlabos regression cavity written against the C++ API
    (setVelocityConditionOnBlockBoundaries of createLocalBoundaryCondition3D:
    faces, edges and corners; lattice.initialize() runs the processors once):
    populations bit-identical to compiled Palabos after 30 steps, and its
    computeAverageEnergy / computeAverageDensity."""
    name = f"doctest_cavity_{size}_{'f64' if precision == 'double' else 'f32'}"
    case, _, data = load(name, GOLDEN_DIR)
    out = tmp_path / "f.bin"
    r = _run(["cylinder3d", case.nx, case.ny, case.nz, 30, out, precision])
    got = np.fromfile(out, dtype=np.float64 if precision == "double" else np.float32)
    assert hashlib.sha256(got.tobytes()).hexdigest() == str(data["sha256_s30"])
    tol = 1e-12 if precision == "double" else 1e-6
    assert r["avg_energy"] == pytest.approx(data["stats_s30"][3], rel=tol)
    assert r["avg_density"] == pytest.approx(data["stats_s30"][2], rel=tol)


@pytest.mark.parametrize("precision", ["double", "float"])
def test_palabos_coprocessor_path_runs_on_the_engine(gpu, precision):
    """The reference's own co-processor path with the engine behind it
    (oracle/coprocessor_driver.cpp, compiled against the reference's sources): an
    adapter deriving from the real plb::CoProcessor3D<T>, reached through
    global::defaultCoProcessor3D<T>(), two uneven periodic blocks offloaded by
    initiateCoProcessors and stepped through transferToCoProcessors /
    MultiBlockLattice3D::collideAndStream / transferFromCoProcessors: every
    population equals Palabos's CPU run after 6 steps, bit for bit."""
    exe = os.path.join(ROOT, "oracle", "_ref", "coprocessor_driver")
    if not os.path.exists(exe):
        pytest.skip("oracle/_ref/coprocessor_driver is built by oracle/Makefile.ref where the reference sources are")
    res = subprocess.run([exe, precision, "14", "9", "10", "6", "6", "1.55"], capture_output=True, text=True,
                         timeout=300)
    assert res.returncode == 0, res.stderr
    r = json.loads(res.stdout.strip().splitlines()[-1])
    assert r["handles"] == 2
    assert r["differ"] == 0, r
