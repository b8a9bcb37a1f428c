"""GPU parity: the HIP engine against the oracle on the same seeded inputs.

Bar: bit-exact populations for FP64 and FP32 (both follow Palabos's operation
order, both compiled without contraction) and for FP16/BF16 storage (same FP32
arithmetic, same round-to-nearest-even conversions). Statistics are FP64 sums
taken in a different order, so they agree to 1e-12 relative.
"""
import copy
import hashlib
import os

import numpy as np
import pytest

from cases import (Case, TAG_BB, TAG_WALL, cavity_case, channel_case, gpu_lattice, gpu_run, random_state,
                   sphere_tags)
from golden_cases import GOLDEN, all_faces as _all_faces, load, walls_case as _walls
from oracle import oracle

pytestmark = pytest.mark.gpu


def _periodic(storage, dynamics="bgk", omega=1.7, shape=(8, 9, 10)):
    return Case(*shape, storage=storage, dynamics=dynamics, omega=omega, periodic=(True, True, True))


CASES = {
    "periodic_bgk": lambda s: (_periodic(s), 23),
    "periodic_rbgk": lambda s: (_periodic(s, "rbgk", 1.3), 23),
    "periodic_odd_shape": lambda s: (_periodic(s, shape=(5, 3, 70)), 9),
    "walls_bb_nodyn": lambda s: (_walls(s), 15),
    "cavity_zouhe": lambda s: (cavity_case(12, 1 / 0.6, 0.05, storage=s), 40),
    "cavity_regularized": lambda s: (cavity_case(12, 1 / 0.6, 0.05, storage=s, family="regularized"), 40),
    "channel_zouhe_mebb": lambda s: (channel_case(16, 10, 11, 1.8, 0.05, storage=s, sphere=((6, 5, 5), 2.2)), 30),
    "channel_reg_bb_rbgk": lambda s: (channel_case(16, 10, 11, 1.8, 0.05, storage=s, family="regularized",
                                                   sphere=((6, 5, 5), 2.2), sphere_tag=TAG_BB, dynamics="rbgk"), 30),
    "channel_wall_link": lambda s: (channel_case(16, 10, 11, 1.8, 0.05, storage=s, sphere=((6, 5, 5), 2.2),
                                                 sphere_tag=TAG_WALL), 30),
    "faces_zouhe": lambda s: (_all_faces(s, "zouhe"), 12),
    "faces_zouhe_swapped": lambda s: (_all_faces(s, "zouhe", True), 12),
    "faces_regularized": lambda s: (_all_faces(s, "regularized"), 12),
    "faces_regularized_swapped": lambda s: (_all_faces(s, "regularized", True), 12),
    "guo_force": lambda s: (Case(8, 9, 10, storage=s, omega=1.7, periodic=(True, True, True),
                                 force=(1e-4, -5e-5, 2e-5)), 11),
    "brinkman_rbgk": lambda s: (Case(8, 9, 10, storage=s, dynamics="rbgk", omega=1.6, periodic=(True, False, True),
                                     force=(2e-5, 0, 0),
                                     alpha=0.5 + 0.5 * np.sin(np.arange(720).reshape(8, 9, 10) * 0.37)), 11),
    "channel_porous_walls_rbgk": lambda s: (_porous_channel(s), 14),
    "forced_porous_walls_rbgk": lambda s: (_forced_walls(s), 14),
}


def _porous_channel(storage):
    """Zou-He planes, a link-wise wall sphere, RBGK with porosity: plane cells and
    cells next to the wall take the damped collision too."""
    case = channel_case(16, 10, 11, 1.8, 0.05, storage=storage, sphere=((6, 5, 5), 2.2), sphere_tag=TAG_WALL,
                        dynamics="rbgk")
    case.alpha = 0.2 + 0.2 * np.cos(np.arange(16 * 10 * 11).reshape(16, 10, 11) * 0.23)
    return case


def _forced_walls(storage):
    """Body force, porosity and a link-wise wall sphere (cells next to the wall
    take the forced collision), periodic in x and y."""
    case = Case(12, 10, 11, storage=storage, dynamics="rbgk", omega=1.7, periodic=(True, True, False),
                force=(1e-5, 0.0, 2e-6))
    case.tags = sphere_tags(case.shape, (6, 5, 5), 2.6, TAG_WALL)
    case.alpha = 0.2 + 0.2 * np.cos(np.arange(12 * 10 * 11).reshape(12, 10, 11) * 0.23)
    return case


def _state(case, seed=7):
    return oracle.to_storage(random_state(case, seed), case.storage)


def _compare(case, got, want):
    if case.tags is not None:
        wall = case.tags == TAG_WALL
        got = got.copy()
        want = want.copy()
        got[wall] = 0
        want[wall] = 0
    if case.storage in ("f16", "bf16"):
        assert np.array_equal(got, want), f"{np.count_nonzero(got != want)} populations differ"
    else:
        diff = np.abs(got.astype(np.float64) - want.astype(np.float64))
        assert np.array_equal(got, want), f"max diff {diff.max():.3e}, {np.count_nonzero(diff)} differ"


@pytest.mark.parametrize("storage", ["f64", "f32", "f16", "bf16"])
@pytest.mark.parametrize("name", sorted(CASES))
def test_gpu_matches_oracle(gpu, name, storage):
    case, steps = CASES[name](storage)
    print(f"[case] {name} {storage} {case.shape} steps={steps}", flush=True)
    f0 = _state(case)
    want, wst = oracle.run(case, f0, steps)
    got, gst = gpu_run(case, f0, steps)
    _compare(case, got, want)
    # statistics: same values, different FP64 summation order
    assert gst["count"] == wst["count"]
    np.testing.assert_allclose(gst["avg_rhobar"], wst["avg_rhobar"], rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(gst["avg_usqr"], wst["avg_usqr"], rtol=1e-12, atol=1e-18)
    assert gst["max_usqr"] == wst["max_usqr"]
    np.testing.assert_allclose(gst["force"], wst["force"], rtol=1e-11, atol=1e-14)
    np.testing.assert_allclose(gst["porous"], wst["porous"], rtol=1e-11, atol=1e-14)


LAUNCH_VARIANTS = {
    "unstaged": {"PLBAMD_STAGED": "0"},
    "staged": {"PLBAMD_STAGED": "1"},
    "segments64": {"PLBAMD_BULK_TX": "64"},
    "no_xcd_remap": {"PLBAMD_XCD_REMAP": "0"},
    "no_pairs": {"PLBAMD_PAIRS": "0"},
    "row_segments": {"PLBAMD_PACKED": "0"},
    "row_even_steps": {"PLBAMD_LINEAR_EVEN": "0"},
}
# (PLBAMD_LINEAR_EVEN=0 only switches the linear even step off; the multi-block
# tests run it on every storage, the 16-bit single-block cases here by default)


@pytest.mark.parametrize("storage", ["f64", "f16"])
@pytest.mark.parametrize("name", ["periodic_rbgk", "periodic_odd_shape", "walls_bb_nodyn", "channel_zouhe_mebb",
                                  "channel_wall_link", "faces_regularized", "forced_porous_walls_rbgk"])
@pytest.mark.parametrize("variant", sorted(LAUNCH_VARIANTS))
def test_gpu_launch_variants(gpu, monkeypatch, variant, name, storage):
    """The same results with the odd step unstaged (one kernel for both
    parities) or staged through LDS whatever the storage, with 64-cell row
    segments (several segments per row: halo elements between workgroups),
    one row segment per workgroup instead of several whole rows (16-bit pair
    kernel), and without the XCD-aware workgroup order."""
    for k, v in LAUNCH_VARIANTS[variant].items():
        monkeypatch.setenv(k, v)
    case, steps = CASES[name](storage)
    print(f"[case] {variant} {name} {storage} {case.shape} steps={steps}", flush=True)
    f0 = _state(case)
    want, wst = oracle.run(case, f0, steps)
    got, gst = gpu_run(case, f0, steps)
    _compare(case, got, want)
    np.testing.assert_allclose(gst["avg_rhobar"], wst["avg_rhobar"], rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(gst["force"], wst["force"], rtol=1e-11, atol=1e-14)


@pytest.mark.parametrize("storage", ["f64", "f16"])
def test_plane_values_changed_between_runs(gpu, storage):
    """New inlet values set after the lattice has run (refreshed on the device
    without a rebuild) act from the next step on, as in the oracle."""
    case = channel_case(16, 10, 11, 1.8, 0.05, storage=storage, sphere=((6, 5, 5), 2.2))
    f0 = _state(case)
    lat = gpu_lattice(case)
    try:
        lat.setPopulations(f0)
        lat.collideAndStream(5)
        slower = case.planes[0].values * 0.5
        lat.setBoundaryVelocity(0, slower)
        lat.collideAndStream(6)
        got = lat.getPopulations(dtype=np.uint16 if storage == "f16" else np.float64)
    finally:
        lat.close()
    want, _ = oracle.run(case, f0, 5)
    case2 = copy.deepcopy(case)
    case2.planes[0].values = slower
    want, _ = oracle.run(case2, want, 6)
    _compare(case, got, want)


@pytest.mark.parametrize("steps", [0, 1, 2, 3])
def test_transfer_roundtrip_every_parity(gpu, steps):
    """Upload/download work in both AA layouts: download after an even or odd step
    count equals the oracle, and re-uploading the download changes nothing."""
    case = _walls("f64")
    f0 = _state(case)
    lat = gpu_lattice(case)
    try:
        lat.setPopulations(f0)
        lat.collideAndStream(steps)
        got = lat.getPopulations()
        want, _ = oracle.run(case, f0, steps)
        assert np.array_equal(got, want)
        lat.setPopulations(got)
        assert np.array_equal(lat.getPopulations(), got)
        sub = (2, 5, 1, 7, 3, 8)
        part = lat.getPopulations(sub)
        assert np.array_equal(part, got[2:6, 1:8, 3:9])
        lat.collideAndStream(1)
        want2, _ = oracle.run(case, want, 1)
        assert np.array_equal(lat.getPopulations(), want2)
    finally:
        lat.close()


def test_equilibrium_and_moments(gpu):
    case = _periodic("f64")
    rng = np.random.RandomState(3)
    rho = 1 + 0.01 * rng.rand(*case.shape)
    u = 0.04 * (rng.rand(*case.shape, 3) - 0.5)
    lat = gpu_lattice(case)
    try:
        lat.initializeAtEquilibrium(rho, u)
        f = lat.getPopulations()
        np.testing.assert_array_equal(f, oracle.equilibrium(rho, u))
        r, v = lat.macroscopic()
        np.testing.assert_allclose(r, rho, rtol=1e-14)
        np.testing.assert_allclose(v, u, rtol=1e-12, atol=1e-16)
        np.testing.assert_allclose(lat.averageDensity(), rho.mean(), rtol=1e-13)
    finally:
        lat.close()


def test_statistics_off_and_time(gpu):
    case = _periodic("f32")
    lat = gpu_lattice(case)
    try:
        lat.toggleInternalStatistics(False)
        lat.collideAndStream(3)
        assert lat.time == 3
        import palabos_amd
        with pytest.raises(palabos_amd.PlbError):
            lat.getStatistics()
    finally:
        lat.close()


GOLDEN_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("name", sorted(n for n, s in GOLDEN.items()
                                        if not s.get("summary") and "wrapped" not in n and "external" not in s and
                                        not n.startswith(("blockbc_", "pressure_neumann"))))
def test_gpu_matches_palabos_golden(gpu, name):
    """The HIP engine reproduces compiled Palabos's populations bit for bit."""
    case, init, data = load(name, GOLDEN_DIR)
    print(f"[golden] {name}", flush=True)
    lat = gpu_lattice(case)
    try:
        lat.setPopulations(oracle.to_storage(init, case.storage))
        done = 0
        for s in GOLDEN[name]["dumps"]:
            lat.collideAndStream(s - done)
            done = s
            got = lat.getPopulations(dtype=np.float64 if case.storage == "f64" else np.float32)
            want = data[f"s{s}"]
            assert np.array_equal(got, want), f"step {s}: {np.count_nonzero(got != want)} differ"
            st = lat.getStatistics()
            np.testing.assert_allclose(st["force"], data[f"stats_s{s}"][4:7], rtol=1e-10, atol=1e-14)
    finally:
        lat.close()


def test_gpu_cavity64_1000_steps_matches_palabos(gpu):
    """Lid-driven cavity 64^3, Zou-He lid u = 0.05, omega = 1/0.6, 1000 steps, FP64:
    the full population field equals compiled Palabos's (SHA-256 of the array)."""
    case, init, data = load("cavity64_f64", GOLDEN_DIR)
    lat = gpu_lattice(case)
    try:
        lat.setPopulations(init)
        lat.collideAndStream(1000)
        got = lat.getPopulations()
        assert hashlib.sha256(got.tobytes()).hexdigest() == str(data["sha256_s1000"])
        st = lat.getStatistics()
        np.testing.assert_allclose(1 + st["avg_rhobar"], data["stats_s1000"][0], rtol=1e-12)
        np.testing.assert_allclose(0.5 * st["avg_usqr"], data["stats_s1000"][1], rtol=1e-11)
        np.testing.assert_allclose(lat.averageDensity(), data["stats_s1000"][2], rtol=1e-12)
        np.testing.assert_allclose(lat.averageEnergy(), data["stats_s1000"][3], rtol=1e-11)
    finally:
        lat.close()
