"""External moments on the GPU: BoxRhoBarJfunctional3D and
ExternalRhoJcollideAndStream3D (basicDynamics/dynamicsProcessor3D.hh:41-209,
dataProcessors/dataAnalysisFunctional3D.hh:528-561) against compiled Palabos's
golden vectors and the oracle, bit for bit; the step with an unmodified field
equals the plain step; the refusals."""
import os

import numpy as np
import pytest

import palabos_amd as plb
from cases import Case, TAG_BB, TAG_NODYN, channel_case, gpu_lattice, random_state
from golden_cases import GOLDEN, load
from oracle import oracle
from test_oracle_golden import external_cycle

pytestmark = pytest.mark.gpu

GOLDEN_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
EXTERNAL = sorted(n for n, s in GOLDEN.items() if "external" in s)


def _dtype(storage):
    return {"f64": np.float64, "f32": np.float32}.get(storage, np.uint16)


def _tagged_case(storage, dynamics, shape=(10, 9, 12)):
    case = Case(*shape, storage=storage, dynamics=dynamics, omega=1.7, periodic=(True, True, True))
    tags = np.zeros(shape, np.uint8)
    tags[3:6, 2:5, 4:7] = TAG_BB
    tags[7, 6:8, 1:3] = TAG_NODYN
    case.tags = tags
    return case


@pytest.mark.parametrize("name", EXTERNAL)
def test_gpu_external_cycle_matches_palabos(gpu, name):
    """The coupling cycle (field computed, j_x += a (1 + rhoBar), external step)
    reproduces compiled Palabos's populations and field after every dump."""
    case, init, data = load(name, GOLDEN_DIR)
    a = GOLDEN[name]["external"]
    print(f"[external golden] {name}", flush=True)
    lat = gpu_lattice(case)
    try:
        lat.setPopulations(oracle.to_storage(init, case.storage))
        done = 0
        for s in GOLDEN[name]["dumps"]:
            while done < s:
                _, (rb, j) = external_cycle(case, None, a, lattice=lat)
                done += 1
            got = lat.getPopulations(dtype=_dtype(case.storage))
            assert np.array_equal(got, data[f"s{s}"]), f"step {s}: {np.count_nonzero(got != data[f's{s}'])} differ"
            field = data[f"rhobarj_s{s}"]
            assert np.array_equal(rb, field[..., 0]) and np.array_equal(j, field[..., 1:])
    finally:
        lat.close()


@pytest.mark.parametrize("storage", ["f64", "f32", "f16", "bf16"])
@pytest.mark.parametrize("dynamics", ["bgk", "rbgk"])
def test_gpu_external_cycle_matches_oracle(gpu, storage, dynamics):
    """Seven coupling cycles on a periodic lattice with BounceBack and NoDynamics
    cells, every storage (16-bit: the oracle's restatement of the rounding)."""
    case = _tagged_case(storage, dynamics)
    f0 = oracle.to_storage(random_state(case, 23), storage)
    print(f"[external oracle] {storage} {dynamics}", flush=True)
    want = f0
    lat = gpu_lattice(case)
    try:
        lat.setPopulations(f0)
        for _ in range(7):
            want, (rb_o, j_o) = external_cycle(case, want, -2e-4)
            _, (rb_g, j_g) = external_cycle(case, None, -2e-4, lattice=lat)
            assert np.array_equal(rb_g, rb_o) and np.array_equal(j_g, j_o)
        got = lat.getPopulations(dtype=_dtype(storage))
    finally:
        lat.close()
    assert np.array_equal(got, want), f"{np.count_nonzero(got != want)} populations differ"


@pytest.mark.parametrize("storage", ["f64", "f32", "f16", "bf16"])
@pytest.mark.parametrize("dynamics", ["bgk", "rbgk"])
def test_gpu_external_step_with_computed_field_is_the_plain_step(gpu, storage, dynamics):
    """An external step whose field is the populations' own moments equals a
    collideAndStream step, at an even and an odd time (both AA layouts)."""
    case = _tagged_case(storage, dynamics)
    f0 = oracle.to_storage(random_state(case, 29), storage)
    a, b = gpu_lattice(case), gpu_lattice(case)
    try:
        a.setPopulations(f0)
        b.setPopulations(f0)
        for _ in range(3):
            a.computeRhoBarJ()
            a.externalCollideAndStream(1)
            b.collideAndStream(1)
            assert np.array_equal(a.getPopulations(dtype=_dtype(storage)), b.getPopulations(dtype=_dtype(storage)))
    finally:
        a.close()
        b.close()


def test_gpu_rhobar_j_set_get_and_box(gpu):
    """Values set on a box come back rounded to the arithmetic type; the other
    cells keep the computed field; a box computation touches only its box."""
    case = _tagged_case("f32", "bgk")
    lat = gpu_lattice(case)
    try:
        lat.setPopulations(oracle.to_storage(random_state(case, 31), "f32"))
        lat.computeRhoBarJ()
        full_rb, full_j = lat.getRhoBarJ()
        want_rb, want_j = oracle.rhobar_j(case, oracle.to_storage(random_state(case, 31), "f32"))
        assert np.array_equal(full_rb, want_rb) and np.array_equal(full_j, want_j)
        box = (2, 5, 1, 4, 3, 8)
        rng = np.random.RandomState(2)
        rb = rng.standard_normal((4, 4, 6)) * 1e-3
        j = rng.standard_normal((4, 4, 6, 3)) * 1e-3
        lat.setRhoBarJ(rb, j, box=box)
        got_rb, got_j = lat.getRhoBarJ(box=box)
        assert np.array_equal(got_rb, rb.astype(np.float32).astype(np.float64))
        assert np.array_equal(got_j, j.astype(np.float32).astype(np.float64))
        all_rb, _ = lat.getRhoBarJ()
        mask = np.ones(case.shape, bool)
        mask[2:6, 1:5, 3:9] = False
        assert np.array_equal(all_rb[mask], full_rb[mask])
        lat.computeRhoBarJ(box=(0, 0, 0, 0, 0, 0))
        again_rb, _ = lat.getRhoBarJ(box=box)
        assert np.array_equal(again_rb, got_rb)
    finally:
        lat.close()


def test_gpu_external_refusals(gpu):
    """No field yet; a non-periodic lattice; boundary planes; a body force; a
    multi-block lattice that keeps its blocks."""
    case = _tagged_case("f64", "bgk")
    lat = gpu_lattice(case)
    try:
        with pytest.raises(plb.PlbError, match="no rhoBar"):
            lat.externalCollideAndStream(1)
    finally:
        lat.close()
    lat = plb.Block3D(8, 8, 8, plb.D3Q19, plb.BGK(1.6), periodic=(True, False, True))
    try:
        lat.computeRhoBarJ()
        with pytest.raises(plb.PlbError, match="periodic lattices only"):
            lat.externalCollideAndStream(1)
    finally:
        lat.close()
    lat = gpu_lattice(channel_case(12, 8, 9, 1.6, 0.03))
    try:
        with pytest.raises(plb.PlbError, match="single-block lattice of bulk"):
            lat.computeRhoBarJ()
    finally:
        lat.close()
    lat = plb.Block3D(8, 8, 8, plb.D3Q19, plb.BGK(1.6), periodic=True, force=(1e-5, 0, 0))
    try:
        with pytest.raises(plb.PlbError, match="body force"):
            lat.computeRhoBarJ()
    finally:
        lat.close()
    lat = plb.Block3D(8, 8, 8, plb.D3Q19, plb.BGK(1.6), periodic=True,
                      blocks=[(0, 3, 0, 7, 0, 7), (4, 7, 0, 7, 0, 7)], keep_blocks=True)
    try:
        with pytest.raises(plb.PlbError, match="single-block"):
            lat.computeRhoBarJ()
    finally:
        lat.close()
