"""The oracle against golden vectors written by compiled Palabos (tests/golden/).

FP64 and FP32 populations must match bit for bit: the oracle follows the
reference's operation order and both are compiled without contraction.
Palabos's "stored" statistics are FP64 sums over cells taken in its own
traversal order, so they are compared to 1e-12 relative.
"""
import hashlib
import os

import numpy as np
import pytest

from golden_cases import GOLDEN, load
from oracle import oracle

GOLDEN_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SMALL = sorted(n for n, s in GOLDEN.items() if not s.get("summary") and "external" not in s)
EXTERNAL = sorted(n for n, s in GOLDEN.items() if "external" in s)


def external_cycle(case, f, a, lattice=None):
    """One external-moment cycle as the reference driver runs it: rhoBar, j of
    every cell (BoxRhoBarJfunctional3D), j_x += a (1 + rhoBar) in the arithmetic
    type, one ExternalRhoJcollideAndStream3D step. With `lattice` (a Block3D) the
    device does the processors; the oracle otherwise. Returns (state or None,
    the field used)."""
    T = np.float64 if case.storage == "f64" else np.float32
    if lattice is None:
        rb, j = oracle.rhobar_j(case, f)
    else:
        lattice.computeRhoBarJ()
        rb, j = lattice.getRhoBarJ()
    rb, j = rb.astype(T), j.astype(T)
    j[..., 0] = j[..., 0] + T(a) * (T(1) + rb)
    if lattice is None:
        return oracle.external_step(case, f, rb, j), (rb, j)
    lattice.setRhoBarJ(rb, j)
    lattice.externalCollideAndStream(1)
    return None, (rb, j)


def _stats_check(name, ost, ref_row, case):
    stored_rho, stored_energy, avg_rho, avg_energy, fx, fy, fz = ref_row
    tol = 1e-12 if case.storage == "f64" else 1e-6
    np.testing.assert_allclose(1.0 + ost["avg_rhobar"], stored_rho, rtol=tol)
    np.testing.assert_allclose(0.5 * ost["avg_usqr"], stored_energy, rtol=max(tol, 1e-11), atol=1e-18)
    np.testing.assert_allclose(ost["force"], [fx, fy, fz], rtol=1e-10 if case.storage == "f64" else 1e-5,
                               atol=1e-14)


@pytest.mark.parametrize("name", SMALL)
def test_oracle_matches_palabos(name):
    case, init, data = load(name, GOLDEN_DIR)
    spec = GOLDEN[name]
    start = oracle.to_storage(init, case.storage)
    wrapped = spec.get("bcstyle") == "wrapped"
    for s in spec["dumps"]:
        # from the initial state each time: the values Neumann-type boundaries copy
        # are part of the state; the reference driver runs lattice.initialize() (one
        # pass of the processors) before its first step
        got, ost = oracle.run(case, start, s, initialize=True, wrapped=wrapped, snapshot=wrapped)
        want = data[f"s{s}"]
        assert got.dtype == want.dtype
        assert np.array_equal(got, want), (
            f"{name} step {s}: {np.count_nonzero(got != want)} populations differ, "
            f"max {np.abs(got.astype(float) - want.astype(float)).max():.3e}")
        _stats_check(name, ost, data[f"stats_s{s}"], case)


@pytest.mark.parametrize("name", EXTERNAL)
def test_oracle_external_cycle_matches_palabos(name):
    """BoxRhoBarJfunctional3D, a coupling j_x += a (1 + rhoBar), and
    ExternalRhoJcollideAndStream3D (dynamicsProcessor3D.hh:41-209) with BGK or
    regularized BGK, BounceBack and NoDynamics cells on a periodic lattice: the
    populations and the field used match compiled Palabos bit for bit."""
    case, init, data = load(name, GOLDEN_DIR)
    a = GOLDEN[name]["external"]
    f = oracle.to_storage(init, case.storage)
    done = 0
    for s in GOLDEN[name]["dumps"]:
        while done < s:
            f, (rb, j) = external_cycle(case, f, a)
            done += 1
        assert np.array_equal(f, data[f"s{s}"]), f"step {s}: {np.count_nonzero(f != data[f's{s}'])} differ"
        field = data[f"rhobarj_s{s}"]
        assert np.array_equal(rb, field[..., 0]) and np.array_equal(j, field[..., 1:])


def test_external_cycle_without_coupling_is_the_plain_step():
    """With the field left as BoxRhoBarJfunctional3D computes it, one external
    step equals one collideAndStream (same moments, same collision): bit-exact,
    BGK and regularized BGK."""
    for dyn in ("bgk", "rbgk"):
        case, init, _ = load("external_bgk_f64", GOLDEN_DIR)
        case.dynamics = dyn
        f0 = oracle.to_storage(init, "f64")
        ext, _ = external_cycle(case, f0, 0.0)
        plain, _ = oracle.run(case, f0, 1)
        assert np.array_equal(ext, plain), dyn


def test_external_step_refuses_non_periodic_lattices():
    """At a non-periodic edge Palabos's processor streams through the block
    envelope (populations enter from outside the domain): not restated, refused."""
    case, init, _ = load("external_bgk_f64", GOLDEN_DIR)
    case.periodic = (True, False, True)
    rb, j = oracle.rhobar_j(case, init)
    with pytest.raises(ValueError):
        oracle.external_step(case, init, rb, j)


def test_multiblock_equals_single_block():
    """Palabos's 6-block uneven decomposition is bit-identical to one block (the
    oracle), the decomposition invariance that envelopes guarantee."""
    case, init, data = load("multiblock_mebb_f64", GOLDEN_DIR)
    state, ost = oracle.run(case, init, 10)
    assert np.array_equal(state, data["s10"])


def test_cavity64_1000_steps():
    """Lid-driven cavity, 64^3, Zou-He lid u = 0.05, omega = 1/0.6, 1000 steps, FP64:
    the whole population field matches compiled Palabos bit for bit."""
    case, init, data = load("cavity64_f64", GOLDEN_DIR)
    state, ost = oracle.run(case, init, 1000)
    assert hashlib.sha256(state.tobytes()).hexdigest() == str(data["sha256_s1000"])
    np.testing.assert_array_equal(state.reshape(-1, 19).sum(axis=0), data["popsum_s1000"])
    _stats_check("cavity64", ost, data["stats_s1000"], case)


def test_wrapped_and_dynamics_based_agree_after_completion():
    """The wrapped Zou-He variant stores completed boundary populations; completing
    the dynamics-based snapshot gives the same field."""
    _, _, wrapped = load("cavity_zouhe_wrapped_f64", GOLDEN_DIR)
    case, _, dyn = load("cavity_zouhe_f64", GOLDEN_DIR)
    assert np.array_equal(oracle.complete(case, dyn["s50"]), wrapped["s50"])
    assert not np.array_equal(dyn["s50"], wrapped["s50"])


DOCTEST = sorted(n for n in GOLDEN if n.startswith("doctest_cavity_"))


@pytest.mark.parametrize("name", DOCTEST)
def test_doctest_cavity_matches_palabos(name):
    """The jlabos regression cavity (regularized velocity conditions of the wrapped
    factory on every face, edge and corner; jlabos/src/testing/test_cylinder.py:78-130)
    after 30 steps: the whole population field equals compiled Palabos's (SHA-256)."""
    case, init, data = load(name, GOLDEN_DIR)
    state, ost = oracle.run(case, oracle.to_storage(init, case.storage), 30, initialize=True, wrapped=True,
                            snapshot=True)
    assert hashlib.sha256(state.tobytes()).hexdigest() == str(data["sha256_s30"])
    _stats_check(name, ost, data["stats_s30"], case)
