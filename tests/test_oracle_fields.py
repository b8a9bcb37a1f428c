"""The oracle's analysis fields (oracle.fields) against the fields compiled
Palabos computed (tests/golden/fields_*.npz) on the golden cases without
planes: bulk BGK / RBGK (FP32), BounceBack and NoDynamics walls (FP64), Guo
forcing (FP64 and FP32). Bit for bit."""
import os

import numpy as np
import pytest

from golden_cases import build
from oracle import oracle

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("name", ["periodic_rbgk_f32", "walls_bb_nodyn_f64", "guo_f64", "guo_f32"])
def test_oracle_fields_match_palabos(name):
    z = np.load(os.path.join(GOLDEN, f"fields_{name}.npz"), allow_pickle=False)
    step, want = int(z["step"]), z["fields"].astype(np.float64)
    case, init, _ = build(name)
    f, _ = oracle.run(case, oracle.to_storage(init, case.storage), step)
    got = oracle.fields(case, f)   # the reference driver builds BounceBack(1), NoDynamics(1)
    for k in range(7):
        assert np.array_equal(got[..., k], want[..., k]), f"{name}: field {k} differs in {np.count_nonzero(got[..., k] != want[..., k])} cells"


def test_oracle_fields_refuse_planes():
    case, init, _ = build("cavity_zouhe_f64")
    with pytest.raises(ValueError):
        oracle.fields(case, init)


def test_rhobar_is_the_sequential_sum():
    """computeRhoBar of a bulk cell is the FP32 sum f_0 + f_1 + ... + f_18 in
    index order (a restatement independent of the oracle): Palabos's field and
    the oracle's both equal it bit for bit, which pins the summation order."""
    z = np.load(os.path.join(GOLDEN, "fields_periodic_rbgk_f32.npz"), allow_pickle=False)
    case, init, _ = build("periodic_rbgk_f32")
    f, _ = oracle.run(case, oracle.to_storage(init, case.storage), int(z["step"]))
    f32 = f.astype(np.float32)
    rb = np.zeros(case.shape, np.float32)
    for i in range(19):
        rb = rb + f32[..., i]
    assert np.array_equal(oracle.fields(case, f)[..., 1], z["fields"][..., 1].astype(np.float64))
    assert np.array_equal(rb.astype(np.float64), z["fields"][..., 1].astype(np.float64))
