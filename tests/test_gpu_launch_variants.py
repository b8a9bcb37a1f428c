"""Launch variants of the odd steps that the engine picks by lattice size and
shape, forced on small lattices through its overrides (read when a lattice is
created): nontemporal or ordinary z-moving stores (PLBAMD_CACHED_Z; small
lattices take the ordinary ones by default) and rows split into several
segments (PLBAMD_BULK_TX=64, so that a periodic row no longer wraps through
LDS and the halo loads and stores run). Every variant must give the oracle's
populations bit for bit: periodic rows, bounce-back z walls (row-end
reflections), and link-wall spheres (tagged kernels), in FP64, FP32 and FP16
storage (the staged kernel and the pair kernel)."""
import contextlib
import os

import numpy as np
import pytest

from cases import TAG_WALL, Case, gpu_lattice, random_state, sphere_tags
from oracle import oracle

pytestmark = pytest.mark.gpu

VARIANTS = [
    {},
    {"PLBAMD_CACHED_Z": "0"},
    {"PLBAMD_CACHED_Z": "1"},
    {"PLBAMD_BULK_TX": "64"},
    {"PLBAMD_BULK_TX": "64", "PLBAMD_CACHED_Z": "0"},
]


@contextlib.contextmanager
def _env(settings):
    old = {k: os.environ.get(k) for k in settings}
    os.environ.update(settings)
    try:
        yield
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


def _dtype(storage):
    return {"f64": np.float64, "f32": np.float32}.get(storage, np.uint16)


@pytest.mark.parametrize("storage", ["f64", "f32", "f16"])
@pytest.mark.parametrize("geometry", ["periodic", "zwalls", "linkwalls"])
def test_odd_step_launch_variants_match_the_oracle(gpu, storage, geometry):
    shape = (8, 6, 256)
    periodic = (True, True, False) if geometry == "zwalls" else (True, True, True)
    case = Case(*shape, storage=storage, dynamics="bgk", omega=1.7, periodic=periodic)
    if geometry == "linkwalls":
        case.tags = sphere_tags(shape, (4.0, 3.0, 128.0), 2.5, TAG_WALL)
    f0 = oracle.to_storage(random_state(case, 41), storage)
    want, _ = oracle.run(case, f0, 6)
    solid = case.tags == TAG_WALL if case.tags is not None else None
    if solid is not None:
        want = want.copy()
        want[solid] = 0
    for v in VARIANTS:
        print(f"[launch variant] {storage} {geometry} {v}", flush=True)
        with _env(v):
            lat = gpu_lattice(case)
        try:
            lat.setPopulations(f0)
            lat.collideAndStream(6)
            got = lat.getPopulations(dtype=_dtype(storage))
        finally:
            lat.close()
        if solid is not None:
            got[solid] = 0
        assert np.array_equal(got, want), f"{v}: {np.count_nonzero(got != want)} populations differ"
