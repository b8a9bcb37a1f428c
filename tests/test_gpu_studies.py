"""The configuration studies at reduced size: they run end to end on the
engine and their physics checks hold (Taylor-Green decay against the exact
2D-mode solution, storage errors ordered by precision, drag positive and
steady, Darcy flow through a multi-block porous medium)."""
import numpy as np
import pytest

from palabos_amd import studies

pytestmark = pytest.mark.gpu


def test_taylor_green_precision_small(gpu):
    r = studies.taylor_green_precision(n=32, steps=200, storages=("f32", "f16", "bf16"), modes=("2d",))
    m = r["modes"]["2d"]
    print(m, flush=True)
    # FP64 follows the exact solution to the lattice's O(dx^2, Ma^2) accuracy
    assert m["f64"]["velocity_l2_vs_analytic"] < 2e-2
    assert abs(m["f64"]["energy_error_vs_analytic"]) < 4e-2
    # errors against FP64 ordered by precision (measured at this size: FP32
    # 1.7e-5, FP16 2.7e-4, BF16 3.4e-3 relative), all far below the
    # discretisation error against the exact solution
    assert m["f32"]["velocity_l2_vs_f64"] < 1e-4
    assert m["f32"]["velocity_l2_vs_f64"] < m["f16"]["velocity_l2_vs_f64"] < m["bf16"]["velocity_l2_vs_f64"]
    assert m["bf16"]["velocity_l2_vs_f64"] < 2e-2
    assert m["bf16"]["velocity_l2_vs_f64"] < m["f64"]["velocity_l2_vs_analytic"]


def test_sphere_drag_small(gpu):
    r = studies.sphere_drag(nx=64, ny=32, nz=32, radius=4.0, center=(16.0, 16.0, 16.0), u_max=0.05, re=20.0,
                            steps=1500, average_last=200, ramp=500)
    print(r, flush=True)
    assert r["sphere_cells"] > 0
    assert r["fx_mean"] > 0
    assert r["fx_std"] < 0.1 * r["fx_mean"]   # the ramped start keeps the pressure waves small
    assert 1.0 < r["cd_u_mean"] < 100.0
    assert r["cd_u_max_min"] <= r["cd_u_max"] <= r["cd_u_max_max"]
    assert r["cd_u_max_std"] > 0 and r["swing_pct"] > 0


def test_porous_permeability_small(gpu):
    shape = (64, 48, 48)
    blocks = studies.setups.uneven_blocks(shape, (2, 2, 2), 16, 48, 3)
    r = studies.porous_permeability(shape=shape, count=12, steps=1200, storage="f16", g=1e-5, blocks=blocks,
                                    check_every=300)
    print(r, flush=True)
    h = np.array(r["ux_history"])
    assert r["blocks"] == 8
    assert 0.5 < r["porosity"] < 1.0
    assert np.all(h > 0) and r["permeability"] > 0
    # approaching a steady Darcy flow: increments shrink
    assert abs(h[-1] - h[-2]) < abs(h[1] - h[0])


def test_porous_adjoint_small_fd(gpu):
    """Radius-scale gradient of the permeability of a smooth-sphere porous medium
    (regularized BGK, Guo force, 8 uneven blocks, FP64) by the adjoint against
    central finite differences in the scale."""
    shape = (48, 32, 32)
    blocks = studies.setups.uneven_blocks(shape, (2, 2, 2), 12, 36, 3)
    r = studies.porous_adjoint(shape=shape, count=6, steps=200, checkpoint_every=20, window=100, storage="f64",
                               g=1e-4, blocks=blocks, fd_eps=1e-3)
    print(r, flush=True)
    assert r["dK_dscale_adjoint"] < 0      # larger spheres, lower permeability
    assert r["relative_difference"] < 1e-4


def test_porous_walls_adjoint_small_fd(gpu):
    """The adjoint through link-wise walls on a small sphere medium with a smooth
    Brinkman overlay (FP64, 8 uneven blocks): dK/ds against central finite
    differences in the overlay's scale."""
    shape = (48, 32, 32)
    blocks = studies.setups.uneven_blocks(shape, (2, 2, 2), 12, 36, 3)
    r = studies.porous_walls_adjoint(shape=shape, count=6, steps=200, checkpoint_every=20, window=100,
                                     storage="f64", g=1e-4, blocks=blocks, fd_eps=1e-3)
    print(r, flush=True)
    assert r["porosity"] < 1.0
    assert r["dK_ds_adjoint"] < 0          # more resistance, lower permeability
    assert r["relative_difference"] < 1e-4
