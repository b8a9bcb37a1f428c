"""Pins of the oracle that do not need a Palabos run: storage roundings against
numpy / exhaustive bit patterns, conservation laws, fixed points, streaming as
a permutation, link-wise walls against the Palabos-pinned edge reflection,
Guo and Brinkman momentum balances, and the analytic Taylor-Green decay."""
import numpy as np
import pytest

from cases import TAG_NODYN, TAG_WALL, Case, random_state, uniform
from oracle import oracle

C = np.array([[0, 0, 0], [-1, 0, 0], [0, -1, 0], [0, 0, -1], [-1, -1, 0], [-1, 1, 0], [-1, 0, -1],
              [-1, 0, 1], [0, -1, -1], [0, -1, 1], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 0],
              [1, -1, 0], [1, 0, 1], [1, 0, -1], [0, 1, 1], [0, 1, -1]])
OPP = [0] + list(range(10, 19)) + list(range(1, 10))


# -- storage roundings ---------------------------------------------------------

def test_f16_decode_exhaustive():
    bits = np.arange(65536, dtype=np.uint16)
    got = oracle.f16_to_f32(bits)
    want = bits.view(np.float16).astype(np.float32)
    nan = np.isnan(want)
    assert np.array_equal(np.isnan(got), nan)
    assert np.array_equal(got[~nan].view(np.uint32), want[~nan].view(np.uint32))


def test_f16_encode_matches_numpy_rne():
    # every half value, the midpoints between neighbours (ties), and 1 ulp either side
    h = np.arange(0, 0x7c00, dtype=np.uint16).view(np.float16).astype(np.float64)
    mids = (h[:-1] + h[1:]) / 2
    x = np.concatenate([h, mids, np.nextafter(mids, 0), np.nextafter(mids, 1), [70000.0, 65520.0, 65519.0]])
    x = x.astype(np.float32)
    x = np.concatenate([x, -x, uniform((100000,), 3).astype(np.float32) * 1e-3])
    got = oracle.f32_to_f16(x)
    with np.errstate(over="ignore"):
        want = x.astype(np.float16).view(np.uint16)
    assert np.array_equal(got, want)


def test_bf16_roundings():
    bits = np.arange(65536, dtype=np.uint16)
    dec = oracle.bf16_to_f32(bits)
    assert np.array_equal(dec.view(np.uint32), bits.astype(np.uint32) << 16)
    x = np.concatenate([uniform((200000,), 9) * 1e-2, uniform((1000,), 10) * 1e30]).astype(np.float32)
    enc = oracle.f32_to_bf16(x)
    val = oracle.bf16_to_f32(enc).astype(np.float64)
    # the two bf16 neighbours of x: truncation and truncation + 1 ulp (away from zero)
    lo_bits = (x.view(np.uint32) >> 16).astype(np.uint16)
    lo = oracle.bf16_to_f32(lo_bits).astype(np.float64)
    hi = oracle.bf16_to_f32(lo_bits + 1).astype(np.float64)
    xd = x.astype(np.float64)
    err = np.abs(val - xd)
    assert np.all(err <= np.minimum(np.abs(lo - xd), np.abs(hi - xd)))
    tie = np.abs(lo - xd) == np.abs(hi - xd)
    assert np.all(enc[tie] % 2 == 0)


# -- conservation and fixed points ---------------------------------------------

def _totals(f):
    f = f.astype(np.float64).reshape(-1, 19)
    return f.sum(), (f @ C).sum(axis=0)


@pytest.mark.parametrize("dynamics,omega", [("bgk", 1.7), ("rbgk", 1.3), ("bgk", 0.6)])
def test_collide_and_stream_conserve_mass_and_momentum(dynamics, omega):
    case = Case(8, 9, 10, dynamics=dynamics, omega=omega, periodic=(True, True, True))
    f0 = random_state(case, 21)
    m0, j0 = _totals(f0)
    f1, _ = oracle.run(case, f0, 3)
    m1, j1 = _totals(f1)
    assert abs(m1 - m0) < 1e-13 * case.nx * case.ny * case.nz
    np.testing.assert_allclose(j1, j0, atol=1e-13 * case.nx * case.ny * case.nz)


@pytest.mark.parametrize("dynamics", ["bgk", "rbgk"])
def test_uniform_equilibrium_is_a_fixed_point(dynamics):
    case = Case(5, 6, 7, dynamics=dynamics, omega=1.4, periodic=(True, True, True))
    rho = np.full(case.shape, 1.01)
    u = np.broadcast_to([0.03, -0.02, 0.01], case.shape + (3,))
    f0 = oracle.equilibrium(rho, u)
    f1, _ = oracle.run(case, f0, 2)
    np.testing.assert_allclose(f1, f0, rtol=0, atol=1e-16)


# -- streaming -----------------------------------------------------------------

def test_streaming_is_a_periodic_shift():
    """With NoDynamics everywhere a step is pure streaming: f_i(x) <- f_i(x - c_i)."""
    case = Case(6, 7, 8, periodic=(True, True, True))
    case.tags = np.full(case.shape, TAG_NODYN, np.uint8)
    f0 = random_state(case, 5)
    f1, _ = oracle.run(case, f0, 1)
    want = np.stack([np.roll(f0[..., i], shift=tuple(C[i]), axis=(0, 1, 2)) for i in range(19)], -1)
    assert np.array_equal(f1, want)


def test_edge_reflection_is_same_step_bounce_back():
    """At a non-periodic edge a population whose source lies outside is the
    cell's own opposite population (blockLattice3D.hh:403-427)."""
    case = Case(5, 6, 7, periodic=(False, True, False))
    case.tags = np.full(case.shape, TAG_NODYN, np.uint8)
    f0 = random_state(case, 6)
    f1, _ = oracle.run(case, f0, 1)
    nx, ny, nz = case.shape
    for i in range(1, 19):
        for x, y, z in [(0, 2, 0), (nx - 1, 0, nz - 1), (2, 3, 0), (0, 0, 3)]:
            sx, sy, sz = x - C[i][0], (y - C[i][1]) % ny, z - C[i][2]
            inside = 0 <= sx < nx and 0 <= sz < nz
            want = f0[sx, sy, sz, i] if inside else f0[x, y, z, OPP[i]]
            assert f1[x, y, z, i] == want


@pytest.mark.parametrize("storage", ["f64", "f16"])
def test_link_wall_equals_domain_edge(storage):
    """A periodic lattice whose one-cell outer shell is a link-wise wall evolves its
    interior exactly like the non-periodic lattice with Palabos's edge reflection."""
    inner = Case(6, 7, 8, storage=storage, omega=1.6)
    outer = Case(8, 9, 10, storage=storage, omega=1.6, periodic=(True, True, True))
    tags = np.full(outer.shape, TAG_WALL, np.uint8)
    tags[1:-1, 1:-1, 1:-1] = 0
    outer.tags = tags
    f_in = oracle.to_storage(random_state(inner, 8), storage)
    f_out = np.zeros(outer.shape + (19,), dtype=f_in.dtype)
    f_out[1:-1, 1:-1, 1:-1] = f_in
    a, sa = oracle.run(inner, f_in, 9)
    b, sb = oracle.run(outer, f_out, 9)
    assert np.array_equal(a, b[1:-1, 1:-1, 1:-1])
    assert sa["avg_rhobar"] == pytest.approx(sb["avg_rhobar"], rel=1e-13)


def test_link_wall_force_is_reflected_momentum():
    """Force on a link-wise wall = total momentum the fluid loses to it per step,
    measured as the fluid momentum change in a closed periodic box."""
    case = Case(10, 9, 8, omega=1.5, periodic=(True, True, True))
    tags = np.zeros(case.shape, np.uint8)
    tags[4:6, 3:6, 3:5] = TAG_WALL
    case.tags = tags
    f0 = random_state(case, 30)
    f0[tags == TAG_WALL] = 0
    j0 = _totals(f0)[1]
    f1, st = oracle.run(case, f0, 1)
    f1[tags == TAG_WALL] = 0
    j1 = _totals(f1)[1]
    np.testing.assert_allclose(j0 - j1, st["force"], atol=1e-13)


# -- body force and porosity ---------------------------------------------------

def test_guo_force_momentum_balance():
    """Guo forcing: the equilibrium is taken at u = j/rho + F/2 and the source adds
    (1 - omega/2) F, so a collision adds F (1 + omega (rho - 1)/2) to each cell."""
    F = np.array([2e-5, -1e-5, 5e-6])
    omega = 1.3
    case = Case(6, 7, 8, omega=omega, periodic=(True, True, True), force=tuple(F))
    f0 = random_state(case, 12)
    j0 = _totals(f0)[1]
    rho = f0.reshape(-1, 19).sum(1) + 1
    f1, _ = oracle.run(case, f0, 1)
    j1 = _totals(f1)[1]
    want = F * (1 + omega * (rho - 1) / 2).sum()
    np.testing.assert_allclose(j1 - j0, want, rtol=1e-9, atol=1e-15)


@pytest.mark.parametrize("dynamics", ["bgk", "rbgk"])
def test_brinkman_solid_relaxes_momentum_geometrically(dynamics):
    """alpha = 1 everywhere: the equilibrium is at rest, so BGK gives j' = (1 - omega) j
    per collision and the porous drag sums omega * j; regularized BGK sets the first
    moment to the equilibrium's, j' = 0, and the drag is j."""
    omega = 1.25
    case = Case(5, 6, 7, dynamics=dynamics, omega=omega, periodic=(True, True, True),
                alpha=np.ones((5, 6, 7)))
    f0 = random_state(case, 14)
    j0 = _totals(f0)[1]
    f1, st = oracle.run(case, f0, 1)
    j1 = _totals(f1)[1]
    keep = (1 - omega) if dynamics == "bgk" else 0.0
    removed = omega if dynamics == "bgk" else 1.0
    np.testing.assert_allclose(j1, keep * j0, rtol=1e-11, atol=1e-14)
    np.testing.assert_allclose(st["porous"], removed * j0, rtol=1e-11, atol=1e-15)


@pytest.mark.parametrize("dynamics", ["bgk", "rbgk"])
def test_porosity_zero_reduces_to_plain_dynamics(dynamics):
    base = Case(5, 6, 7, dynamics=dynamics, omega=1.5, periodic=(True, True, True))
    por = Case(5, 6, 7, dynamics=dynamics, omega=1.5, periodic=(True, True, True), alpha=np.zeros((5, 6, 7)))
    f0 = random_state(base, 15)
    a, _ = oracle.run(base, f0, 4)
    b, st = oracle.run(por, f0, 4)
    np.testing.assert_allclose(b, a, rtol=0, atol=2e-16)
    np.testing.assert_allclose(st["porous"], 0, atol=1e-15)


# -- analytic ------------------------------------------------------------------

def test_taylor_green_2d_mode_decays_at_the_lattice_viscosity():
    """u = u0 (sin kx cos ky, -cos kx sin ky, 0) is an exact Navier-Stokes solution
    with kinetic energy E0 exp(-4 nu k^2 t); nu = (1/omega - 1/2)/3."""
    n, steps, omega, u0 = 48, 300, 1.2, 0.01
    case = Case(n, n, 2, omega=omega, periodic=(True, True, True))
    k = 2 * np.pi / n
    x, y, _ = np.meshgrid(np.arange(n), np.arange(n), np.arange(2), indexing="ij")
    u = np.zeros((n, n, 2, 3))
    u[..., 0] = u0 * np.sin(k * x) * np.cos(k * y)
    u[..., 1] = -u0 * np.cos(k * x) * np.sin(k * y)
    f0 = oracle.equilibrium(np.ones((n, n, 2)), u)

    def energy(f):
        f = f.reshape(-1, 19)
        rho = f.sum(1) + 1
        v = (f @ C) / rho[:, None]
        return 0.5 * (v ** 2).sum()

    e0 = energy(f0)
    f1, _ = oracle.run(case, f0, steps)
    nu = (1 / omega - 0.5) / 3
    measured = -np.log(energy(f1) / e0) / (4 * k * k * steps)
    assert measured == pytest.approx(nu, rel=5e-3)


# -- forward-mode derivatives (the adjoint's CPU reference) --------------------

def test_dual_number_derivatives_match_finite_differences():
    """The oracle's dual-number derivatives of the adjoint objective equal central
    finite differences of plain FP64 oracle runs (porosity cells and inlet scale)."""
    import copy
    from cases import channel_case, rest_state
    from palabos_amd import setups
    case = channel_case(16, 8, 8, 1.6, 0.04)
    alpha = setups.brinkman_blobs(case.shape, n_blobs=4, sigma=(2, 3), seed=1)
    alpha[0] = 0
    alpha[-1] = 0
    case.alpha = alpha
    f0 = rest_state(case)
    seeds = np.random.RandomState(0).choice(np.flatnonzero(alpha.ravel() > 0.1), 16, replace=False)
    T, xin, xout = 12, 1, 14
    J, dJa, dJs = oracle.objective_dual(case, f0, T, xin, xout, 1.0, 1.0, seeds)

    def J_of(c):
        f, tot = f0.copy(), 0.0
        for _ in range(T):
            f, _ = oracle.run(c, f, 1)
            rho = f.sum(-1) + 1
            jx = f @ C[:, 0]
            plane = c.ny * c.nz
            tot += ((rho[xin].sum() - rho[xout].sum()) / (3 * plane) + c.omega * (c.alpha * jx).sum()) / T
        return tot

    assert J == pytest.approx(J_of(case), rel=1e-13)
    eps = 1e-6
    for k in range(3):
        cp, cm = copy.deepcopy(case), copy.deepcopy(case)
        cp.alpha.ravel()[seeds[k]] += eps
        cm.alpha.ravel()[seeds[k]] -= eps
        fd = (J_of(cp) - J_of(cm)) / (2 * eps)
        assert dJa[k] == pytest.approx(fd, rel=2e-3, abs=1e-11)
    cp, cm = copy.deepcopy(case), copy.deepcopy(case)
    cp.planes[0].values = case.planes[0].values * (1 + eps)
    cm.planes[0].values = case.planes[0].values * (1 - eps)
    assert dJs == pytest.approx((J_of(cp) - J_of(cm)) / (2 * eps), rel=1e-7)


# -- closed-form post-collision moments: RBGK + Guo, Brinkman at 0 < alpha < 1 ---
#
# Palabos defines BGK + Guo (externalForceDynamics.hh:215-232) but neither a
# regularized BGK with a Guo force nor a Brinkman term, so these collisions are
# pinned here by what the moment algebra of D3Q19 fixes per cell. With
# rho = 1 + rhoBar, j and Pi = sum_i c_i c_i f-bar_i taken before the collision,
#   u0 = j / rho + F / 2,  u = (1 - alpha) u0,  jeq = rho u,
#   Pi_eq = cs2 rhoBar I + jeq jeq^T / rho     (second moment of the shifted feq),
#   G1 = (1 - omega/2) F,  G2 = (1 - omega/2) (u F^T + F u^T)   (Guo source, moments 1, 2),
# the post-collision moments are
#   BGK:  rhoBar' = rhoBar,  j' = (1 - omega) j + omega jeq + G1,
#         Pi' = (1 - omega) Pi + omega Pi_eq + G2;
#   RBGK: rhoBar' = rhoBar,  j' = jeq + G1,
#         Pi' = Pi_eq + (1 - omega) (Pi - cs2 rhoBar I - j j^T / rho) + G2,
# (the regularized non-equilibrium part carries PiNeq of the pre-collision
# moments, dynamicsTemplates3D.h:1902-1973 / momentTemplates3D.h:517-539, and no
# first moment), and the porous statistic sums omega rho alpha u0 (BGK) or
# rho alpha u0 (RBGK) over the cells. Each identity uses only the isotropy of
# the D3Q19 weights up to fourth order, which the lattice has.


def _cell_moments(f):
    f = f.reshape(-1, 19).astype(np.float64)
    rho_bar = f.sum(1)
    j = f @ C
    pi = np.einsum("ka,kb,nk->nab", C, C, f)
    return rho_bar, j, pi


def _unstream(f1):
    """Post-collision populations of a fully periodic lattice after one step:
    streaming moved f'_i(x) to x + c_i."""
    return np.stack([np.roll(f1[..., i], shift=tuple(-C[i]), axis=(0, 1, 2)) for i in range(19)], -1)


def _closed_form(dynamics, omega, F, alpha, rho_bar, j, pi, *, jeq_from_j=False, guo_u="damped",
                 rlb_with_j=False, pineq_with_jeq=False):
    """Post-collision (j', Pi') and the per-cell porous drag of the collisions
    defined above; the keyword switches give plausible mistakes instead."""
    rho = 1.0 + rho_bar
    u0 = j / rho[:, None] + F / 2
    u = (1 - alpha)[:, None] * u0
    jeq = (1 - alpha)[:, None] * j + rho[:, None] * F / 2 if jeq_from_j else rho[:, None] * u
    ug = {"damped": u, "undamped": u0, "no_half_force": (1 - alpha)[:, None] * j / rho[:, None]}[guo_u]
    g1 = (1 - omega / 2) * F
    g2 = (1 - omega / 2) * (np.einsum("na,b->nab", ug, F) + np.einsum("a,nb->nab", F, ug))
    eye = np.eye(3)[None]
    cs2 = 1.0 / 3.0
    jj = np.einsum("na,nb->nab", jeq, jeq) / rho[:, None, None]
    pi_eq = cs2 * rho_bar[:, None, None] * eye + jj
    if dynamics == "bgk":
        j1 = (1 - omega) * j + omega * jeq + g1
        pi1 = (1 - omega) * pi + omega * pi_eq + g2
        porous = omega * rho[:, None] * alpha[:, None] * u0
    else:
        jr = j if rlb_with_j else jeq
        j1 = jr + g1
        pi_eq_r = cs2 * rho_bar[:, None, None] * eye + np.einsum("na,nb->nab", jr, jr) / rho[:, None, None]
        jn = jeq if pineq_with_jeq else j
        pineq = pi - cs2 * rho_bar[:, None, None] * eye - np.einsum("na,nb->nab", jn, jn) / rho[:, None, None]
        pi1 = pi_eq_r + (1 - omega) * pineq + g2
        porous = rho[:, None] * alpha[:, None] * u0
    return j1, pi1, porous.sum(0)


# Mistakes each closed-form test must tell apart from the definition: forming
# j_eq from damped j instead of damped u, regularizing around j instead of j_eq,
# the undamped velocity or u without F/2 in the Guo source, PiNeq from j_eq.
_MISTAKES = {
    "bgk": [dict(jeq_from_j=True), dict(guo_u="undamped"), dict(guo_u="no_half_force")],
    "rbgk": [dict(jeq_from_j=True), dict(rlb_with_j=True), dict(guo_u="undamped"),
             dict(guo_u="no_half_force"), dict(pineq_with_jeq=True)],
}


@pytest.mark.parametrize("dynamics,omega", [("bgk", 1.3), ("bgk", 1.85), ("rbgk", 1.7), ("rbgk", 0.9)])
def test_forced_porous_collision_moments_closed_form(dynamics, omega):
    """RBGK + Guo and BGK + Guo with a Brinkman porosity 0 < alpha < 1 and random
    rho: per cell, mass, first and second moments after the collision equal the
    closed forms above to rounding, the porous statistic equals its sum, and
    every plausible mistake in _MISTAKES misses by more than 1e3 times the
    tolerance (e.g. regularizing around j instead of j_eq moves j' by
    rho F / 2 - alpha rho u0, damping j instead of u moves j_eq by alpha rho F / 2)."""
    shape = (6, 7, 8)
    F = np.array([3e-4, -2e-4, 1.5e-4])
    alpha = 0.05 + 0.9 * (uniform(shape, 40) * 0.5 + 0.5)
    case = Case(*shape, dynamics=dynamics, omega=omega, periodic=(True, True, True),
                force=tuple(F), alpha=alpha)
    f0 = random_state(case, 41, rho_amp=0.2, u_amp=0.06, noise=2e-3)
    rho_bar, j, pi = _cell_moments(f0)
    f1, st = oracle.run(case, f0, 1)
    rb1, j1, pi1 = _cell_moments(_unstream(f1))
    a = alpha.reshape(-1)
    tol = 2e-15
    assert np.abs(rb1 - rho_bar).max() < tol
    want_j, want_pi, want_porous = _closed_form(dynamics, omega, F, a, rho_bar, j, pi)
    assert np.abs(j1 - want_j).max() < tol
    assert np.abs(pi1 - want_pi).max() < tol
    np.testing.assert_allclose(st["porous"], want_porous, rtol=1e-12, atol=1e-14)
    for wrong in _MISTAKES[dynamics]:
        wj, wpi, _ = _closed_form(dynamics, omega, F, a, rho_bar, j, pi, **wrong)
        miss = max(np.abs(j1 - wj).max(), np.abs(pi1 - wpi).max())
        assert miss > 1e3 * tol, f"closed form cannot tell {wrong} apart"
    # the porous statistic of RBGK carries no omega, that of BGK does
    other = _closed_form("rbgk" if dynamics == "bgk" else "bgk", omega, F, a, rho_bar, j, pi)[2]
    assert np.abs(other - st["porous"]).max() > 1e-6


@pytest.mark.parametrize("dynamics", ["bgk", "rbgk"])
def test_rbgk_guo_without_porosity_closed_form(dynamics):
    """alpha = 0 (no porosity field): j' = (1 - omega) j + omega rho u0 + (1 - omega/2) F for
    BGK (Palabos's GuoExternalForceBGKdynamics, pinned by golden vectors) and
    j' = j + rho F / 2 + (1 - omega/2) F for RBGK + Guo, with the second moments above."""
    shape = (5, 6, 7)
    F = np.array([-2.5e-4, 1e-4, 3e-4])
    omega = 1.55
    case = Case(*shape, dynamics=dynamics, omega=omega, periodic=(True, True, True), force=tuple(F))
    f0 = random_state(case, 43, rho_amp=0.2, u_amp=0.06, noise=2e-3)
    rho_bar, j, pi = _cell_moments(f0)
    f1, _ = oracle.run(case, f0, 1)
    rb1, j1, pi1 = _cell_moments(_unstream(f1))
    zero = np.zeros(rho_bar.shape)
    want_j, want_pi, _ = _closed_form(dynamics, omega, F, zero, rho_bar, j, pi)
    assert np.abs(rb1 - rho_bar).max() < 2e-15
    assert np.abs(j1 - want_j).max() < 2e-15
    assert np.abs(pi1 - want_pi).max() < 2e-15
    if dynamics == "rbgk":
        rho = 1 + rho_bar
        np.testing.assert_allclose(j1, j + rho[:, None] * F / 2 + (1 - omega / 2) * F, rtol=0, atol=2e-15)


@pytest.mark.parametrize("dynamics", ["bgk", "rbgk"])
def test_brinkman_intermediate_alpha_without_force(dynamics):
    """No body force, 0 < alpha < 1: j' = (1 - omega) j + omega (1 - alpha) j (BGK) or
    (1 - alpha) j (RBGK); damping u and damping j coincide here, so this pins the
    damping factor itself (e.g. alpha in place of 1 - alpha fails by omega (1 - 2 alpha) j)."""
    shape = (5, 6, 7)
    omega = 1.45
    alpha = 0.5 * (uniform(shape, 44) * 0.5 + 0.5) + 0.1
    case = Case(*shape, dynamics=dynamics, omega=omega, periodic=(True, True, True), alpha=alpha)
    f0 = random_state(case, 45, rho_amp=0.2, u_amp=0.06, noise=2e-3)
    rho_bar, j, pi = _cell_moments(f0)
    f1, st = oracle.run(case, f0, 1)
    _, j1, pi1 = _cell_moments(_unstream(f1))
    a = alpha.reshape(-1)[:, None]
    keep = (1 - omega) + omega * (1 - a) if dynamics == "bgk" else (1 - a)
    assert np.abs(j1 - keep * j).max() < 2e-15
    wrong = (1 - omega) + omega * a if dynamics == "bgk" else a
    assert np.abs(j1 - wrong * j).max() > 1e-4
    removed = (omega if dynamics == "bgk" else 1.0) * (a * j).sum(0)
    np.testing.assert_allclose(st["porous"], removed, rtol=1e-12, atol=1e-15)
    want_j, want_pi, _ = _closed_form(dynamics, omega, np.zeros(3), a[:, 0], rho_bar, j, pi)
    assert np.abs(pi1 - want_pi).max() < 2e-15


# -- the 16-bit modes' BGK rates -----------------------------------------------

def test_exact_rates_are_exact_and_nearest():
    """The 16-bit storage modes relax BGK with omega' and constants that carry no
    rounding: a_k = t_k omega' exactly, the mass and momentum factors
    om1 + a0 + 6 a1 + 12 a2 and om1 + 3 (2 a1 + 8 a2) are 1 exactly, and omega' is
    the multiple of 36 * 2^-25 nearest to omega. A floor instead of rint fails the
    nearest check; a finer grid (2^-26) makes 36 w inexact above omega = 1 and fails
    the factors; a swapped weight class fails a_k = t_k omega'."""
    rng = np.random.RandomState(3)
    omegas = np.concatenate([rng.uniform(1e-3, 2.0, 20000), [2.0, 1.9999999, 1.0, 0.5, 1.8, 1.3, 0.6, 1e-3]])
    grid = 36.0 * 2.0 ** -25
    for om in omegas.astype(np.float32):
        om1, a0, a1, a2, omp = (float(v) for v in oracle.exact_rates(float(om)))
        assert (a0 * 3, a1 * 18, a2 * 36) == (omp, omp, omp)
        assert om1 + a0 + 6 * a1 + 12 * a2 == 1.0
        assert om1 + 3 * (2 * a1 + 8 * a2) == 1.0
        assert om1 == 1.0 - omp
        q = omp / grid
        assert q == round(q)
        assert abs(omp - float(om)) <= grid / 2 * (1 + 1e-12)


def test_16bit_bgk_collision_closed_form():
    """A spatially uniform periodic lattice streams onto itself, so one step is one
    collision per cell: in FP16 storage it equals the FP64 BGK update
    f (1 - omega') + omega' f_eq(rhoBar, j) to one FP16 rounding of the FP32
    evaluation (at most one ulp of the stored value). Palabos's omega is 5e-7 away
    and inside that tolerance: the rates themselves are pinned above."""
    omega = 1.7
    case = Case(2, 2, 2, storage="f16", omega=omega, periodic=(True, True, True))
    rng = np.random.RandomState(5)
    for trial in range(20):
        rho = 1.0 + 0.02 * (rng.rand() - 0.5)
        u = 0.08 * (rng.rand(3) - 0.5)
        f = oracle.equilibrium(np.array([rho]), u[None, :])[0] + 2e-3 * (rng.rand(19) - 0.5)
        f16 = oracle.to_storage(np.broadcast_to(f, (2, 2, 2, 19)).copy(), "f16")
        fin = oracle.from_storage(f16, "f16")[0, 0, 0]
        out = oracle.from_storage(oracle.run(case, f16, 1)[0], "f16")
        assert np.array_equal(out, np.broadcast_to(out[0, 0, 0], out.shape))
        rb = fin.sum()
        j = C.T @ fin
        omp = float(oracle.exact_rates(omega)[4])
        want = fin * (1 - omp) + omp * oracle.equilibrium(np.array([1 + rb]), (j / (1 + rb))[None, :])[0]
        ulp = np.spacing(np.abs(want).astype(np.float16)).astype(np.float64)
        assert np.all(np.abs(out[0, 0, 0] - want) <= ulp), trial
