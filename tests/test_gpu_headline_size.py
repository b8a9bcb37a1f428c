"""Correctness at the benchmarked size: the 512^3 periodic Taylor-Green lattice
of bench.py (BGK, omega = 1.8, u0 = 0.04; 134,217,728 cells, more than 2^31
population offsets in FP64), FP64 and FP16 storage, 200 steps. Total mass is
conserved; the FP64 kinetic energy follows the early-time viscous decay
E0 exp(-6 nu k^2 t) of the 3D vortex; the FP16-storage velocity field stays
within stated bounds of the FP64 one.

Bounds, from the oracle at 64^3 and 128^3 with the same u0 and omega: the
energy ratio differs from exp(-6 nu k^2 t) by -0.13 % at 200 steps (the
acoustic transient of a uniform-density start plus the lattice's
discretisation error); FP16 storage gives a relative velocity L2 error of
about 2e-4 against FP64 from the first steps on (the rounding of each stored
population to 11 significant bits: rounding the exact initial state alone
costs 1.2e-4), an absolute RMS error of about 4e-6 lattice units, and an
amplitude error (the component along the FP64 field) below 2e-5 after 200
steps; mean-density drift 8e-8 after 50 steps at 128^3."""
import math

import numpy as np
import pytest

from palabos_amd import studies

pytestmark = pytest.mark.gpu

N, STEPS, OMEGA, U0 = 512, 200, 1.8, 0.04


def test_taylor_green_512_fp64_and_fp16(gpu):
    lat64 = studies._tg_lattice(N, "f64", OMEGA, U0, "3d")
    lat16 = studies._tg_lattice(N, "f16", OMEGA, U0, "3d")
    try:
        rho64, e64 = lat64.averageDensity(), lat64.averageEnergy()
        rho16 = lat16.averageDensity()
        print(f"[512^3] initial <rho> {rho64!r} (f16 {rho16!r}), energy {e64:.9e}", flush=True)
        lat64.collideAndStream(STEPS)
        lat16.collideAndStream(STEPS)
        # mass: FP64 to rounding; FP16 storage rounds every population each step
        # (round to nearest even, unbiased), so the mean density drifts by rounding only
        assert abs(lat64.averageDensity() - rho64) < 1e-13
        assert abs(lat16.averageDensity() - rho16) < 1e-6
        nu = (1.0 / OMEGA - 0.5) / 3.0
        k = 2.0 * math.pi / N
        decay = math.exp(-6.0 * nu * k * k * STEPS)
        ratio = lat64.averageEnergy() / e64
        print(f"[512^3] FP64 energy ratio {ratio:.9f}, exp(-6 nu k^2 t) {decay:.9f}", flush=True)
        assert abs(ratio / decay - 1.0) < 5e-3
        d = s = p = 0.0
        for x0, x1 in studies._slabs(N):
            box = (x0, x1 - 1, 0, N - 1, 0, N - 1)
            u64 = lat64.velocity(box).reshape(-1, 3)
            u16 = lat16.velocity(box).reshape(-1, 3)
            d += float(((u16 - u64) ** 2).sum())
            s += float((u64 ** 2).sum())
            p += float(((u16 - u64) * u64).sum())
        rel, rms, amp = math.sqrt(d / s), math.sqrt(d / N ** 3), p / s
        print(f"[512^3] FP16 vs FP64 velocity: relative L2 {rel:.3e}, RMS {rms:.3e}, amplitude {amp:+.3e}", flush=True)
        assert rel < 4e-4
        assert rms < 1e-5
        assert abs(amp) < 1e-4
    finally:
        lat64.close()
        lat16.close()
