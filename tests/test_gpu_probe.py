"""The streaming-bandwidth probe (plbamd_bandwidth_probe) runs both patterns
and reports the bytes it moves (sizes here fit the on-die cache, so no bound
on the rate is asserted)."""
import pytest

from palabos_amd import native

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("pattern,nbytes,per_unit", [("copy", 1 << 28, 32), ("slots", 304 * 64 ** 3, 304),
                                                      ("slots4", 304 * 64 ** 3, 152), ("slots16", 304 * 64 ** 3, 608)])
def test_probe_moves_the_stated_bytes(gpu, pattern, nbytes, per_unit):
    ms, moved, gbs = native.bandwidth_probe(pattern, nbytes, repeats=2)
    assert moved == nbytes // per_unit * per_unit
    assert ms > 0
    assert gbs > 0
