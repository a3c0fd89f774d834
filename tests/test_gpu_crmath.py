"""The device's correctly rounded pow / x^1.5 / hypot / sin / cos (crmath.h) give the same bits as the host
compilation of the same source, which tests/test_crmath.py pins to the nearest double of the exact value: so
the device's fma, sqrt, division and rounding behave as the algorithm assumes."""
import numpy as np
import pytest

from phygas_amd import hip

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("fn", ["pow", "pow15", "hypot", "sin", "cos"])
def test_device_bits_equal_host_bits(fn):
    rng = np.random.default_rng(7)
    n = 1 << 20
    if fn == "pow":
        x = np.exp(rng.uniform(np.log(1e-13), np.log(1e3), n))
        y = rng.choice([2.02, 1.02, 1.5, 0.5, 3.0, -1.25], n)
    elif fn == "pow15":
        x, y = np.exp(rng.uniform(np.log(1e-6), np.log(1e6), n)), None
    elif fn == "hypot":
        x, y = rng.uniform(-40, 40, n), rng.uniform(-40, 40, n)
    else:
        x, y = rng.uniform(-7.0, 7.0, n), None
    dev = hip.crmath(fn, x, y, device=True)
    host = hip.crmath(fn, x, y, device=False)
    assert np.array_equal(dev, host), int(np.sum(dev != host))
