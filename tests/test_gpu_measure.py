"""The measurement utilities behind bench.py's VALU roofline: the issue rates
of single instruction classes and of mixes, on the device."""
import pytest

import cudakpz_amd as K

pytestmark = pytest.mark.gpu

# 256 CUs x 4 SIMDs x 16 lanes at 2.4 GHz, and twice that: no VALU rate can exceed it
NOMINAL = 256 * 4 * 16 * 2.4e9


def test_class_rates():
    slow, fast = K.valu_peak(reps=2)
    assert 0.5 * NOMINAL < slow <= 1.05 * NOMINAL
    assert slow < fast <= 2.05 * NOMINAL


def test_mix_rates_lie_between_the_class_rates_and_fall_with_the_slow_share():
    slow, fast = K.valu_peak(reps=2)
    rates = {n: K.valu_mix_rate(n, reps=2) for n in (0, 2, 3, 4, 5, 8)}
    assert rates[0] > rates[3] > rates[8]
    assert rates[2] >= 0.95 * rates[4]
    for n, r in rates.items():
        assert 0.5 * slow < r <= 1.1 * fast, (n, r)
    with pytest.raises(K.KpzError):
        K.valu_mix_rate(7)
