"""The library's restatement of GLU's MWC_4096 streams against the oracle's,
which tests/test_oracle_pins.py pins to the published algorithms GLU
combines (MT19937's init_genrand table, Marsaglia's complementary
multiply-with-carry).  Host only: the device copies of the streams start
from these states (tests/test_gpu_rng.py checks the device draws)."""
import numpy as np
import pytest

from oracle import oracle as O

import glu_amd as gl


@pytest.mark.parametrize("seed,nth", [(1, 1), (7, 4), (2024, 16), (4294967295, 3)])
def test_host_stream_matches_the_oracle_bit_for_bit(seed, nth):
    for t in sorted({0, nth // 2, nth - 1}):
        assert np.array_equal(gl.rng_draws(seed, nth, t, 1000), O.rng_draws(seed, nth, t, 1000)), (seed, nth, t)


def test_stream_arguments_are_checked():
    with pytest.raises(gl.GfixError):
        gl.rng_draws(0, 4, 0, 10)  # SEED 0 means /dev/urandom in GLU: no reproducible stream
    with pytest.raises(gl.GfixError):
        gl.rng_draws(5, 4, 4, 10)
