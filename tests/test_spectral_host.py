"""The library's host-side window weights and coherent gains
(spx_window_coefficients / spx_window_coherent_gain, the tables every
spectrum uses) against the oracle's restatement of window.rs:24-108.  Both
evaluate the reference's expressions in the same order with the C library's
cos, so they agree bit for bit.  No GPU needed."""
import numpy as np
import pytest

from oracle import spectral_oracle as S
from spicier_amd.spectral import WindowFunction as W


@pytest.mark.parametrize("w", list(W))
@pytest.mark.parametrize("n", [0, 1, 2, 3, 17, 1000, 4096])
def test_window_weights_match_oracle(w, n):
    got = w.coefficients(n)
    want = np.asarray(S.window_coefficients(int(w), n), dtype=np.float64)
    np.testing.assert_array_equal(got, want)
    g, o = w.coherent_gain(n), S.coherent_gain(int(w), n)
    assert (g == o) or (g != g and o != o)
    x = np.linspace(-1.0, 2.0, n)
    np.testing.assert_array_equal(w.apply(x), np.asarray(S.window_apply(int(w), list(x))))
