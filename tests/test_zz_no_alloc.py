"""palx_spectrum_noalloc: the spectrum of palx_spectrum with no device
allocation (static per-GPU staging bins).  Checked through the C ABI against
palx_spectrum on two calls in a row, the second on a smaller grid (fewer
bins), and against the FP64 oracle."""
import numpy as np
import pytest

from oracle import dns_oracle as O

pytestmark = pytest.mark.gpu


def _handle(P, N, prec, seed):
    S = O.create_state(N, 1e4, 5.0, 0.5, seed, "KM3", dtype=np.float64)
    h = P.Handle(N, prec)
    h.load_initial(S.uc0, P.calcom_f32(S.uc0))
    return h, S


def test_noalloc_spectrum_equals_palx_spectrum_twice_second_smaller():
    from palinstrophy_amd import _palx as P
    for N, prec, seed in ((128, 8, 3), (32, 4, 7)):   # the second call has a quarter of the bins
        print(f"case: noalloc spectrum N={N} precision={prec}", flush=True)
        h, S = _handle(P, N, prec, seed)
        nbins = max(32, 2 * N)
        e_ref, c_ref = h.spectrum(nbins)
        e_na, c_na = h.spectrum_noalloc(nbins)
        assert np.array_equal(c_na, c_ref)
        assert np.array_equal(e_na, e_ref)          # same kernel, one fixed summation order
        e_o, c_o, *_ = O.spectrum_bins(S)
        assert np.array_equal(c_na, c_o)
        # FP32 handles hold the same complex64 omega the oracle promotes: FP64 sums either way
        assert np.abs(e_na - e_o).max() <= 1e-12 * e_o.max()
        h.close()


def test_noalloc_spectrum_rejects_more_bins_than_its_array():
    from palinstrophy_amd import _palx as P
    h = P.Handle(32, 8)
    with pytest.raises(ValueError, match="PALX_SPECTRUM_NOALLOC_MAX_BINS"):
        h.spectrum_noalloc(P.SPECTRUM_NOALLOC_MAX_BINS + 1)
    h.close()
