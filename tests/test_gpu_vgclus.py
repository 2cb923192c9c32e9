"""The variogram clustering's kriging-weight heuristic
(vgclus/tvgs_clus.py:92-137) on the GPU class API against the oracle
(GPU)."""
import numpy as np
import pytest

from oracle import spk_oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize('vg,thr', [('1.0 Sph(40.0)', None),
                                    ('0.2 Nug(0.0) + 0.8 Exp(25.0)', 0.02),
                                    ('1.0 Gau(30.0)', 0.05)])
def test_sorted_weights_match_oracle(gpu, vg, thr):
    from spinterps_amd.vgclus import sorted_kriging_weights
    rng = np.random.default_rng(12)
    for trial in range(4):
        n = int(rng.integers(5, 40))
        pts = rng.uniform(0, 100, (n + 1, 2))
        got = sorted_kriging_weights(vg, pts, thr)
        exp = O.srtd_krg_wts(vg, pts, thr)
        assert got.shape == (n,)
        np.testing.assert_allclose(got, exp, rtol=1e-9, atol=1e-11)
        assert np.all(np.diff(got) >= 0) and abs(got.sum() - 1) < 1e-9
