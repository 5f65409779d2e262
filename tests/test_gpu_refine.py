"""Learned-preconditioner refinement machinery, end to end at a small size."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_refinement_end_to_end_small():
    from shotsweep_amd import refine
    print('\n[case] refine: 96 pairs, 2 epochs, 6 held-out models, 3 iterations', flush=True)
    result, net = refine.run(pairs=96, epochs=2, n_eval=6, iters=3, seed=1, log=lambda m: None)
    for key in ('plain_mean_misfit_ratio', 'learned_mean_misfit_ratio'):
        r = np.asarray(result[key])
        assert r.shape == (4,) and r[0] == pytest.approx(1.0)
        assert np.all(np.diff(r) <= 1e-6), key          # backtracking and fallback never raise the misfit
    assert len(result['learned_fallbacks_per_iter']) == 3
    print('  plain', np.round(result['plain_mean_misfit_ratio'], 4),
          'learned', np.round(result['learned_mean_misfit_ratio'], 4))
