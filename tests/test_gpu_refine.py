"""Learned-preconditioner refinement machinery, end to end at a small size."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_refinement_end_to_end_small():
    from shotsweep_amd import refine
    print('\n[case] refine: 96 pairs, 2 epochs, 6 held-out models, 3 iterations', flush=True)
    result, net = refine.run(pairs=96, epochs=2, n_eval=6, iters=3, seed=1, dagger_rounds=1, log=lambda m: None)
    for key in ('plain_mean_misfit_ratio', 'learned_mean_misfit_ratio'):
        r = np.asarray(result[key])
        assert r.shape == (4,) and r[0] == pytest.approx(1.0)
        assert np.all(np.diff(r) <= 1e-6), key          # backtracking and fallback never raise the misfit
    assert len(result['learned_fallbacks_per_iter']) == 3
    print('  plain', np.round(result['plain_mean_misfit_ratio'], 4),
          'learned', np.round(result['learned_mean_misfit_ratio'], 4))


def test_learned_preconditioner_gain():
    """What the trained network gains (the learned-preconditioner workload, reduced and seeded): 2,000
    training pairs plus two rounds of pairs from the network's own refinement
    trajectories, 32 held-out models, 10 iterations.  The learned step reaches the
    plain adjoint-gradient step's 10-iteration misfit in at most half the
    iterations, ends with a lower misfit, and lowers the model RMS error more than
    the plain step does."""
    from shotsweep_amd import refine
    print('\n[case] refine gain: 2000 pairs + 2 aggregation rounds, 32 held-out models, 10 iterations', flush=True)
    r, _ = refine.run(pairs=2000, epochs=8, n_eval=32, iters=10, seed=0, log=lambda m: None)
    plain, learned = np.asarray(r['plain_mean_misfit_ratio']), np.asarray(r['learned_mean_misfit_ratio'])
    print(f"  reach {r['learned_iters_to_reach_plain_final']} iterations; final misfit ratio plain {plain[-1]:.4f} "
          f"learned {learned[-1]:.4f}; model RMS start {r['model_rms_error_start']:.1f} plain "
          f"{r['model_rms_error_plain']:.1f} learned {r['model_rms_error_learned']:.1f} m/s")
    assert r['learned_iters_to_reach_plain_final'] is not None and r['learned_iters_to_reach_plain_final'] <= 5
    assert learned[-1] < plain[-1]
    assert r['model_rms_error_learned'] < r['model_rms_error_plain'] < r['model_rms_error_start']
