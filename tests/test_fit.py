"""CMA-ES of cubnm_amd.fit on analytic functions (CPU): it must find the
minimum of a shifted, rotated ill-conditioned quadratic in 20 dimensions
inside the unit cube, and of a separable function with its minimum on a
face of the cube (clipping)."""
import numpy as np


def _run(f, dim, gens, seed=0):
    from cubnm_amd.fit import CMAES
    rng = np.random.default_rng(seed)
    X = rng.uniform(0, 1, (16, dim))
    y = np.array([f(x) for x in X])
    es = CMAES(X[np.argmin(y)], 0.3, 16, seed=seed)
    best = y.min()
    for _ in range(gens):
        X = es.ask()
        y = np.array([f(x) for x in X])
        es.tell(X, y)
        best = min(best, y.min())
    return best, es


def test_cmaes_rotated_quadratic():
    dim = 20
    rng = np.random.default_rng(3)
    Q, _ = np.linalg.qr(rng.standard_normal((dim, dim)))
    scales = np.logspace(0, 3, dim)
    x_star = rng.uniform(0.3, 0.7, dim)
    f = lambda x: float(np.sum(scales * (Q @ (x - x_star)) ** 2))
    best, es = _run(f, dim, 800)
    assert best < 1e-8
    assert np.abs(es.mean - x_star).max() < 1e-3


def test_cmaes_minimum_on_the_boundary():
    dim = 8
    f = lambda x: float(np.sum((x + 0.2) ** 2))   # minimum at x = -0.2, clipped to 0
    best, es = _run(f, dim, 200)
    assert abs(best - dim * 0.04) < 1e-6
