"""CPU checks of the selector network's double-precision reference (which in turn checks the GPU
kernels in test_gpu_selector.py): backpropagation against central finite differences, and the
plan table shape."""
import numpy as np

from radixsortbidi8_amd import planner as pl


def test_reference_gradient_matches_finite_differences():
    F, H, O = 7, 5, 4
    rng = np.random.default_rng(0)
    p = pl.init_params(F, H, O, 3).astype(np.float64)
    p += rng.normal(0, 0.1, p.size)
    x = rng.normal(0, 1, (6, F))
    t = rng.normal(5, 1, (6, O))
    m = (rng.random((6, O)) < 0.7).astype(np.float64)
    _, g = pl.ref_loss_grad(p, x, t, m, F, H, O)
    eps = 1e-6
    num = np.zeros_like(p)
    for i in range(p.size):
        d = np.zeros_like(p)
        d[i] = eps
        num[i] = (pl.ref_loss_grad(p + d, x, t, m, F, H, O)[0] - pl.ref_loss_grad(p - d, x, t, m, F, H, O)[0]) / (2 * eps)
    assert np.allclose(g, num, rtol=1e-5, atol=1e-7)


def test_plan_list():
    assert pl.NUM_PLANS == len(set(pl.PLANS))
    assert (3, -1, 0) in pl.PLANS
    assert all(pl.plan_valid(p, 1 << 20) == (p[0] != 3) for p in pl.PLANS)
    table = {"8:20": [2, 4, 0], "8:10": [1, 0, 0]}
    assert pl.fixed_plan(table, 8, 1 << 21) == (2, 4, 0)
    assert pl.fixed_plan(table, 8, 1 << 9) == (1, 0, 0)


def test_library_plan_order_matches():
    assert pl.library_plans() == pl.PLANS
