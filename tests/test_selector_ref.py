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
    assert (5, -1, 0) in pl.PLANS and not pl.plan_valid((5, -1, 0), 1 << 20, 2)
    assert not pl.plan_valid((5, -1, 0), (1 << 16) - 1, 8) and pl.plan_valid((5, -1, 0), 1 << 16, 4)
    table = {"8:20": [2, 4, 0], "8:10": [1, 0, 0]}
    assert pl.fixed_plan(table, 8, 1 << 21) == (2, 4, 0)
    assert pl.fixed_plan(table, 8, 1 << 9) == (1, 0, 0)


def test_library_plan_order_matches():
    assert pl.library_plans() == pl.PLANS


def test_segmented_table_covers_every_key_kind():
    """The tuned segmented table names a valid (tile configuration, LDS tier) for each key width and
    transform kind, and the sweep it came from timed every pair."""
    import json
    import os
    from radixsortbidi8_amd import SEGMENTED_TILE_CONFIGS, SMALL_CAPS
    path = os.path.join(os.path.dirname(__file__), "..", "radixsortbidi8_amd", "tuning", "segmented_table.json")
    with open(path) as f:
        t = json.load(f)
    assert set(t["best"]) == {"1:1", "2:1", "4:0", "4:1", "8:0", "8:1"}
    for tc, sc in t["best"].values():
        assert 0 <= tc < len(SEGMENTED_TILE_CONFIGS) and 0 <= sc < len(SMALL_CAPS)
    for name, times in t["times_ms"].items():
        assert len(times) == len(SEGMENTED_TILE_CONFIGS) * len(SMALL_CAPS), name


def test_segmented_automatic_choice_is_the_tuned_winner():
    """The library's automatic segmented configuration equals the sweep's winner for every key
    width and transform kind (integers: injective; floats: not)."""
    import json
    import os
    import numpy as np
    import radixsortbidi8_amd as r
    path = os.path.join(os.path.dirname(__file__), "..", "radixsortbidi8_amd", "tuning", "segmented_table.json")
    with open(path) as f:
        best = json.load(f)["best"]
    for dt, key in ((np.uint8, "1:1"), (np.int16, "2:1"), (np.int32, "4:1"), (np.uint32, "4:1"), (np.float32, "4:0"),
                    (np.uint64, "8:1"), (np.int64, "8:1"), (np.float64, "8:0")):
        assert list(r.segmented_default_config(dt)) == best[key], dt
