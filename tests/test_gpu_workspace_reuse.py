"""One workspace, many sorts of different sizes and plans: every region of the workspace must hold
one kind of data whatever the size of the sort at hand (tile records at fixed addresses, the
indirect sorts' arrays at the end), so that no look-back read can mistake tile counts, tables or
gathered keys left by a sort of another size for a published look-back word. Sorted inputs give
large per-tile digit counts; fresh workspaces give small look-back tags. Every result is checked
against numpy (keys only, ascending) or the oracle (indices)."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
r = pytest.importorskip("radixsortbidi8_amd")


def keys_of(nb, n, dist, rng):
    dt = {4: np.uint32, 8: np.uint64}[nb]
    a = rng.integers(0, np.iinfo(dt).max, n, dtype=dt, endpoint=True)
    if dist == "sorted":
        a.sort()
    elif dist == "reverse":
        a = np.sort(a)[::-1].copy()
    elif dist == "fewtop":
        a &= dt((1 << (8 * nb - 8)) - 1)
    return a


@pytest.mark.parametrize("nb", [4, 8])
def test_sizes_and_plans_on_one_workspace(nb):
    rng = np.random.default_rng(1234 + nb)
    cap = 1 << 22
    ws = r.Workspace(r.U32 if nb == 4 else r.U64, cap)
    tdt = torch.uint32 if nb == 4 else torch.uint64
    out = torch.empty(cap, dtype=tdt, device="cuda")
    buf = torch.empty(cap, dtype=tdt, device="cuda")
    sizes = [(1 << 12) + 3, 1 << 16, (1 << 18) + 77, 1 << 20, 1 << 22, 1 << 17, 5000, (1 << 21) + 1]
    for rnd in range(2):
        for n in sizes:
            for algo in (2, 1, 4):
                for dist in ("sorted", "uniform", "reverse", "fewtop"):
                    prev = r.set_algorithm(algo)
                    try:
                        a = keys_of(nb, n, dist, rng)
                        d = torch.from_numpy(a).cuda()
                        r.radixsortcopynoalloc(d, out[:n], buf[:n], ws)
                        got = out[:n].cpu().numpy()
                    finally:
                        r.set_algorithm(prev)
                    assert np.array_equal(got, np.sort(a)), (rnd, n, algo, dist)


def test_indirect_sizes_on_one_workspace():
    rng = np.random.default_rng(99)
    cap = 1 << 20
    ws = r.Workspace(r.U64, cap, r.LAYOUT_INDIRECT)
    stride = 24
    for n in (1 << 10, 1 << 20, 3000, (1 << 19) + 5, 1 << 14, 1 << 20):
        rec = rng.integers(0, 256, n * stride, dtype=np.uint8)
        idx = rng.permutation(n).astype(np.uint32)
        d_rec = torch.from_numpy(rec).cuda()
        d_in = torch.from_numpy(idx).cuda()
        d_out = torch.empty_like(d_in)
        r.sort_indices(d_rec, stride, 8, d_in, d_out, ws, np.uint64)
        exp = oracle.sort_indirect(rec, stride, 8, idx, np.uint64, r.decode(r.U64, 64, 0, True))
        assert np.array_equal(d_out.cpu().numpy(), exp), n
