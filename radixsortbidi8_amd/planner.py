"""Plan selection for the rsbd8 GPU sort: the autotuned fixed plan and the learned selector.

A *plan* is (algorithm, tile configuration, presort check):
  algorithm 1  onesweep: each scatter pass resolves its tile offsets by decoupled look-back
  algorithm 2  count + chained scan + scatter per pass
  algorithm 3  one block sorts everything in LDS (keys only, n <= 16384)
  algorithm 4  two MSD partitions by the top two digits, then one LDS sort per bucket (keys only,
               4/8-byte keys; narrower keys and skewed inputs end in algorithm 2)
  tile config  index into TILE_CONFIGS (threads x keys per thread of the scatter kernels)
  presort      first test whether the input is already in order (or strictly reversed)
Every plan produces the same, bit-identical output; they differ only in time.

The *fixed plan* is the autotuner's table: for every key width and size bucket, the plan that
was fastest on uniformly random keys.  The *learned selector* is a small ReLU network that reads
24 features of a sample of the keys (rsbd8_hip_features) and predicts log2 of the time of every
plan; it is trained on the GPU by the library's own forward/backward/Adam kernels on timings the
sorter generates, and picks the plan with the lowest predicted time.  The reference has no
counterpart: it picks its CPU threading plan from compile-time thresholds
(Radixsortbidi8.hpp:12948-13152) and count-based branches (H:56797-57052).
"""
from __future__ import annotations

import ctypes
import json
import math
import os
import time

import numpy as np

from . import (F32, F64, I8, I16, I32, I64, TILE_CONFIGS, U8, U16, U32, U64, Workspace, _check, _load, _ptr,
               _stream_ptr, _torch, decode, keytype_of)

HERE = os.path.dirname(os.path.abspath(__file__))
TUNING_DIR = os.path.join(HERE, "tuning")
NUM_FEATURES = 24
HIDDEN = 32
LDS_MAX = 16384

PLANS = [(a, c, chk) for a in (1, 2, 4) for c in range(len(TILE_CONFIGS)) for chk in (0, 1)] + [(5, -1, 0), (3, -1, 0)]
SAMPLE_MIN = 1 << 16  # algorithm 5 (the sample partition plan whatever the keys): 4- and 8-byte keys from 2^16
NUM_PLANS = len(PLANS)

DTYPES = ("u8", "u16", "u32", "u64", "f32", "f64")
DISTS = ("uniform", "sorted", "reverse", "unique16", "zipf", "low2bytes")


class _Plan(ctypes.Structure):
    _fields_ = [("algo", ctypes.c_int), ("tile_cfg", ctypes.c_int), ("presort_check", ctypes.c_int),
                ("reserved", ctypes.c_int)]


def _lib():
    L = _load()
    if not getattr(L, "_planner_ready", False):
        vp, sz, ci, cu, cf = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_uint, ctypes.c_float
        sig = {
            "rsbd8_hip_radixsortcopynoalloc_plan": ([ci, cu, sz, vp, sz, vp, vp, vp, ctypes.POINTER(_Plan), vp], ci),
            "rsbd8_hip_features": ([ci, cu, sz, vp, vp, vp], ci),
            "rsbd8_hip_mlp_param_count": ([ci, ci, ci], ci),
            "rsbd8_hip_mlp_forward": ([ci, ci, ci, vp, vp, ci, vp, vp, vp], ci),
            "rsbd8_hip_mlp_backward": ([ci, ci, ci, vp, vp, ci, vp, vp, vp, vp, cf, vp, vp, vp], ci),
            "rsbd8_hip_mlp_adam": ([ci, vp, vp, vp, vp, cf, cf, cf, cf, ci, cf, vp], ci),
            "rsbd8_hip_plan_count": ([], ci),
            "rsbd8_hip_plan_at": ([ci, ctypes.POINTER(_Plan)], ci),
            "rsbd8_hip_select_plan": ([ci, cu, sz, vp, vp, ci, ctypes.POINTER(ci), vp], ci),
            "rsbd8_hip_radixsortcopynoalloc_learned": ([ci, cu, sz, vp, sz, vp, vp, vp, vp, ci, sz, ctypes.POINTER(_Plan),
                                                         ctypes.POINTER(ci), vp], ci),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        L._planner_ready = True
    return L


def plan_valid(plan, n: int, nbytes: int = 0) -> bool:
    """Whether plan runs on n keys of nbytes bytes (0: any width), as the library's selection masks it."""
    if plan[0] == 5:
        return n >= SAMPLE_MIN and nbytes in (0, 4, 8)
    return plan[0] != 3 or n <= LDS_MAX


def sort_with_plan(keys, out, buf, ws: Workspace, plan, flags: int, stream=None):
    """rsbd8_hip_radixsortcopynoalloc_plan on torch tensors."""
    kt = keytype_of(keys.dtype)
    p = _Plan(int(plan[0]), int(plan[1]), int(plan[2]), 0)
    _check(_lib().rsbd8_hip_radixsortcopynoalloc_plan(kt, flags, keys.numel(), ws.ptr(), ws.nbytes, _ptr(keys),
                                                       _ptr(out), _ptr(buf), ctypes.byref(p), _stream_ptr(stream)))


def library_plans():
    """The plan list as the library indexes it (must equal PLANS)."""
    L = _lib()
    out = []
    for i in range(L.rsbd8_hip_plan_count()):
        p = _Plan()
        _check(L.rsbd8_hip_plan_at(i, ctypes.byref(p)))
        out.append((p.algo, p.tile_cfg, p.presort_check))
    return out


def features(keys, flags: int, out=None, stream=None):
    """24 float features of a key array (device tensor)."""
    torch = _torch()
    if out is None:
        out = torch.empty(NUM_FEATURES, dtype=torch.float32, device=keys.device)
    _check(_lib().rsbd8_hip_features(keytype_of(keys.dtype), flags, keys.numel(), _ptr(keys), _ptr(out),
                                     _stream_ptr(stream)))
    return out


# ---- workloads ------------------------------------------------------------------------------

def torch_dtype(name):
    torch = _torch()
    return {"u8": torch.uint8, "u16": torch.uint16, "u32": torch.uint32, "u64": torch.uint64,
            "f32": torch.float32, "f64": torch.float64}[name]


_ZIPF_CACHE = {}


def _zipf_table(device):
    torch = _torch()
    if device not in _ZIPF_CACHE:
        ranks = torch.arange(1, (1 << 20) + 1, dtype=torch.float64, device=device)
        w = ranks.pow(-1.1)
        _ZIPF_CACHE[device] = torch.cumsum(w, 0) / w.sum()
    return _ZIPF_CACHE[device]


def _random_bits(n, nbytes, gen, device):
    torch = _torch()
    words = (n * nbytes + 3) // 4
    w = torch.randint(-(1 << 31), (1 << 31) - 1, (words,), dtype=torch.int32, device=device, generator=gen)
    return w.view(torch.uint8)[: n * nbytes]


def _signed_view(t):
    """Same-size view that torch can index and flip on the GPU (its unsigned 16/32/64-bit types
    have no indexing kernels)."""
    torch = _torch()
    return t.view({1: torch.uint8, 2: torch.int16, 4: torch.int32, 8: torch.int64}[t.element_size()])


def make_keys(dtype_name: str, n: int, dist: str, gen, device="cuda"):
    """A seeded synthetic key array of the request mix (uniform, sorted, reverse sorted, 16 unique
    values, Zipf(1.1) over 2^20 ranks, low two bytes varying)."""
    torch = _torch()
    td = torch_dtype(dtype_name)
    nb = torch.tensor([], dtype=td).element_size()
    if dtype_name in ("f32", "f64"):
        base = torch.rand(n, dtype=torch.float64, device=device, generator=gen).to(td)
    else:
        base = _random_bits(n, nb, gen, device).view(td)
    if dist == "uniform":
        return base.contiguous()
    if dist in ("sorted", "reverse"):
        sv = _sortable(base)
        order = torch.argsort(sv, stable=True)
        s = _signed_view(base)[order]
        return (s.flip(0) if dist == "reverse" else s).contiguous().view(td)
    if dist == "unique16":
        pick = torch.randint(0, 16, (n,), device=device, generator=gen)
        return _signed_view(base)[:16][pick % max(1, min(16, n))].contiguous().view(td)
    if dist == "zipf":
        u = torch.rand(n, dtype=torch.float64, device=device, generator=gen)
        r = torch.searchsorted(_zipf_table(device), u).clamp_(max=(1 << 20) - 1)
        vals = _random_bits(1 << 20, nb, gen, device).view(td) if dtype_name not in ("f32", "f64") else \
            torch.rand(1 << 20, dtype=torch.float64, device=device, generator=gen).to(td)
        return _signed_view(vals)[r].contiguous().view(td)
    if dist == "low2bytes":
        if nb <= 2:
            return base.contiguous()
        b = base.view(torch.uint8).view(n, nb).clone()
        b[:, 2:] = 0x5A
        return b.view(-1).view(td).contiguous()
    raise ValueError(dist)


def _sortable(t):
    """A torch-sortable int64 view ordered like the native ascending rsbd8 order (floats by value)."""
    torch = _torch()
    if t.dtype in (torch.float32, torch.float64):
        return t.to(torch.float64)
    if t.dtype == torch.uint64:
        return t.view(torch.int64) ^ torch.tensor(-(1 << 63), dtype=torch.int64, device=t.device)
    if t.dtype == torch.uint32:
        return t.view(torch.int32).to(torch.int64) & 0xFFFFFFFF
    if t.dtype == torch.uint16:
        return t.view(torch.int16).to(torch.int64) & 0xFFFF
    return t.to(torch.int64)


def request_stream(count: int, seed: int, lo_log2: float, hi_log2: float):
    """The seeded request mix: dtype uniform over the six, size log-uniform, distribution uniform."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        dt = DTYPES[int(rng.integers(0, len(DTYPES)))]
        n = int(2.0 ** rng.uniform(lo_log2, hi_log2))
        dist = DISTS[int(rng.integers(0, len(DISTS)))]
        out.append((dt, max(2, n), dist, int(rng.integers(0, 2 ** 62))))
    return out


# ---- timing -----------------------------------------------------------------------------------

class Arena:
    """Reusable device arrays sized for the largest request of a stream."""

    def __init__(self, max_n: int, device="cuda"):
        torch = _torch()
        self.max_n = max_n
        self.out = torch.empty(max_n * 8, dtype=torch.uint8, device=device)
        self.buf = torch.empty(max_n * 8, dtype=torch.uint8, device=device)
        self.ws = Workspace(U8, max_n)
        self.feat = torch.empty(NUM_FEATURES, dtype=torch.float32, device=device)

    def views(self, keys):
        n, nb = keys.numel(), keys.element_size()
        return self.out[: n * nb].view(keys.dtype), self.buf[: n * nb].view(keys.dtype)


def time_call(fn, reps: int = 3):
    """Median completion time (ms) of fn from an idle device: the host time from the call to the
    end of its last kernel, the cost a caller that waits for the result pays (it includes the launch
    overhead that a caller pipelining many sorts would hide)."""
    torch = _torch()
    times = []
    torch.cuda.synchronize()
    for _ in range(reps):
        b = torch.cuda.Event()
        t0 = time.perf_counter()
        fn()
        b.record()
        b.synchronize()
        times.append((time.perf_counter() - t0) * 1e3)
    return float(np.median(times))


def time_plans(keys, flags: int, arena: Arena, reps: int = 3):
    """Median ms of every valid plan on this key array (NaN for invalid plans)."""
    out, buf = arena.views(keys)
    res = np.full(NUM_PLANS, np.nan)
    for i, p in enumerate(PLANS):
        if not plan_valid(p, keys.numel(), keys.element_size()):
            continue
        sort_with_plan(keys, out, buf, arena.ws, p, flags)   # warm the plan once
        res[i] = time_call(lambda: sort_with_plan(keys, out, buf, arena.ws, p, flags), reps)
    return res


# ---- the network (host orchestration of the HIP kernels) --------------------------------------

class Selector:
    """features (24) -> 32 -> 32 -> log2 time in microseconds of every plan."""

    def __init__(self, params: np.ndarray | None = None, seed: int = 1, device="cuda"):
        torch = _torch()
        self.F, self.H, self.O = NUM_FEATURES, HIDDEN, NUM_PLANS
        self.count = _lib().rsbd8_hip_mlp_param_count(self.F, self.H, self.O)
        if params is None:
            params = init_params(self.F, self.H, self.O, seed)
        assert params.size == self.count
        self.params = torch.from_numpy(params.astype(np.float32)).to(device)
        self.device = device
        self.min_count = 0  # below this many keys the fixed plan is used without a selection

    def forward(self, x, acts=None):
        torch = _torch()
        x = x.contiguous()
        B = x.shape[0]
        y = torch.empty(B, self.O, dtype=torch.float32, device=self.device)
        _check(_lib().rsbd8_hip_mlp_forward(self.F, self.H, self.O, _ptr(self.params), _ptr(x), B, _ptr(acts), _ptr(y),
                                            _stream_ptr()))
        return y

    def gradient(self, x, target, mask):
        """Gradient of the masked mean squared error on the GPU (returns (loss, grad))."""
        torch = _torch()
        B = x.shape[0]
        acts = torch.empty(B, 2 * self.H, dtype=torch.float32, device=self.device)
        y = self.forward(x, acts)
        per = torch.empty(B, self.count, dtype=torch.float32, device=self.device)
        grad = torch.empty(self.count, dtype=torch.float32, device=self.device)
        msum = float(mask.sum().item())
        _check(_lib().rsbd8_hip_mlp_backward(self.F, self.H, self.O, _ptr(self.params), _ptr(x.contiguous()), B,
                                             _ptr(acts), _ptr(y), _ptr(target.contiguous()), _ptr(mask.contiguous()),
                                             msum, _ptr(per), _ptr(grad), _stream_ptr()))
        loss = float((((y - target) ** 2) * mask).sum().item() / max(msum, 1.0))
        return loss, grad

    def train(self, x, target, mask, steps: int = 3000, lr: float = 3e-3, decay: float = 1e-4, log=None):
        torch = _torch()
        m = torch.zeros_like(self.params)
        v = torch.zeros_like(self.params)
        loss = float("nan")
        for step in range(1, steps + 1):
            loss, grad = self.gradient(x, target, mask)
            lr_t = lr * (0.1 ** (step / steps))
            _check(_lib().rsbd8_hip_mlp_adam(self.count, _ptr(self.params), _ptr(grad), _ptr(m), _ptr(v), lr_t, 0.9,
                                             0.999, 1e-8, step, decay, _stream_ptr()))
            if log and (step % 500 == 0 or step == 1):
                log(f"selector step {step} loss {loss:.4f}")
        return loss

    def select(self, keys, flags: int, stream=None) -> int:
        """rsbd8_hip_select_plan: the plan index the library's fused selection kernel picks."""
        idx = ctypes.c_int(-1)
        _check(_lib().rsbd8_hip_select_plan(keytype_of(keys.dtype), flags, keys.numel(), _ptr(keys), _ptr(self.params),
                                            self.H, ctypes.byref(idx), _stream_ptr(stream)))
        return idx.value

    def sort(self, keys, out, buf, ws: Workspace, flags: int, below=None, stream=None) -> int:
        """rsbd8_hip_radixsortcopynoalloc_learned: select on the device, then sort; returns the plan
        index, or -1 when the request is below the gate (min_count) and the plan below was used."""
        idx = ctypes.c_int(-1)
        bp = ctypes.byref(_Plan(int(below[0]), int(below[1]), int(below[2]), 0)) if below is not None else None
        _check(_lib().rsbd8_hip_radixsortcopynoalloc_learned(keytype_of(keys.dtype), flags, keys.numel(), ws.ptr(),
                                                             ws.nbytes, _ptr(keys), _ptr(out), _ptr(buf),
                                                             _ptr(self.params), self.H, int(self.min_count), bp,
                                                             ctypes.byref(idx), _stream_ptr(stream)))
        return idx.value

    def choose(self, feats, n: int, nbytes: int = 0):
        """Index of the plan with the lowest predicted time among those valid for n keys of nbytes."""
        y = self.forward(feats.view(1, -1))[0].cpu().numpy()
        for i, p in enumerate(PLANS):
            if not plan_valid(p, n, nbytes):
                y[i] = np.inf
        return int(np.argmin(y))

    def save(self, path):
        np.savez(path, params=self.params.cpu().numpy(), plans=np.array(PLANS), F=self.F, H=self.H, O=self.O,
                 min_count=self.min_count)

    @classmethod
    def load(cls, path, device="cuda"):
        d = np.load(path, allow_pickle=False)
        assert [tuple(p) for p in d["plans"].tolist()] == PLANS
        sel = cls(d["params"], device=device)
        sel.min_count = int(d["min_count"]) if "min_count" in d else 0
        return sel

    def calibrate_gate(self, requests, X, T, M, table: dict, overhead_ms: float):
        """Chooses min_count (a power of two, or 0) minimising, over requests with measured per-plan
        timings, the fixed plan's time below the gate plus the picked plan's time and the selection
        overhead above it. Returns {gate_log2: total ms}."""
        torch = _torch()
        y = self.forward(torch.from_numpy(X).to(self.device)).cpu().numpy()
        ms = np.where(M > 0, np.exp2(T) * 1e-3, np.nan)
        tf, tl, ns = [], [], []
        for i, (dt, n, _, _) in enumerate(requests):
            nb = torch.tensor([], dtype=torch_dtype(dt)).element_size()
            tf.append(float(ms[i, PLANS.index(fixed_plan(table, nb, n))]))
            tl.append(float(ms[i, int(np.argmin(np.where(M[i] > 0, y[i], np.inf)))]) + overhead_ms)
            ns.append(n)
        tf, tl, ns = np.array(tf), np.array(tl), np.array(ns)
        totals = {g: float(np.where(ns < (1 << g), tf, tl).sum()) for g in range(0, 31)}
        best = min(totals, key=lambda g: (totals[g], g))
        self.min_count = 0 if best == 0 else (1 << best)
        return totals


def init_params(F, H, O, seed):
    rng = np.random.default_rng(seed)
    w1 = rng.normal(0, math.sqrt(2.0 / F), (H, F))
    w2 = rng.normal(0, math.sqrt(2.0 / H), (H, H))
    w3 = rng.normal(0, math.sqrt(1.0 / H), (O, H))
    return np.concatenate([w1.ravel(), np.zeros(H), w2.ravel(), np.zeros(H), w3.ravel(), np.full(O, 6.0)]).astype(np.float32)


# ---- CPU double-precision reference of the network (checks the GPU kernels) ------------------

def ref_unpack(p, F, H, O):
    i = 0
    out = []
    for shape in ((H, F), (H,), (H, H), (H,), (O, H), (O,)):
        k = int(np.prod(shape))
        out.append(p[i:i + k].reshape(shape))
        i += k
    return out


def ref_forward(p, x, F, H, O):
    w1, b1, w2, b2, w3, b3 = ref_unpack(np.asarray(p, np.float64), F, H, O)
    h1 = np.maximum(x @ w1.T + b1, 0.0)
    h2 = np.maximum(h1 @ w2.T + b2, 0.0)
    return h1, h2, h2 @ w3.T + b3


def ref_loss_grad(p, x, t, m, F, H, O):
    """Masked MSE and its gradient by backpropagation, in float64."""
    x = np.asarray(x, np.float64)
    w1, b1, w2, b2, w3, b3 = ref_unpack(np.asarray(p, np.float64), F, H, O)
    h1, h2, y = ref_forward(p, x, F, H, O)
    msum = max(float(m.sum()), 1.0)
    loss = float((((y - t) ** 2) * m).sum() / msum)
    dy = 2.0 * m * (y - t) / msum
    gw3, gb3 = dy.T @ h2, dy.sum(0)
    d2 = (dy @ w3) * (h2 > 0)
    gw2, gb2 = d2.T @ h1, d2.sum(0)
    d1 = (d2 @ w2) * (h1 > 0)
    gw1, gb1 = d1.T @ x, d1.sum(0)
    return loss, np.concatenate([gw1.ravel(), gb1, gw2.ravel(), gb2, gw3.ravel(), gb3])


# ---- autotuner (the fixed plan) -----------------------------------------------------------------

def size_bucket(n: int) -> int:
    return int(round(math.log2(max(2, n))))


def autotune(log2_sizes=range(10, 29, 2), dtypes=("u8", "u16", "u32", "u64"), reps: int = 3, log=None, seed: int = 7):
    """Times every plan without the presort check on uniform keys for each key width and size and
    keeps the fastest: {"<bytes>:<log2 n>": [algo, cfg, check]}."""
    torch = _torch()
    gen = torch.Generator(device="cuda").manual_seed(seed)
    table = {}
    max_n = 1 << max(log2_sizes)
    arena = Arena(max_n)
    for dt in dtypes:
        for lg in log2_sizes:
            n = 1 << lg
            keys = make_keys(dt, n, "uniform", gen)
            flags = decode(keytype_of(keys.dtype))
            best, best_t = None, float("inf")
            for p in PLANS:
                if p[2] or not plan_valid(p, n, keys.element_size()):
                    continue
                out, buf = arena.views(keys)
                sort_with_plan(keys, out, buf, arena.ws, p, flags)
                t = time_call(lambda: sort_with_plan(keys, out, buf, arena.ws, p, flags), reps)
                if t < best_t:
                    best, best_t = p, t
            table[f"{keys.element_size()}:{lg}"] = list(best)
            if log:
                log(f"autotune {dt} 2^{lg}: {best} {best_t:.3f} ms")
            del keys
    return table


def fixed_plan(table: dict, nbytes: int, n: int):
    lg = size_bucket(n)
    keys = sorted(int(k.split(":")[1]) for k in table if int(k.split(":")[0]) == nbytes)
    if not keys:
        return (0, -1, 0)
    for b in sorted(keys, key=lambda v: (abs(v - lg), v)):
        p = tuple(table[f"{nbytes}:{b}"])
        if plan_valid(p, n, nbytes):
            return p
    return PLANS[0]


def load_table(path=None):
    with open(path or os.path.join(TUNING_DIR, "autotune_table.json")) as f:
        return json.load(f)


# ---- segmented batches: tile configuration and LDS tier ------------------------------------------

def segmented_batch(nseg: int, total_log2=None, seed: int = 0x5EED0005):
    """CSR offsets of nseg segments with lengths floor(2^U(4, 20)) (the 65,536-array batch workload:
    about 6.2e9 keys for 65,536 segments), or those lengths scaled to about 2^total_log2 keys in
    total when total_log2 is given."""
    rng = np.random.default_rng(seed)
    lens = np.floor(2.0 ** rng.uniform(4, 20, nseg))
    if total_log2 is not None:
        lens = np.maximum(1, np.floor(lens * ((1 << total_log2) / lens.sum())))
    lens = lens.astype(np.int64)
    off = np.zeros(nseg + 1, np.uint64)
    off[1:] = np.cumsum(lens)
    return off


def autotune_segmented(keys, offsets, direction=0, mode=64, reps: int = 3, log=None):
    """Times the captured plan of one batch for every (tile configuration, LDS tier) pair and keeps
    the fastest: returns ((tile_cfg, small_cfg), {"tile_cfg:small_cfg": ms})."""
    torch = _torch()
    from . import SegmentedPlan, SEGMENTED_TILE_CONFIGS, SMALL_CAPS
    out, buf = torch.empty_like(keys), torch.empty_like(keys)
    np_dt = {torch.int32: np.int32, torch.uint32: np.uint32, torch.int64: np.int64, torch.uint64: np.uint64,
             torch.float32: np.float32, torch.float64: np.float64, torch.int16: np.int16, torch.uint16: np.uint16,
             torch.int8: np.int8, torch.uint8: np.uint8}[keys.dtype]
    times = {}
    for sc in range(len(SMALL_CAPS)):
        for tc in range(len(SEGMENTED_TILE_CONFIGS)):
            plan = SegmentedPlan(np_dt, offsets, direction, mode, tile_cfg=tc, small_cfg=sc)
            plan.capture(keys, out, buf)
            plan.launch()
            times[f"{tc}:{sc}"] = time_call(plan.launch, reps)
            plan.close()
            if log:
                log(f"segmented {keys.dtype} tile_cfg {tc} small_cfg {sc}: {times[f'{tc}:{sc}']:.3f} ms")
    best = min(times, key=times.get)
    return tuple(int(v) for v in best.split(":")), times


def load_segmented_table(path=None):
    """{"<key bytes>": [tile_cfg, small_cfg]} from the segmented sweep (tuning/segmented_table.json)."""
    with open(path or os.path.join(TUNING_DIR, "segmented_table.json")) as f:
        return json.load(f)["best"]


# ---- training data and evaluation ----------------------------------------------------------------

def collect(requests, reps: int = 3, log=None):
    """Features and per-plan log2(us) timings of a list of requests (dtype, n, dist, seed)."""
    torch = _torch()
    max_n = max(r[1] for r in requests)
    arena = Arena(max_n)
    X = np.zeros((len(requests), NUM_FEATURES), np.float32)
    T = np.zeros((len(requests), NUM_PLANS), np.float32)
    M = np.zeros((len(requests), NUM_PLANS), np.float32)
    for i, (dt, n, dist, seed) in enumerate(requests):
        gen = torch.Generator(device="cuda").manual_seed(seed)
        keys = make_keys(dt, n, dist, gen)
        flags = decode(keytype_of(keys.dtype))
        X[i] = features(keys, flags, arena.feat).cpu().numpy()
        t = time_plans(keys, flags, arena, reps)
        ok = ~np.isnan(t)
        T[i, ok] = np.log2(np.maximum(t[ok] * 1e3, 1e-3))
        M[i, ok] = 1.0
        if log and (i % 50 == 0):
            log(f"collect {i}/{len(requests)} {dt} n={n} {dist}")
        del keys
    return X, T, M


def request_weights(T, M, power: float):
    """Per-request training weights: (best time / median best time) ** power, clipped to [1/16, 16]
    (T holds log2 microseconds, M marks the timed plans). A squared error in log2 time costs the
    same relative time at every size, but a stream's total is its large requests'; power 0 weighs
    every request alike."""
    best = np.where(M > 0, T, np.inf).min(axis=1)
    best = np.where(np.isfinite(best), best, np.nan)
    med = np.nanmedian(best)
    w = np.exp2(power * (np.nan_to_num(best, nan=med) - med))
    return np.clip(w, 1.0 / 16, 16.0).astype(np.float32)


def selection_overhead_ms(selector: Selector, reps: int = 50) -> float:
    """Median wall time of rsbd8_hip_select_plan (sampling, network, the choice's round trip) on a
    2^20-key array: the fixed cost a learned choice must earn back."""
    torch = _torch()
    gen = torch.Generator(device="cuda").manual_seed(99)
    keys = make_keys("u32", 1 << 20, "uniform", gen)
    flags = decode(keytype_of(keys.dtype))
    selector.select(keys, flags)
    torch.cuda.synchronize()
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        selector.select(keys, flags)
        times.append((time.perf_counter() - t0) * 1e3)
    return float(np.median(times))


def regret(requests, X, T, M, selector: Selector, table: dict):
    """Offline comparison on measured per-plan timings (no selection overhead): total time of the
    fixed plan, of the network's pick and of the best plan per request, in ms."""
    torch = _torch()
    y = selector.forward(torch.from_numpy(X).to(selector.device)).cpu().numpy()
    ms = np.where(M > 0, np.exp2(T) * 1e-3, np.nan)
    tot = {"fixed_ms": 0.0, "learned_ms": 0.0, "best_ms": 0.0}
    for i, (dt, n, _, _) in enumerate(requests):
        nb = torch.tensor([], dtype=torch_dtype(dt)).element_size()
        pf = fixed_plan(table, nb, n)
        pred = np.where(M[i] > 0, y[i], np.inf)
        tot["fixed_ms"] += float(ms[i, PLANS.index(pf)]) if pf in PLANS else float("nan")
        tot["learned_ms"] += float(ms[i, int(np.argmin(pred))])
        tot["best_ms"] += float(np.nanmin(ms[i]))
    tot["learned_regret"] = tot["learned_ms"] / tot["best_ms"] - 1.0
    tot["fixed_regret"] = tot["fixed_ms"] / tot["best_ms"] - 1.0
    return tot


def evaluate(requests, selector: Selector, table: dict, verify: bool = True, log=None, max_n=None, reps: int = 1):
    """Sorts every request with the fixed and the learned plan. Returns the total completion times
    (time_call: from an idle device to the end of the sort, in ms; with reps > 1 the median of reps
    timings per request, the two plans timed alternately) and checks that both outputs are
    bit-identical. The learned time is that of the whole rsbd8_hip_radixsortcopynoalloc_learned
    call: sampling, network and the host round trip of the choice (above the size gate), then the
    sort. runs holds the totals of each of the reps timing rounds on its own (the run-to-run
    spread of the speed-up)."""
    torch = _torch()
    max_n = max_n or max(r[1] for r in requests)
    arena = Arena(max_n)
    out2 = torch.empty(max_n * 8, dtype=torch.uint8, device="cuda")
    tot_fixed = tot_learned = 0.0
    run_fixed, run_learned = [0.0] * reps, [0.0] * reps
    chosen = {}
    mismatches = 0
    for i, (dt, n, dist, seed) in enumerate(requests):
        gen = torch.Generator(device="cuda").manual_seed(seed)
        keys = make_keys(dt, n, dist, gen)
        flags = decode(keytype_of(keys.dtype))
        out, buf = arena.views(keys)
        nb = keys.element_size()
        pf = fixed_plan(table, nb, n)

        def fixed():
            sort_with_plan(keys, out, buf, arena.ws, pf, flags)

        def learned():
            return selector.sort(keys, out2[: n * nb].view(keys.dtype), buf, arena.ws, flags, below=pf)

        fixed()
        idx = learned()
        tf, tl = [], []
        for k in range(reps):
            tf.append(time_call(fixed, 1))
            tl.append(time_call(learned, 1))
            run_fixed[k] += tf[-1]
            run_learned[k] += tl[-1]
        fixed()
        idx = learned()
        key = PLANS[idx] if idx >= 0 else "below gate (fixed plan)"
        chosen[key] = chosen.get(key, 0) + 1
        if verify and not torch.equal(out.view(torch.uint8)[: n * nb], out2[: n * nb]):
            mismatches += 1
        tot_fixed += float(np.median(tf))
        tot_learned += float(np.median(tl))
        if log and i % 256 == 0:
            log(f"eval {i}/{len(requests)} fixed {tot_fixed:.1f} ms learned {tot_learned:.1f} ms")
        del keys
    runs = [{"fixed_ms": f, "learned_ms": l, "speedup": f / max(l, 1e-9)} for f, l in zip(run_fixed, run_learned)]
    return {"requests": len(requests), "reps": reps, "fixed_ms": tot_fixed, "learned_ms": tot_learned,
            "speedup": tot_fixed / max(tot_learned, 1e-9), "mismatches": mismatches, "runs": runs,
            "chosen": {str(k): v for k, v in sorted(chosen.items(), key=lambda kv: -kv[1])}}
