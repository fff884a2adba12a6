"""radixsortbidi8_amd -- the rsbd8 stable LSD radix sort on MI355X (gfx950).

Python mirror of the reference's four sorting templates (Radixsortbidi8.hpp = H):

  radixsort(array, direction, mode)                       H:60397-60466  (in place, allocating)
  radixsortcopy(input, output, direction, mode)           H:60468-60588  (copy, allocating)
  radixsortnoalloc(input, buffer, workspace, movetobuffer, direction, mode)    H:60155-60395
  radixsortcopynoalloc(input, output, buffer, workspace, direction, mode)      H:60121-60242

plus the indirection forms (pointer arrays with byte offsets, H:60590-61738, and records addressed
by 32-bit indices) and key/payload pairs.  Arrays are torch tensors on the GPU (host numpy arrays
are accepted by the allocating forms and staged through device memory).  The sorting itself runs
in the HIP library ``lib/librsbd8_hip.so`` (built from ``csrc/`` by ``build.py``); importing this
package without that library raises.

``sortingmode`` / ``sortingdirection`` carry the reference's enum values (H:212-265) and
``decode`` reproduces its compile-time mode decoding (H:59778-59904).
"""
from __future__ import annotations

import ctypes
import enum
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "librsbd8_hip.so")


class sortingmode(enum.IntEnum):
    """rsbd8::sortingmode (H:212-250)."""
    native = 64
    nativeabs = 65
    nativetieredabs = 66
    forceunsigned = 0
    specialsigned = 0
    forcesigned = 2
    forceabssigned = 3
    forcefloatingp = 6
    forceabsfloatingp = 7
    specialunsigned = 7
    specialfloatingp = 4
    forcetieredabsfloatingp = 5
    forcetieredabssigned = 1


class sortingdirection(enum.IntEnum):
    """rsbd8::sortingdirection (H:252-265)."""
    ascfwdorder = 0
    dscrevorder = 3
    dscfwdorder = 1
    ascrevorder = 2


ABS, SIGN, FLTP, DESC, REV = 1, 2, 4, 8, 16
LAYOUT_KEYS, LAYOUT_PAIRS, LAYOUT_INDIRECT = 0, 1, 2
U8, U16, U32, U64, I8, I16, I32, I64, F32, F64 = range(10)
# keys wider than 64 bits (H:1061-1804, H:2186-2404): the 80-bit extended float in its 10-, 12- and
# 16-byte records (longdoubletest80/96/128; the 16-byte one is x86-64's long double, numpy's
# longdouble) and 128-bit integers, passed as byte records
LD80, LD96, LD128, U128, I128 = 10, 11, 12, 13, 14
WIDE_RECORD_BYTES = {LD80: 10, LD96: 12, LD128: 16, U128: 16, I128: 16}
TILE_CONFIGS = ((256, 8), (256, 16), (512, 8), (1024, 12), (256, 24))
SEGMENTED_TILE_CONFIGS = ((256, 8), (256, 16), (512, 8), (512, 16), (256, 24))

_lib = None


class Rsbd8Error(RuntimeError):
    pass


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise Rsbd8Error(f"HIP library missing: {LIB_PATH} (run radixsortbidi8_amd/build.py)")
    L = ctypes.CDLL(LIB_PATH)
    vp, sz, ci, cu = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_uint
    pd = ctypes.c_ssize_t
    sig = {
        "rsbd8_hip_decode": ([ci, cu, cu, ci, ctypes.POINTER(cu)], ci),
        "rsbd8_hip_workspace_size": ([ci, sz, ci], sz),
        "rsbd8_hip_workspace_init": ([vp, sz, vp], ci),
        "rsbd8_hip_radixsortcopynoalloc": ([ci, cu, sz, vp, sz, vp, vp, vp, vp], ci),
        "rsbd8_hip_radixsortnoalloc": ([ci, cu, sz, vp, sz, vp, vp, ci, vp], ci),
        "rsbd8_hip_radixsortcopy": ([ci, cu, sz, vp, vp, vp], ci),
        "rsbd8_hip_radixsort": ([ci, cu, sz, vp, vp], ci),
        "rsbd8_hip_radixsortcopynoalloc_indirect": ([ci, cu, sz, vp, sz, vp, vp, pd, ci, pd, sz, vp], ci),
        "rsbd8_hip_radixsortnoalloc_indirect": ([ci, cu, sz, vp, sz, vp, vp, ci, pd, ci, pd, sz, vp], ci),
        "rsbd8_hip_sort_indices": ([ci, cu, sz, vp, sz, vp, sz, sz, vp, vp, vp], ci),
        "rsbd8_hip_sort_pairs": ([ci, cu, sz, vp, sz, vp, vp, vp, vp, vp, vp, vp], ci),
        "rsbd8_hip_set_tile_config": ([ci], ci),
        "rsbd8_hip_set_algorithm": ([ci], ci),
        "rsbd8_hip_set_sample_target": ([ci], ci),
        "rsbd8_hip_last_error": ([], ctypes.c_char_p),
        "rsbd8_hip_segmented_plan_create": ([ci, cu, sz, vp, ci, ci, ctypes.POINTER(vp)], ci),
        "rsbd8_hip_segmented_run": ([vp, vp, vp, vp, vp, vp], ci),
        "rsbd8_hip_segmented_capture": ([vp, vp, vp, vp, vp, vp], ci),
        "rsbd8_hip_segmented_launch_graph": ([vp, vp], ci),
        "rsbd8_hip_segmented_plan_info": ([vp, ctypes.POINTER(ctypes.c_uint64)], ci),
        "rsbd8_hip_segmented_default_config": ([ci, cu, ctypes.POINTER(ci), ctypes.POINTER(ci)], ci),
        "rsbd8_hip_segmented_plan_destroy": ([vp], None),
        "rsbd8_hip_profile_begin": ([], ci),
        "rsbd8_hip_profile_end": ([ctypes.POINTER(ctypes.c_double), ctypes.POINTER(cu), ci], ci),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def lib():
    return _load()


def exported_symbols():
    """Names the C ABI exports (include/rsbd8_hip.h)."""
    return sorted(n for n in dir(_load()) if n.startswith("rsbd8_hip_"))


def _torch():
    import torch
    return torch


_TORCH_KEYTYPE = None


def keytype_of(dtype) -> int:
    """Key type code of a torch or numpy dtype (bool sorts as its unsigned byte, H:201-202)."""
    global _TORCH_KEYTYPE
    try:
        nd = np.dtype(dtype)
        table = {np.dtype(np.uint8): U8, np.dtype(np.bool_): U8, np.dtype(np.uint16): U16,
                 np.dtype(np.uint32): U32, np.dtype(np.uint64): U64, np.dtype(np.int8): I8,
                 np.dtype(np.int16): I16, np.dtype(np.int32): I32, np.dtype(np.int64): I64,
                 np.dtype(np.float32): F32, np.dtype(np.float64): F64}
        if nd in table:
            return table[nd]
    except TypeError:
        pass
    torch = _torch()
    if _TORCH_KEYTYPE is None:
        _TORCH_KEYTYPE = {torch.uint8: U8, torch.bool: U8, torch.uint16: U16, torch.uint32: U32,
                          torch.uint64: U64, torch.int8: I8, torch.int16: I16, torch.int32: I32,
                          torch.int64: I64, torch.float32: F32, torch.float64: F64}
    if dtype in _TORCH_KEYTYPE:
        return _TORCH_KEYTYPE[dtype]
    raise TypeError(f"unsupported key dtype {dtype}")


def decode(keytype: int, mode=sortingmode.native, direction=sortingdirection.ascfwdorder,
           indirect: bool = False) -> int:
    """The five decoded flags of H:59778-59904 as bits ABS|SIGN|FLTP|DESC|REV."""
    out = ctypes.c_uint(0)
    _check(_load().rsbd8_hip_decode(int(keytype), int(mode), int(direction), int(bool(indirect)), ctypes.byref(out)))
    return out.value


def _check(rc: int):
    if rc != 0:
        msg = _load().rsbd8_hip_last_error().decode(errors="replace")
        raise Rsbd8Error(f"rsbd8_hip error {rc}: {msg}")


def _stream_ptr(stream=None):
    torch = _torch()
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def _ptr(t, device: bool = True):
    """Address of a tensor's data for the C ABI. The noalloc forms hand it to GPU kernels, so the
    tensor must live on the GPU (device=True); a non-contiguous view would sort the wrong elements,
    so it must be contiguous. The allocating forms also take host tensors (device=False)."""
    if t is None:
        return None
    if device and not t.is_cuda:
        raise Rsbd8Error("array must be a GPU tensor (the noalloc forms launch kernels on it); "
                         "radixsort/radixsortcopy accept host arrays")
    if not t.is_contiguous():
        raise Rsbd8Error("array must be contiguous")
    return ctypes.c_void_p(t.data_ptr())


def _is_tensor(x):
    try:
        torch = _torch()
    except ImportError:
        return False
    return isinstance(x, torch.Tensor)


def set_tile_config(cfg: int) -> int:
    """Force a scatter tile configuration (index into TILE_CONFIGS) or -1 for automatic."""
    prev = _load().rsbd8_hip_set_tile_config(int(cfg))
    if prev < -1:
        _check(prev)
    return prev


ALGO_AUTO, ALGO_ONESWEEP, ALGO_COUNTSCAN, ALGO_MSD2 = 0, 1, 2, 4


def set_algorithm(algo: int) -> int:
    """Force the pass algorithm (ALGO_ONESWEEP, ALGO_COUNTSCAN, or ALGO_MSD2: two MSD partitions and
    a per-bucket LDS sort for 4/8-byte keys, by the top two digits for keys-only integer sorts under
    an injective transform, else by splitters sampled from the keys; 1/2-byte keys fall back to
    ALGO_COUNTSCAN) or ALGO_AUTO; returns the previous."""
    prev = _load().rsbd8_hip_set_algorithm(int(algo))
    if prev < 0:
        _check(prev)
    return prev


def set_sample_target(keys: int) -> int:
    """Mean bucket length the splitter-sampled partitions of ALGO_MSD2 aim at (0 restores the
    default, 4096); returns the previous setting."""
    prev = _load().rsbd8_hip_set_sample_target(int(keys))
    if prev < 0:
        _check(prev)
    return prev


def profile_begin():
    """Start recording HIP-event timings of every sort launch (on its own stream)."""
    _check(_load().rsbd8_hip_profile_begin())


def profile_end():
    """Stop recording; returns {kind: (total_ms, launches)} for kinds 'hist', 'scatter'
    (executed passes), 'scatter_skipped', 'countscan' (count + scan of executed passes),
    'countscan_skipped', and for the two-level MSD plan when it ran 'msd_scatter' (its two
    partition scatters), 'msd_local' (the per-bucket LDS sort) and 'msd_countscan'."""
    kinds = ("hist", "scatter", "scatter_skipped", "countscan", "countscan_skipped", "msd_scatter", "msd_local",
             "msd_countscan")
    ms = (ctypes.c_double * len(kinds))()
    cnt = (ctypes.c_uint * len(kinds))()
    _check(_load().rsbd8_hip_profile_end(ms, cnt, len(kinds)))
    return {k: (ms[i], cnt[i]) for i, k in enumerate(kinds)}


# ---- workspace (the pzeroedindices equivalent, allocatearray H:59921-60065) -----------------

class Workspace:
    """Device workspace for the noalloc sorts; zeroed on creation, left reusable by every sort."""

    def __init__(self, keytype: int, count: int, layout: int = LAYOUT_KEYS, device=None):
        torch = _torch()
        self.keytype, self.count, self.layout = keytype, count, layout
        self.nbytes = int(_load().rsbd8_hip_workspace_size(keytype, count, layout))
        self.tensor = torch.empty(self.nbytes, dtype=torch.uint8, device=device or "cuda")
        _check(_load().rsbd8_hip_workspace_init(_ptr(self.tensor), self.nbytes, _stream_ptr()))

    def ptr(self):
        return _ptr(self.tensor)


def allocatearray(dtype, count: int, layout: int = LAYOUT_KEYS, device=None) -> Workspace:
    return Workspace(keytype_of(dtype), count, layout, device)


def _flags(dtype, mode, direction, indirect=False, keytype=None):
    kt = keytype_of(dtype) if keytype is None else keytype
    return kt, decode(kt, mode, direction, indirect)


def _count(a, keytype: int) -> int:
    """Number of keys in an array: its elements, or for the wide key types its bytes over the
    record size (wide records travel as uint8 tensors or arrays, WIDE_RECORD_BYTES each)."""
    if keytype in WIDE_RECORD_BYTES:
        nbytes = a.numel() * a.element_size() if _is_tensor(a) else a.nbytes
        return nbytes // WIDE_RECORD_BYTES[keytype]
    return a.numel() if _is_tensor(a) else a.size


# ---- the four templates without indirection ----------------------------------------------

def radixsortcopynoalloc(input, output, buffer, workspace: Workspace,
                         direction=sortingdirection.ascfwdorder, mode=sortingmode.native, stream=None, keytype=None):
    """rsbd8::radixsortcopynoalloc (H:60121-60153; 8-bit keys H:60186-60242, buffer may be None).
    keytype overrides the key type of input's dtype (the wide types: uint8 records)."""
    kt, f = _flags(input.dtype, mode, direction, keytype=keytype)
    _check(_load().rsbd8_hip_radixsortcopynoalloc(kt, f, _count(input, kt), workspace.ptr(), workspace.nbytes,
                                                   _ptr(input), _ptr(output), _ptr(buffer), _stream_ptr(stream)))


def radixsortnoalloc(input, buffer, workspace: Workspace, movetobuffer: bool = False,
                     direction=sortingdirection.ascfwdorder, mode=sortingmode.native, stream=None, keytype=None):
    """rsbd8::radixsortnoalloc (H:60155-60184; 8-bit keys H:60244-60395)."""
    kt, f = _flags(input.dtype, mode, direction, keytype=keytype)
    _check(_load().rsbd8_hip_radixsortnoalloc(kt, f, _count(input, kt), workspace.ptr(), workspace.nbytes,
                                               _ptr(input), _ptr(buffer), int(bool(movetobuffer)),
                                               _stream_ptr(stream)))


def _host_call(fn, kt, f, arrays, count, stream):
    rc = fn(kt, f, count, *arrays, stream)
    if rc == -2:  # allocation failure: the reference returns false (H:60463)
        return False
    _check(rc)
    return True


def radixsortcopy(input, output, direction=sortingdirection.ascfwdorder, mode=sortingmode.native,
                  stream=None, keytype=None) -> bool:
    """rsbd8::radixsortcopy (H:60468-60588): torch GPU tensors or host numpy arrays."""
    if _is_tensor(input):
        kt, f = _flags(input.dtype, mode, direction, keytype=keytype)
        return _host_call(_load().rsbd8_hip_radixsortcopy, kt, f, (_ptr(input, False), _ptr(output, False)),
                          _count(input, kt), _stream_ptr(stream))
    inp = np.ascontiguousarray(input)
    assert output.flags.c_contiguous and output.dtype == inp.dtype and output.size == inp.size
    kt, f = _flags(inp.dtype, mode, direction, keytype=keytype)
    return _host_call(_load().rsbd8_hip_radixsortcopy, kt, f,
                      (inp.ctypes.data_as(ctypes.c_void_p), output.ctypes.data_as(ctypes.c_void_p)), _count(inp, kt),
                      None if stream is None else _stream_ptr(stream))


def radixsort(input, direction=sortingdirection.ascfwdorder, mode=sortingmode.native, stream=None,
              keytype=None) -> bool:
    """rsbd8::radixsort (H:60397-60466), in place: torch GPU tensor or host numpy array."""
    if _is_tensor(input):
        kt, f = _flags(input.dtype, mode, direction, keytype=keytype)
        return _host_call(_load().rsbd8_hip_radixsort, kt, f, (_ptr(input, False),), _count(input, kt),
                          _stream_ptr(stream))
    assert input.flags.c_contiguous
    kt, f = _flags(input.dtype, mode, direction, keytype=keytype)
    return _host_call(_load().rsbd8_hip_radixsort, kt, f, (input.ctypes.data_as(ctypes.c_void_p),), _count(input, kt),
                      None if stream is None else _stream_ptr(stream))


# ---- indirection and pairs -----------------------------------------------------------------

def radixsortcopynoalloc_pointers(input_ptrs, output_ptrs, workspace: Workspace, key_dtype,
                                  indirection1: int = 0, direction=sortingdirection.ascfwdorder,
                                  mode=sortingmode.native, second_level: bool = False,
                                  indirection2: int = 0, index2: int = 0, stream=None):
    """T**/type+offset/member indirection (H:60590-60627, H:61360-61393): input_ptrs is an int64
    tensor of device addresses; output_ptrs receives them sorted by the key they point at. key_dtype
    is a dtype or a key type code (the wide types)."""
    kt = key_dtype if isinstance(key_dtype, int) else keytype_of(key_dtype)
    f = decode(kt, mode, direction, True)
    _check(_load().rsbd8_hip_radixsortcopynoalloc_indirect(
        kt, f, input_ptrs.numel(), workspace.ptr(), workspace.nbytes, _ptr(input_ptrs), _ptr(output_ptrs),
        int(indirection1), int(bool(second_level)), int(indirection2), int(index2), _stream_ptr(stream)))


def radixsortnoalloc_pointers(input_ptrs, buffer_ptrs, workspace: Workspace, key_dtype, indirection1: int = 0,
                              movetobuffer: bool = False, direction=sortingdirection.ascfwdorder,
                              mode=sortingmode.native, second_level: bool = False, indirection2: int = 0,
                              index2: int = 0, stream=None):
    """In-place pointer sort (H:61017-61045, H:61395-61427)."""
    kt = key_dtype if isinstance(key_dtype, int) else keytype_of(key_dtype)
    f = decode(kt, mode, direction, True)
    _check(_load().rsbd8_hip_radixsortnoalloc_indirect(
        kt, f, input_ptrs.numel(), workspace.ptr(), workspace.nbytes, _ptr(input_ptrs), _ptr(buffer_ptrs),
        int(bool(movetobuffer)), int(indirection1), int(bool(second_level)), int(indirection2), int(index2),
        _stream_ptr(stream)))


def sort_indices(records, stride: int, key_offset: int, idx_in, idx_out, workspace: Workspace, key_dtype,
                 direction=sortingdirection.ascfwdorder, mode=sortingmode.native, stream=None):
    """Sort 32-bit record indices by the key at records + idx*stride + key_offset."""
    kt = keytype_of(key_dtype)
    f = decode(kt, mode, direction, True)
    _check(_load().rsbd8_hip_sort_indices(kt, f, idx_in.numel(), workspace.ptr(), workspace.nbytes,
                                           _ptr(records), int(stride), int(key_offset), _ptr(idx_in),
                                           _ptr(idx_out), _stream_ptr(stream)))


def sort_pairs(keys_in, keys_out, keys_buffer, payload_in, payload_out, payload_buffer, workspace: Workspace,
               direction=sortingdirection.ascfwdorder, mode=sortingmode.native, indirect: bool = True, stream=None,
               keytype=None):
    """Stable key sort moving a u32 payload along (payload_in None: the source index)."""
    kt = keytype_of(keys_in.dtype) if keytype is None else keytype
    f = decode(kt, mode, direction, indirect)
    _check(_load().rsbd8_hip_sort_pairs(kt, f, _count(keys_in, kt), workspace.ptr(), workspace.nbytes, _ptr(keys_in),
                                         _ptr(keys_out), _ptr(keys_buffer), _ptr(payload_in), _ptr(payload_out),
                                         _ptr(payload_buffer), _stream_ptr(stream)))


# ---- segmented batches --------------------------------------------------------------------

SMALL_CAPS = (1024, 4096, 8192)


def segmented_default_config(key_dtype, direction=sortingdirection.ascfwdorder, mode=sortingmode.native):
    """The library's automatic (tile_cfg, small_cfg) for a segmented plan of this key type and mode."""
    kt = keytype_of(key_dtype)
    tc, sc = ctypes.c_int(), ctypes.c_int()
    _check(_load().rsbd8_hip_segmented_default_config(kt, decode(kt, mode, direction), ctypes.byref(tc), ctypes.byref(sc)))
    return tc.value, sc.value


class SegmentedPlan:
    """A batch of independent arrays (CSR offsets) sorted with one set of flags. Short segments are
    sorted in LDS by one persistent kernel, long ones by batched onesweep passes; ``capture`` records
    the whole sequence into a hipGraph that ``launch`` replays with one launch."""

    def __init__(self, key_dtype, offsets: np.ndarray, direction=sortingdirection.ascfwdorder,
                 mode=sortingmode.native, tile_cfg: int = -1, small_cfg: int = -1):
        torch = _torch()
        self.keytype = keytype_of(key_dtype)
        self.flags = decode(self.keytype, mode, direction)
        off = np.ascontiguousarray(offsets, dtype=np.uint64)
        self.nseg = off.size - 1
        self.total = int(off[-1] - off[0]) if self.nseg > 0 else 0
        self.dev_offsets = torch.from_numpy(off.view(np.int64)).cuda()
        h = ctypes.c_void_p()
        _check(_load().rsbd8_hip_segmented_plan_create(self.keytype, self.flags, self.nseg,
                                                       off.ctypes.data_as(ctypes.c_void_p), int(tile_cfg),
                                                       int(small_cfg), ctypes.byref(h)))
        self._h = h
        self._keep = None

    def info(self):
        v = (ctypes.c_uint64 * 7)()
        _check(_load().rsbd8_hip_segmented_plan_info(self._h, v))
        return dict(zip(("short", "long", "tiles", "tile_keys", "lds_capacity", "total", "plan_bytes"), list(v)))

    def run(self, keys_in, keys_out, buffer, stream=None):
        _check(_load().rsbd8_hip_segmented_run(self._h, _ptr(keys_in), _ptr(keys_out), _ptr(buffer),
                                               _ptr(self.dev_offsets), _stream_ptr(stream)))

    def capture(self, keys_in, keys_out, buffer, stream=None):
        torch = _torch()
        s = stream if stream is not None else torch.cuda.Stream()
        self._keep = (keys_in, keys_out, buffer)
        torch.cuda.synchronize()
        _check(_load().rsbd8_hip_segmented_capture(self._h, _ptr(keys_in), _ptr(keys_out), _ptr(buffer),
                                                   _ptr(self.dev_offsets), _stream_ptr(s)))

    def launch(self, stream=None):
        _check(_load().rsbd8_hip_segmented_launch_graph(self._h, _stream_ptr(stream)))

    def close(self):
        if self._h:
            _load().rsbd8_hip_segmented_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
