"""MI355X-native QSketch: Python binding of the C ABI in include/qsketch.h.

The compute path is the HIP library ``libqsketch_amd.so`` built in-tree by
``qsketch_amd.build``.  Importing this package fails loudly when the library is
missing; there is no CPU fallback.

``Sketch_GPU_Sub_Warp_Level`` mirrors the reference's C++ class of the same
name (include/sketch/device_sketch/device_sketch_v3.h:524-634) and its base
interface ``Sketch<Key_T, Count_T>`` (include/sketch/sketch.h:5-28):
``insert(keys, n)``, ``search(keys, n, counts)``, ``clear()``.  Keys and counts
are device buffers (torch tensors on a HIP device, or raw device addresses).
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libqsketch_amd.so")

QS_OK = 0
QS_ERR_PARAM = -1
QS_ERR_NOMEM = -2
QS_ERR_POOL = -3
QS_ERR_HIP = -4
QS_ERR_STATE = -5
QS_ERR_NODEVICE = -6
QS_ERR_TABLE = -7
QS_ERR_INTERNAL = -8
QS_LITERAL = 0
QS_PAPER = 1
MLP_PARAMS = 8385
NOALLOC_MAX_KEYS = 1 << 20     # include/qsketch.h QS_NOALLOC_MAX_KEYS
NOALLOC_MAX_ALLOCS = 1 << 16   # include/qsketch.h QS_NOALLOC_MAX_ALLOCS

# lib/default.cu:39-54 and src/main.cu:390-391
NEXT_LEVEL_ID_START = 1024
HASH_MASK_ONES = 3
HASH_MASK_TABLE_SZ = 1021
SEED_SZ = 1


class QSketchError(RuntimeError):
    def __init__(self, code: int, what: str = ""):
        self.code = code
        msg = _lib().qs_strerror(code).decode() if _LIB is not None else str(code)
        super().__init__(f"{what}: {msg} ({code})" if what else f"{msg} ({code})")


class qs_params(C.Structure):
    _fields_ = [
        ("n_low", C.c_uint64), ("m_low", C.c_uint32), ("m_high", C.c_uint32),
        ("n_high", C.c_uint64), ("seed_sz", C.c_uint32), ("n_masks", C.c_uint32),
        ("mask_ones", C.c_uint32), ("link_start", C.c_uint32), ("low_limit", C.c_uint32),
        ("key_bytes", C.c_uint32), ("semantics", C.c_int32), ("hash_seed", C.c_uint32),
        ("rng_seed", C.c_uint64), ("chunk_keys", C.c_uint64),
    ]


class qs_stats(C.Structure):
    _fields_ = [
        ("n_low", C.c_uint64), ("n_high", C.c_uint64), ("n_masks", C.c_uint64),
        ("bytes_low", C.c_uint64), ("bytes_high", C.c_uint64), ("keys_inserted", C.c_uint64),
        ("chunks", C.c_uint64), ("chunks_with_fixup", C.c_uint64), ("fixup_keys", C.c_uint64),
        ("linked_buckets", C.c_uint64), ("next_id", C.c_uint32), ("seed", C.c_uint32),
        ("ordered_segments", C.c_uint64), ("long_segments", C.c_uint64),
        ("n_sketches", C.c_uint32), ("pool_error", C.c_uint32), ("hot_keys", C.c_uint64),
    ]


TIMING_KERNELS = ["sort_up", "sort_scan", "sort_down", "seg", "ordered", "assign", "tail", "query", "hot"]


class qs_timing(C.Structure):
    _fields_ = [
        ("insert_kernel_ms", C.c_double), ("insert_kernel_launches", C.c_uint64),
        ("insert_keys", C.c_uint64), ("query_kernel_ms", C.c_double),
        ("query_kernel_launches", C.c_uint64), ("query_keys", C.c_uint64),
        ("kernel_ms", C.c_double * 9), ("kernel_launches", C.c_uint64 * 9),
    ]


class qs_tune_entry(C.Structure):
    _fields_ = [("width", C.c_uint32), ("items_per_lane", C.c_uint32), ("seg_ms", C.c_double),
                ("insert_ms", C.c_double), ("checksum", C.c_uint64)]


class qs_tune_result(C.Structure):
    _fields_ = [("n", C.c_uint32), ("best", C.c_uint32), ("consistent", C.c_uint32), ("pad", C.c_uint32),
                ("e", qs_tune_entry * 16)]


# every symbol declared in include/qsketch.h and include/qsketch_gen.h etc.
_SIGS = {
    "qs_default_params": (C.c_int, [C.POINTER(qs_params), C.c_uint64, C.c_double]),
    "qs_find_greatest_prime": (C.c_uint64, [C.c_uint64]),
    "qs_create": (C.c_int, [C.POINTER(qs_params), C.POINTER(C.c_void_p)]),
    "qs_destroy": (None, [C.c_void_p]),
    "qs_insert": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "qs_query": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]),
    "qs_clear": (C.c_int, [C.c_void_p]),
    "qs_insert_async": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]),
    "qs_query_async": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p]),
    "qs_sync": (C.c_int, [C.c_void_p]),
    "qs_insert_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]),
    "qs_set_hot_filter": (C.c_int, [C.c_void_p, C.c_int]),
    "qs_export_state": (C.c_int, [C.c_void_p] + [C.c_void_p] * 5),
    "qs_import_state": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32]),
    "qs_get_stats": (C.c_int, [C.c_void_p, C.POINTER(qs_stats)]),
    "qs_strerror": (C.c_char_p, [C.c_int]),
    "qs_set_timing": (C.c_int, [C.c_void_p, C.c_int]),
    "qs_get_timing": (C.c_int, [C.c_void_p, C.POINTER(qs_timing)]),
    # sets of sketches and captured inserts
    "qs_multi_create": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(C.c_void_p)]),
    "qs_multi_count": (C.c_uint32, [C.c_void_p]),
    "qs_multi_params": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(qs_params)]),
    "qs_multi_insert": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "qs_multi_query": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "qs_multi_insert_async": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "qs_multi_query_async": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p,
                                       C.c_void_p]),
    "qs_multi_export_state": (C.c_int, [C.c_void_p, C.c_uint32] + [C.c_void_p] * 5),
    "qs_multi_import_state": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p,
                                        C.c_uint32, C.c_uint32]),
    "qs_graph_capture_insert": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                          C.POINTER(C.c_void_p)]),
    "qs_graph_launch": (C.c_int, [C.c_void_p, C.c_void_p]),
    "qs_graph_destroy": (None, [C.c_void_p]),
    # learned heavy-hitter form (include/qsketch_learned.h)
    "qs_learned_create": (C.c_int, [C.POINTER(qs_params), C.c_uint32, C.c_uint64, C.POINTER(C.c_void_p)]),
    "qs_learned_destroy": (None, [C.c_void_p]),
    "qs_learned_sketch": (C.c_void_p, [C.c_void_p]),
    "qs_mlp_get_params": (C.c_int, [C.c_void_p, C.c_void_p]),
    "qs_mlp_set_params": (C.c_int, [C.c_void_p, C.c_void_p]),
    "qs_mlp_step": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_float, C.c_float, C.c_float,
                              C.c_float, C.POINTER(C.c_float), C.c_void_p]),
    "qs_mlp_grad": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.POINTER(C.c_float),
                              C.c_void_p]),
    "qs_mlp_score": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p, C.c_int]),
    "qs_learned_set_threshold": (C.c_int, [C.c_void_p, C.c_float]),
    "qs_learned_threshold": (C.c_float, [C.c_void_p]),
    "qs_learned_insert": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64)]),
    "qs_learned_query": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p]),
    "qs_learned_clear": (C.c_int, [C.c_void_p]),
    "qs_learned_table": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64)]),
    "qs_learned_table_full": (C.c_int, [C.c_void_p]),
    "qs_learned_memory": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                    C.POINTER(C.c_uint64)]),
    # single-level sub-warp sketch (include/qsketch_single.h)
    "qs_sl_create": (C.c_int, [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64,
                               C.POINTER(C.c_void_p)]),
    "qs_sl_destroy": (None, [C.c_void_p]),
    "qs_sl_insert": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "qs_sl_query": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]),
    "qs_sl_clear": (C.c_int, [C.c_void_p]),
    "qs_sl_export": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "qs_set_insert_shape": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32]),
    "qs_get_insert_shape": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "qs_autotune_insert": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                                     C.POINTER(qs_tune_result)]),
    # SF sketch (include/qsketch_sf.h)
    "qs_sf_create": (C.c_int, [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64,
                               C.POINTER(C.c_void_p)]),
    "qs_sf_destroy": (None, [C.c_void_p]),
    "qs_sf_insert": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "qs_sf_query": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]),
    "qs_sf_clear": (C.c_int, [C.c_void_p]),
    "qs_sf_export": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    # calibration kernels (include/qsketch_calib.h)
    "qs_calib_copy": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p]),
    "qs_calib_fill": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint32, C.c_int, C.c_void_p]),
    "qs_calib_gather": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint32, C.c_uint64, C.c_uint64, C.c_void_p,
                                  C.c_void_p]),
    "qs_calib_rmw": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint32, C.c_uint64, C.c_uint64, C.c_int,
                                C.c_void_p]),
    # generators (include/qsketch_gen.h)
    "qs_zipf_create": (C.c_int, [C.c_double, C.c_uint64, C.c_int, C.POINTER(C.c_void_p)]),
    "qs_zipf_destroy": (None, [C.c_void_p]),
    "qs_gen_zipf_u32": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint64, C.c_void_p, C.c_void_p]),
    "qs_gen_zipf_u32_host": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint64, C.c_void_p]),
    "qs_gen_zipf_rank_host": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint64, C.c_void_p]),
    "qs_gen_uniform_u32": (C.c_int, [C.c_uint64, C.c_uint64, C.c_uint64, C.c_void_p, C.c_void_p]),
    "qs_gen_uniform_u32_host": (C.c_int, [C.c_uint64, C.c_uint64, C.c_uint64, C.c_void_p]),
    "qs_rank_to_key32": (C.c_uint32, [C.c_uint64, C.c_uint32]),
    "qs_gen_zipf_u64_noisy": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint64, C.c_double,
                                        C.c_void_p, C.c_void_p]),
    "qs_gen_zipf_u64_noisy_host": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint64, C.c_double,
                                             C.c_void_p]),
}

_LIB = None


def _lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m qsketch_amd.build` "
                "(there is no CPU fallback)")
        try:
            # torch ships its own HIP runtime: load it first, so the library
            # binds to the runtime already in the process instead of loading
            # a second one (two runtimes in one process leave torch no GPU)
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def lib():
    """The loaded C ABI (ctypes.CDLL)."""
    return _lib()


def check(rc: int, what: str = "") -> None:
    if rc != QS_OK:
        raise QSketchError(rc, what)


def _ptr(x) -> int:
    """Device address of a torch tensor (contiguous) or an int address."""
    if x is None:
        return 0
    if isinstance(x, int):
        return x
    if hasattr(x, "data_ptr"):
        if not x.is_contiguous():
            raise ValueError("tensor must be contiguous")
        if not x.is_cuda:
            raise ValueError("keys and counts must be device tensors")
        return x.data_ptr()
    raise TypeError(f"unsupported buffer {type(x)!r}")


def _keys_arg(keys, n, key_bytes: int, what: str):
    """(device address, n) of a key buffer.  For a tensor, n (default: all of
    it) must fit in it and its elements must be key_bytes wide; a raw address
    needs an explicit n, and the caller vouches for its size."""
    if hasattr(keys, "numel"):
        n = keys.numel() if n is None else int(n)
        if n < 0 or n > keys.numel():
            raise ValueError(f"{what}: n = {n} but the key tensor holds {keys.numel()} elements")
        if keys.element_size() != key_bytes:
            raise ValueError(f"{what}: the sketch takes {key_bytes}-byte keys, the tensor has "
                             f"{keys.element_size()}-byte elements ({keys.dtype})")
        if keys.dtype.is_floating_point or keys.dtype.is_complex:
            raise ValueError(f"{what}: keys must be an integer tensor, not {keys.dtype}")
    elif n is None:
        raise ValueError(f"{what}: n is required with a raw device address")
    return _ptr(keys), int(n)


def _counts_arg(counts, n: int, what: str, dtype_bytes: int = 4, floating: bool = False):
    """Device address of an output buffer of at least n elements of a 32-bit
    integer type (or float32 when floating)."""
    if hasattr(counts, "numel"):
        if counts.element_size() != dtype_bytes or counts.dtype.is_floating_point != floating:
            want = "float32" if floating else "32-bit integer"
            raise ValueError(f"{what}: the output must be a {want} tensor, not {counts.dtype}")
        if counts.numel() < n:
            raise ValueError(f"{what}: the output holds {counts.numel()} elements, {n} needed")
    return _ptr(counts)


def _offs_keys(o, keys, key_bytes: int, what: str):
    """Check a batch's offsets against its key tensor; returns the key address."""
    ptr, _ = _keys_arg(keys, int(o[-1]), key_bytes, what)
    return ptr


def default_params(expected_keys: int, wlf: float = 0.5) -> qs_params:
    """Reference sizing rule (device_sketch.h:371-376, lib/qsketch.cu:50-115)."""
    p = qs_params()
    check(_lib().qs_default_params(C.byref(p), int(expected_keys), float(wlf)), "qs_default_params")
    return p


def find_greatest_prime(x: int) -> int:
    return int(_lib().qs_find_greatest_prime(int(x)))


def largest_prime_at_most(x: int) -> int:
    """Largest prime <= x (deterministic Miller-Rabin for 64-bit x).  The bucket
    count must stay prime: with the multiplicative hash k * s mod 2^w, an even
    modulus lets low key bits steer the bucket choice."""
    def is_prime(n: int) -> bool:
        if n < 2:
            return False
        for p in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
            if n % p == 0:
                return n == p
        d, r = n - 1, 0
        while d % 2 == 0:
            d //= 2
            r += 1
        for a in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
            x_ = pow(a, d, n)
            if x_ in (1, n - 1):
                continue
            for _ in range(r - 1):
                x_ = x_ * x_ % n
                if x_ == n - 1:
                    break
            else:
                return False
        return True
    x = int(x)
    while x >= 2 and not is_prime(x):
        x -= 1
    return x if x >= 2 else 0


def number_of_buckets(insert_keys_sz: int, m: int = 8, factor: float = 0.5) -> int:
    """Sketch_GPU::number_of_buckets (device_sketch.h:371-376)."""
    if m != 8:
        row = int(((insert_keys_sz / factor) * HASH_MASK_ONES) / m)
        return find_greatest_prime(row)
    return int(default_params(insert_keys_sz, factor).n_low)


class Sketch_GPU_Sub_Warp_Level:
    """Two-level MSWS on the current HIP device.

    Constructor arguments follow v3:544-547:
    ``(n_low, m_low=8, n_high=1024, m_high=32, seed_sz=1, n_hash_mask_table=1021)``;
    build additions are keyword-only.
    """

    def __init__(self, n_low: int, m_low: int = 8, n_high: int = 1024, m_high: int = 32,
                 seed_sz: int = SEED_SZ, n_hash_mask_table: int = HASH_MASK_TABLE_SZ, *,
                 key_bytes: int = 4, semantics: int = QS_LITERAL, rng_seed: int = 1,
                 hash_seed: int = 0, chunk_keys: int = 0):
        p = qs_params()
        _lib().qs_default_params(C.byref(p), 1 << 20, 0.5)
        p.n_low, p.m_low, p.n_high, p.m_high = int(n_low), int(m_low), int(n_high), int(m_high)
        p.seed_sz, p.n_masks = int(seed_sz), int(n_hash_mask_table)
        p.key_bytes, p.semantics, p.rng_seed = int(key_bytes), int(semantics), int(rng_seed)
        p.hash_seed, p.chunk_keys = int(hash_seed) & 0xffffffff, int(chunk_keys)
        self.params = p
        self.key_bytes = int(key_bytes)
        h = C.c_void_p()
        check(_lib().qs_create(C.byref(p), C.byref(h)), "qs_create")
        self._h = h

    # -- Sketch<Key_T, Count_T> interface (sketch.h:7-14) ------------------
    def insert(self, keys, keys_sz: int | None = None) -> int:
        kp, n = _keys_arg(keys, keys_sz, self.key_bytes, "qs_insert")
        rc = _lib().qs_insert(self._h, kp, n)
        check(rc, "qs_insert")
        return rc

    def search(self, keys, keys_sz: int | None = None, counts=None):
        kp, n = _keys_arg(keys, keys_sz, self.key_bytes, "qs_query")
        if counts is None:
            import torch
            counts = torch.empty(n, dtype=torch.int32, device=keys.device)
        check(_lib().qs_query(self._h, kp, n, _counts_arg(counts, n, "qs_query")), "qs_query")
        return counts

    query = search

    def clear(self) -> None:
        check(_lib().qs_clear(self._h), "qs_clear")

    def destroy(self) -> None:
        if getattr(self, "_h", None):
            _lib().qs_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    # -- build additions -----------------------------------------------------
    def insert_async(self, keys, n: int | None = None, stream: int = 0) -> None:
        kp, n = _keys_arg(keys, n, self.key_bytes, "qs_insert_async")
        check(_lib().qs_insert_async(self._h, kp, n, C.c_void_p(stream)), "qs_insert_async")

    def search_async(self, keys, n: int | None, counts, stream: int = 0) -> None:
        kp, n = _keys_arg(keys, n, self.key_bytes, "qs_query_async")
        check(_lib().qs_query_async(self._h, kp, n, _counts_arg(counts, n, "qs_query_async"),
                                    C.c_void_p(stream)), "qs_query_async")

    def insert_noalloc(self, keys, n: int | None = None, stream: int = 0) -> None:
        """Stream-ordered insert whose scratch is the library's static device
        array (no allocation): n <= NOALLOC_MAX_KEYS, min(n_high, n) <=
        NOALLOC_MAX_ALLOCS."""
        kp, n = _keys_arg(keys, n, self.key_bytes, "qs_insert_noalloc")
        check(_lib().qs_insert_noalloc(self._h, kp, n, C.c_void_p(stream)), "qs_insert_noalloc")

    def sync(self) -> None:
        check(_lib().qs_sync(self._h), "qs_sync")

    def set_shape(self, width: int, items_per_lane: int) -> None:
        """Segment-pass shape: sub-wavefront width x items per lane (new sketches: (64, 4); (0, 0) = the 64-lane one-item kernel)."""
        check(_lib().qs_set_insert_shape(self._h, int(width), int(items_per_lane)), "qs_set_insert_shape")

    def set_hot_filter(self, enable: bool) -> None:
        """Hot-key filter of the blocking insert (on by default; same result either way)."""
        check(_lib().qs_set_hot_filter(self._h, 1 if enable else 0), "qs_set_hot_filter")

    def shape(self):
        w, i = C.c_uint32(), C.c_uint32()
        check(_lib().qs_get_insert_shape(self._h, C.byref(w), C.byref(i)), "qs_get_insert_shape")
        return w.value, i.value

    def autotune(self, keys, n: int | None = None, reps: int = 2) -> dict:
        """Time all 16 segment-pass shapes on a twin sketch with these keys; keep the fastest."""
        import numpy as np
        kp, n = _keys_arg(keys, n, self.key_bytes, "qs_autotune_insert")
        o = np.array([0, n], dtype=np.uint64)
        r = qs_tune_result()
        rc = _lib().qs_autotune_insert(self._h, kp, o.ctypes.data, None, int(reps), C.byref(r))
        table = [{"width": r.e[i].width, "items_per_lane": r.e[i].items_per_lane,
                  "seg_ms": round(r.e[i].seg_ms, 4), "insert_ms": round(r.e[i].insert_ms, 4)}
                 for i in range(r.n)]
        check(rc, "qs_autotune_insert")
        return {"best": table[r.best], "consistent": bool(r.consistent), "table": table}

    def capture_insert(self, keys, n: int | None = None) -> "Graph":
        """Record the insert of n keys at this address into a hipGraph."""
        import numpy as np
        kp, n = _keys_arg(keys, n, self.key_bytes, "qs_graph_capture_insert")
        o = np.array([0, n], dtype=np.uint64)
        g = C.c_void_p()
        check(_lib().qs_graph_capture_insert(self._h, kp, o.ctypes.data, None, C.byref(g)),
              "qs_graph_capture_insert")
        return Graph(g, self)

    def export_state(self):
        import numpy as np
        p = self.params
        low = np.zeros(p.n_low * 8, dtype=np.uint32)
        high = np.zeros(p.n_high * 32, dtype=np.uint32)
        masks = np.zeros(p.n_masks, dtype=np.uint32)
        seed = C.c_uint32()
        nxt = C.c_uint32()
        check(_lib().qs_export_state(self._h, low.ctypes.data, high.ctypes.data, masks.ctypes.data,
                                     C.addressof(seed), C.addressof(nxt)), "qs_export_state")
        return {"low": low, "high": high, "masks": masks, "seed": seed.value, "next": nxt.value}

    def import_state(self, low, high, masks, seed: int, next_id: int) -> None:
        import numpy as np
        low = np.ascontiguousarray(low, dtype=np.uint32)
        high = np.ascontiguousarray(high, dtype=np.uint32)
        masks = np.ascontiguousarray(masks, dtype=np.uint32)
        check(_lib().qs_import_state(self._h, low.ctypes.data, high.ctypes.data, masks.ctypes.data,
                                     int(seed) & 0xffffffff, int(next_id)), "qs_import_state")

    def stats(self) -> dict:
        st = qs_stats()
        check(_lib().qs_get_stats(self._h, C.byref(st)), "qs_get_stats")
        return {f: getattr(st, f) for f, _ in qs_stats._fields_}

    def set_timing(self, enable: bool = True) -> None:
        """Record HIP events around every insert/query kernel launch (resets totals)."""
        check(_lib().qs_set_timing(self._h, 1 if enable else 0), "qs_set_timing")

    def timing(self) -> dict:
        t = qs_timing()
        check(_lib().qs_get_timing(self._h, C.byref(t)), "qs_get_timing")
        d = {f: getattr(t, f) for f, _ in qs_timing._fields_[:6]}
        d["kernels"] = {k: {"ms": t.kernel_ms[i], "launches": t.kernel_launches[i]}
                        for i, k in enumerate(TIMING_KERNELS)}
        return d

    @property
    def handle(self) -> int:
        return self._h.value

    def memory_bytes(self) -> int:
        """Low + high table bytes (show_memory_usage, device_sketch.h:403-407)."""
        return int(self.params.n_low) * 32 + int(self.params.n_high) * 128


# convenience alias
QSketch = Sketch_GPU_Sub_Warp_Level


def make_params(n_low: int, n_high: int = 1024, *, n_masks: int = HASH_MASK_TABLE_SZ, seed_sz: int = SEED_SZ,
                key_bytes: int = 4, semantics: int = QS_LITERAL, rng_seed: int = 1, hash_seed: int = 0,
                chunk_keys: int = 0) -> qs_params:
    """A parameter set with the reference's fixed fields (m_low 8, m_high 32, ...)."""
    p = qs_params()
    _lib().qs_default_params(C.byref(p), 1 << 20, 0.5)
    p.n_low, p.n_high, p.n_masks, p.seed_sz = int(n_low), int(n_high), int(n_masks), int(seed_sz)
    p.key_bytes, p.semantics, p.rng_seed = int(key_bytes), int(semantics), int(rng_seed)
    p.hash_seed, p.chunk_keys = int(hash_seed) & 0xffffffff, int(chunk_keys)
    return p


def _offs_array(offs):
    import numpy as np
    o = np.ascontiguousarray(offs, dtype=np.uint64)
    if o.ndim != 1 or o.size < 2 or o[0] != 0 or np.any(np.diff(o.astype(np.int64)) < 0):
        raise ValueError("offsets must start at 0 and be non-decreasing")
    return o


class SketchSet:
    """A set of independent MSWS sketches updated together (include/qsketch.h,
    qs_multi_*): sketch i owns keys[offs[i]:offs[i+1]] of every batch and ends in
    the state the plain sketch with its parameters reaches on its own stream."""

    def __init__(self, params):
        arr = (qs_params * len(params))(*params)
        self.params = list(arr)
        self.n = len(params)
        self.key_bytes = int(params[0].key_bytes)
        h = C.c_void_p()
        check(_lib().qs_multi_create(arr, self.n, C.byref(h)), "qs_multi_create")
        self._h = h

    def _d_offs(self, d_offs, what):
        if hasattr(d_offs, "numel") and (d_offs.numel() != self.n + 1 or d_offs.element_size() != 8):
            raise ValueError(f"{what}: device offsets must be {self.n + 1} 64-bit values")
        return _ptr(d_offs)

    def _h_offs(self, offs, what):
        o = _offs_array(offs)
        if o.size != self.n + 1:
            raise ValueError(f"{what}: {self.n + 1} offsets needed, {o.size} given")
        return o

    def insert(self, keys, offs) -> None:
        o = self._h_offs(offs, "qs_multi_insert")
        kp = _offs_keys(o, keys, self.key_bytes, "qs_multi_insert")
        check(_lib().qs_multi_insert(self._h, kp, o.ctypes.data), "qs_multi_insert")

    def insert_async(self, keys, offs, d_offs, stream: int = 0) -> None:
        o = self._h_offs(offs, "qs_multi_insert_async")
        kp = _offs_keys(o, keys, self.key_bytes, "qs_multi_insert_async")
        check(_lib().qs_multi_insert_async(self._h, kp, o.ctypes.data, self._d_offs(d_offs, "qs_multi_insert_async"),
                                           C.c_void_p(stream)), "qs_multi_insert_async")

    def search(self, keys, offs, counts=None):
        import torch
        o = self._h_offs(offs, "qs_multi_query")
        kp = _offs_keys(o, keys, self.key_bytes, "qs_multi_query")
        if counts is None:
            counts = torch.empty(int(o[-1]), dtype=torch.int32, device=keys.device)
        check(_lib().qs_multi_query(self._h, kp, o.ctypes.data, _counts_arg(counts, int(o[-1]), "qs_multi_query")),
              "qs_multi_query")
        return counts

    query = search

    def search_async(self, keys, d_offs, n: int, counts, stream: int = 0) -> None:
        kp, n = _keys_arg(keys, n, self.key_bytes, "qs_multi_query_async")
        check(_lib().qs_multi_query_async(self._h, kp, self._d_offs(d_offs, "qs_multi_query_async"), n,
                                          _counts_arg(counts, n, "qs_multi_query_async"), C.c_void_p(stream)),
              "qs_multi_query_async")

    def capture_insert(self, keys, offs, d_offs) -> "Graph":
        """Record the insert of one batch layout into a hipGraph (see Graph)."""
        o = self._h_offs(offs, "qs_graph_capture_insert")
        kp = _offs_keys(o, keys, self.key_bytes, "qs_graph_capture_insert")
        g = C.c_void_p()
        check(_lib().qs_graph_capture_insert(self._h, kp, o.ctypes.data, self._d_offs(d_offs, "qs_graph_capture_insert"),
                                             C.byref(g)),
              "qs_graph_capture_insert")
        return Graph(g, self)

    def clear(self) -> None:
        check(_lib().qs_clear(self._h), "qs_clear")

    def sync(self) -> None:
        check(_lib().qs_sync(self._h), "qs_sync")

    def set_shape(self, width: int, items_per_lane: int) -> None:
        """Segment-pass shape: sub-wavefront width x items per lane (new sketches: (64, 4); (0, 0) = the 64-lane one-item kernel)."""
        check(_lib().qs_set_insert_shape(self._h, int(width), int(items_per_lane)), "qs_set_insert_shape")

    def shape(self):
        w, i = C.c_uint32(), C.c_uint32()
        check(_lib().qs_get_insert_shape(self._h, C.byref(w), C.byref(i)), "qs_get_insert_shape")
        return w.value, i.value

    def autotune(self, keys, offs, d_offs, reps: int = 2) -> dict:
        """Time all 16 shapes on a twin set with this batch; keep the fastest."""
        o = self._h_offs(offs, "qs_autotune_insert")
        kp = _offs_keys(o, keys, self.key_bytes, "qs_autotune_insert")
        r = qs_tune_result()
        rc = _lib().qs_autotune_insert(self._h, kp, o.ctypes.data, self._d_offs(d_offs, "qs_autotune_insert"),
                                       int(reps), C.byref(r))
        table = [{"width": r.e[i].width, "items_per_lane": r.e[i].items_per_lane,
                  "seg_ms": round(r.e[i].seg_ms, 4), "insert_ms": round(r.e[i].insert_ms, 4),
                  "checksum": r.e[i].checksum} for i in range(r.n)]
        check(rc, "qs_autotune_insert")
        return {"best": table[r.best], "consistent": bool(r.consistent), "table": table}

    def export_state(self, i: int):
        import numpy as np
        p = self.params[i]
        low = np.zeros(p.n_low * 8, dtype=np.uint32)
        high = np.zeros(p.n_high * 32, dtype=np.uint32)
        masks = np.zeros(p.n_masks, dtype=np.uint32)
        seed = C.c_uint32()
        nxt = C.c_uint32()
        check(_lib().qs_multi_export_state(self._h, int(i), low.ctypes.data, high.ctypes.data,
                                           masks.ctypes.data, C.addressof(seed), C.addressof(nxt)),
              "qs_multi_export_state")
        return {"low": low, "high": high, "masks": masks, "seed": seed.value, "next": nxt.value}

    def stats(self) -> dict:
        st = qs_stats()
        check(_lib().qs_get_stats(self._h, C.byref(st)), "qs_get_stats")
        return {f: getattr(st, f) for f, _ in qs_stats._fields_}

    def set_timing(self, enable: bool = True) -> None:
        check(_lib().qs_set_timing(self._h, 1 if enable else 0), "qs_set_timing")

    def timing(self) -> dict:
        t = qs_timing()
        check(_lib().qs_get_timing(self._h, C.byref(t)), "qs_get_timing")
        d = {f: getattr(t, f) for f, _ in qs_timing._fields_[:6]}
        d["kernels"] = {k: {"ms": t.kernel_ms[i], "launches": t.kernel_launches[i]}
                        for i, k in enumerate(TIMING_KERNELS)}
        return d

    def memory_bytes(self) -> int:
        return sum(int(p.n_low) * 32 + int(p.n_high) * 128 for p in self.params)

    @property
    def handle(self) -> int:
        return self._h.value

    def destroy(self) -> None:
        if getattr(self, "_h", None):
            _lib().qs_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class Sketch_GPU_Sub_Warp:
    """Single-level sub-warp sketch (include/qsketch_single.h), mirroring the
    reference's Sketch_GPU_Sub_Warp(n, m, seed_sz, n_hash_mask_table)
    (device_sketch_v2.h:307-309): n buckets of m u32 counters, count-min query."""

    def __init__(self, n: int, m: int = 8, seed_sz: int = SEED_SZ, n_hash_mask_table: int = HASH_MASK_TABLE_SZ,
                 *, key_bytes: int = 4, rng_seed: int = 1):
        h = C.c_void_p()
        check(_lib().qs_sl_create(int(n), int(m), int(seed_sz), int(n_hash_mask_table), int(key_bytes),
                                  int(rng_seed), C.byref(h)), "qs_sl_create")
        self._h = h
        self.n, self.m, self.n_masks, self.key_bytes = int(n), int(m), int(n_hash_mask_table), int(key_bytes)

    def insert(self, keys, keys_sz: int | None = None) -> int:
        kp, n = _keys_arg(keys, keys_sz, self.key_bytes, "qs_sl_insert")
        check(_lib().qs_sl_insert(self._h, kp, n), "qs_sl_insert")
        return 0

    def search(self, keys, keys_sz: int | None = None, counts=None):
        import torch
        kp, n = _keys_arg(keys, keys_sz, self.key_bytes, "qs_sl_query")
        if counts is None:
            counts = torch.empty(n, dtype=torch.int32, device=keys.device)
        check(_lib().qs_sl_query(self._h, kp, n, _counts_arg(counts, n, "qs_sl_query")), "qs_sl_query")
        return counts

    query = search

    def clear(self) -> None:
        check(_lib().qs_sl_clear(self._h), "qs_sl_clear")

    def export_state(self):
        import numpy as np
        table = np.zeros(self.n * self.m, np.uint32)
        masks = np.zeros(self.n_masks, np.uint32)
        seed = C.c_uint32()
        check(_lib().qs_sl_export(self._h, table.ctypes.data, masks.ctypes.data, C.addressof(seed)), "qs_sl_export")
        return {"table": table, "masks": masks, "seed": seed.value}

    def memory_bytes(self) -> int:
        return self.n * self.m * 4

    def destroy(self) -> None:
        if getattr(self, "_h", None):
            _lib().qs_sl_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class Sketch_GPU_SF_Host_Sub_Warp:
    """SF sketch with split counters (include/qsketch_sf.h), mirroring the
    reference's Sketch_GPU_SF_Host_Sub_Warp(n, m, times, seed_sz,
    n_hash_mask_table) (device_sketch_v42.h:245-276): n buckets of m counters,
    each split `times` ways; a counter reads as its largest split, an estimate
    is the smallest of the mask's counters."""

    def __init__(self, n: int, m: int = 8, times: int = 3, seed_sz: int = SEED_SZ,
                 n_hash_mask_table: int = HASH_MASK_TABLE_SZ, *, key_bytes: int = 4, rng_seed: int = 1):
        h = C.c_void_p()
        check(_lib().qs_sf_create(int(n), int(m), int(times), int(seed_sz), int(n_hash_mask_table),
                                  int(key_bytes), int(rng_seed), C.byref(h)), "qs_sf_create")
        self._h = h
        self.n, self.m, self.times = int(n), int(m), int(times)
        self.n_masks, self.key_bytes = int(n_hash_mask_table), int(key_bytes)

    def insert(self, keys, keys_sz: int | None = None) -> int:
        kp, n = _keys_arg(keys, keys_sz, self.key_bytes, "qs_sf_insert")
        check(_lib().qs_sf_insert(self._h, kp, n), "qs_sf_insert")
        return 0

    def search(self, keys, keys_sz: int | None = None, counts=None):
        import torch
        kp, n = _keys_arg(keys, keys_sz, self.key_bytes, "qs_sf_query")
        if counts is None:
            counts = torch.empty(n, dtype=torch.int32, device=keys.device)
        check(_lib().qs_sf_query(self._h, kp, n, _counts_arg(counts, n, "qs_sf_query")), "qs_sf_query")
        return counts

    query = search

    def clear(self) -> None:
        check(_lib().qs_sf_clear(self._h), "qs_sf_clear")

    def export_state(self):
        import numpy as np
        split = np.zeros(self.n * self.m * self.times, np.uint32)
        search = np.zeros(self.n * self.m, np.uint32)
        masks = np.zeros(self.n_masks, np.uint32)
        index = np.zeros(self.n_masks * 3, np.uint8)
        seeds = np.zeros(3, np.uint64)
        check(_lib().qs_sf_export(self._h, split.ctypes.data, search.ctypes.data, masks.ctypes.data,
                                  index.ctypes.data, seeds.ctypes.data), "qs_sf_export")
        return {"split": split, "search": search, "masks": masks, "index": index,
                "seeds": [int(x) for x in seeds]}

    def memory_bytes(self) -> int:
        return self.n * self.m * (self.times + 1) * 4

    def destroy(self) -> None:
        if getattr(self, "_h", None):
            _lib().qs_sf_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class LearnedSketch:
    """Learned heavy-hitter form (include/qsketch_learned.h): an MLP routes the
    keys it scores at or above a threshold to an exact GPU table; the others go,
    in stream order, to a plain sketch with parameters `params`."""

    def __init__(self, params: qs_params, table_log2: int = 17, init_seed: int = 1):
        h = C.c_void_p()
        check(_lib().qs_learned_create(C.byref(params), int(table_log2), int(init_seed), C.byref(h)),
              "qs_learned_create")
        self._h = h
        self.params = params
        self.key_bytes = int(params.key_bytes)
        self.table_log2 = int(table_log2)

    # -- network ---------------------------------------------------------------
    def get_params(self):
        import numpy as np
        a = np.zeros(MLP_PARAMS, np.float32)
        check(_lib().qs_mlp_get_params(self._h, a.ctypes.data), "qs_mlp_get_params")
        return a

    def set_params(self, a) -> None:
        import numpy as np
        a = np.ascontiguousarray(a, np.float32)
        assert a.size == MLP_PARAMS
        check(_lib().qs_mlp_set_params(self._h, a.ctypes.data), "qs_mlp_set_params")

    def _pairs(self, keys, targets, what):
        kp, n = _keys_arg(keys, None, self.key_bytes, what)
        if n < 1 or n > 0xFFFFFFFF:
            raise ValueError(f"{what}: 1 .. 2^32 - 1 pairs per batch, {n} given")
        return kp, _counts_arg(targets, n, what, floating=True), n

    def grad(self, keys, targets):
        """(loss, FP32 gradient) of a batch; no update."""
        import numpy as np
        kp, tp, n = self._pairs(keys, targets, "qs_mlp_grad")
        g = np.zeros(MLP_PARAMS, np.float32)
        loss = C.c_float()
        check(_lib().qs_mlp_grad(self._h, kp, tp, n, C.byref(loss), g.ctypes.data), "qs_mlp_grad")
        return loss.value, g

    def step(self, keys, targets, lr: float = 1e-3, beta1: float = 0.9, beta2: float = 0.999,
             eps: float = 1e-8, want_loss: bool = True):
        kp, tp, n = self._pairs(keys, targets, "qs_mlp_step")
        loss = C.c_float()
        check(_lib().qs_mlp_step(self._h, kp, tp, n, lr, beta1, beta2, eps,
                                 C.byref(loss) if want_loss else None, None), "qs_mlp_step")
        return loss.value if want_loss else None

    def score(self, keys, mfma: bool = True, out=None):
        import torch
        kp, n = _keys_arg(keys, None, self.key_bytes, "qs_mlp_score")
        if out is None:
            out = torch.empty(n, dtype=torch.float32, device=keys.device)
        check(_lib().qs_mlp_score(self._h, kp, n, _counts_arg(out, n, "qs_mlp_score", floating=True),
                                  1 if mfma else 0), "qs_mlp_score")
        return out

    # -- routing ---------------------------------------------------------------
    @property
    def threshold(self) -> float:
        return float(_lib().qs_learned_threshold(self._h))

    @threshold.setter
    def threshold(self, v: float) -> None:
        check(_lib().qs_learned_set_threshold(self._h, float(v)), "qs_learned_set_threshold")

    def insert(self, keys, n: int | None = None) -> int:
        """Insert; returns how many keys went to the exact table."""
        kp, n = _keys_arg(keys, n, self.key_bytes, "qs_learned_insert")
        nh = C.c_uint64()
        check(_lib().qs_learned_insert(self._h, kp, n, C.byref(nh)), "qs_learned_insert")
        return nh.value

    def search(self, keys, n: int | None = None, counts=None):
        import torch
        kp, n = _keys_arg(keys, n, self.key_bytes, "qs_learned_query")
        if counts is None:
            counts = torch.empty(n, dtype=torch.int32, device=keys.device)
        check(_lib().qs_learned_query(self._h, kp, n, _counts_arg(counts, n, "qs_learned_query")),
              "qs_learned_query")
        return counts

    query = search

    def clear(self) -> None:
        check(_lib().qs_learned_clear(self._h), "qs_learned_clear")

    def table(self):
        import numpy as np
        cap = 1 << self.table_log2
        k = np.zeros(cap, np.uint64)
        c = np.zeros(cap, np.uint32)
        used = C.c_uint64()
        check(_lib().qs_learned_table(self._h, k.ctypes.data, c.ctypes.data, cap, C.byref(used)),
              "qs_learned_table")
        return k[:used.value], c[:used.value]

    @property
    def table_full(self) -> bool:
        """True once the table has turned a key away to the sketch (until clear)."""
        r = _lib().qs_learned_table_full(self._h)
        if r < 0:
            check(r, "qs_learned_table_full")
        return bool(r)

    def memory(self) -> dict:
        t, m, s = C.c_uint64(), C.c_uint64(), C.c_uint64()
        check(_lib().qs_learned_memory(self._h, C.byref(t), C.byref(m), C.byref(s)), "qs_learned_memory")
        return {"table": t.value, "mlp": m.value, "sketch": s.value, "total": t.value + m.value + s.value}

    def sketch_stats(self) -> dict:
        st = qs_stats()
        check(_lib().qs_get_stats(_lib().qs_learned_sketch(self._h), C.byref(st)), "qs_get_stats")
        return {f: getattr(st, f) for f, _ in qs_stats._fields_}

    def export_sketch_state(self):
        import numpy as np
        p = self.params
        low = np.zeros(p.n_low * 8, dtype=np.uint32)
        high = np.zeros(p.n_high * 32, dtype=np.uint32)
        masks = np.zeros(p.n_masks, dtype=np.uint32)
        seed, nxt = C.c_uint32(), C.c_uint32()
        check(_lib().qs_export_state(_lib().qs_learned_sketch(self._h), low.ctypes.data, high.ctypes.data,
                                     masks.ctypes.data, C.addressof(seed), C.addressof(nxt)), "qs_export_state")
        return {"low": low, "high": high, "masks": masks, "seed": seed.value, "next": nxt.value}

    def destroy(self) -> None:
        if getattr(self, "_h", None):
            _lib().qs_learned_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class Graph:
    """A captured insert (qs_graph_capture_insert).  Replays read the keys and
    device offsets captured, so refill them in place between launches."""

    def __init__(self, h, owner):
        self._h = h
        self._owner = owner  # keeps the sketch alive while the graph is

    def launch(self, stream: int = 0) -> None:
        check(_lib().qs_graph_launch(self._h, C.c_void_p(stream)), "qs_graph_launch")

    def destroy(self) -> None:
        if getattr(self, "_h", None):
            _lib().qs_graph_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class Zipf:
    """Alias table of Zipf(alpha) over ranks 1..universe (include/qsketch_gen.h)."""

    def __init__(self, alpha: float, universe: int, device: bool = True):
        h = C.c_void_p()
        check(_lib().qs_zipf_create(float(alpha), int(universe), 1 if device else 0, C.byref(h)),
              "qs_zipf_create")
        self._h = h
        self.alpha, self.universe = alpha, universe

    def destroy(self) -> None:
        if getattr(self, "_h", None):
            _lib().qs_zipf_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    def keys_u32(self, seed: int, n: int, out=None, offset: int = 0, device: str = "cuda"):
        import torch
        if out is None:
            out = torch.empty(n, dtype=torch.int32, device=device)
        check(_lib().qs_gen_zipf_u32(self._h, seed, offset, n, _counts_arg(out, n, "qs_gen_zipf_u32"), None),
              "qs_gen_zipf_u32")
        return out

    def keys_u32_host(self, seed: int, n: int, offset: int = 0):
        import numpy as np
        out = np.zeros(n, dtype=np.uint32)
        check(_lib().qs_gen_zipf_u32_host(self._h, seed, offset, n, out.ctypes.data), "zipf host")
        return out

    def ranks_host(self, seed: int, n: int, offset: int = 0):
        import numpy as np
        out = np.zeros(n, dtype=np.uint32)
        check(_lib().qs_gen_zipf_rank_host(self._h, seed, offset, n, out.ctypes.data), "zipf ranks")
        return out

    def keys_u64_noisy(self, seed: int, n: int, sigma: float = 1.5, out=None, offset: int = 0,
                       device: str = "cuda"):
        import torch
        if out is None:
            out = torch.empty(n, dtype=torch.int64, device=device)
        check(_lib().qs_gen_zipf_u64_noisy(self._h, seed, offset, n, float(sigma),
                                           _counts_arg(out, n, "qs_gen_zipf_u64_noisy", dtype_bytes=8), None),
              "qs_gen_zipf_u64_noisy")
        return out

    def keys_u64_noisy_host(self, seed: int, n: int, sigma: float = 1.5, offset: int = 0):
        import numpy as np
        out = np.zeros(n, dtype=np.uint64)
        check(_lib().qs_gen_zipf_u64_noisy_host(self._h, seed, offset, n, float(sigma), out.ctypes.data),
              "zipf u64 host")
        return out


def uniform_u32(seed: int, n: int, out=None, offset: int = 0, device: str = "cuda"):
    import torch
    if out is None:
        out = torch.empty(n, dtype=torch.int32, device=device)
    check(_lib().qs_gen_uniform_u32(seed, offset, n, _counts_arg(out, n, "qs_gen_uniform_u32"), None),
          "qs_gen_uniform_u32")
    return out


def uniform_u32_host(seed: int, n: int, offset: int = 0):
    import numpy as np
    out = np.zeros(n, dtype=np.uint32)
    check(_lib().qs_gen_uniform_u32_host(seed, offset, n, out.ctypes.data), "uniform host")
    return out


def rank_to_key32(seed: int, rank: int) -> int:
    return int(_lib().qs_rank_to_key32(seed, rank))

