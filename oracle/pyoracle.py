"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this module, and only as the checker / CPU baseline.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

LITERAL, PAPER = 0, 1


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


class _Sk(C.Structure):
    _fields_ = [
        ("n_low", C.c_uint64), ("n_high", C.c_uint64), ("n_masks", C.c_uint32),
        ("key_bytes", C.c_uint32), ("semantics", C.c_uint32), ("seed", C.c_uint32),
        ("masks", C.POINTER(C.c_uint32)), ("low", C.POINTER(C.c_uint32)),
        ("high", C.POINTER(C.c_uint32)), ("next", C.c_uint32), ("n_triggers", C.c_uint64),
        ("seed64", C.c_uint64),
    ]


class _Minstd(C.Structure):
    _fields_ = [("x", C.c_uint64)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        u64, u32, vp, sz = C.c_uint64, C.c_uint32, C.c_void_p, C.c_size_t
        L.orq_minstd_seed.argtypes = [C.POINTER(_Minstd), u64]
        L.orq_minstd_next.argtypes = [C.POINTER(_Minstd)]
        L.orq_minstd_next.restype = u64
        L.orq_seed_seq_generate.argtypes = [C.POINTER(u32), sz, C.POINTER(u32), sz]
        L.orq_minstd_seed_bytes.argtypes = [C.POINTER(_Minstd), C.POINTER(C.c_int8), sz]
        L.orq_uniform_u32.argtypes = [C.POINTER(_Minstd), u32, u32]
        L.orq_uniform_u32.restype = u32
        L.orq_seed_buffer.argtypes = [u64, u64, C.POINTER(C.c_int8)]
        L.orq_find_greatest_prime.argtypes = [u64]
        L.orq_find_greatest_prime.restype = u64
        L.orq_number_of_buckets.argtypes = [u64, u32, C.c_double]
        L.orq_number_of_buckets.restype = u64
        L.orq_generate_hashmask.argtypes = [C.POINTER(_Minstd), u32, u32, u32, u32, C.POINTER(C.c_uint8)]
        L.orq_create.argtypes = [C.POINTER(_Sk), u64, u64, u32, u32, u32, u32, u64]
        L.orq_create.restype = C.c_int
        L.orq_create_explicit.argtypes = [C.POINTER(_Sk), u64, u64, u32, u32, C.POINTER(u32), u32, u32]
        L.orq_create_explicit.restype = C.c_int
        L.orq_destroy.argtypes = [C.POINTER(_Sk)]
        L.orq_clear.argtypes = [C.POINTER(_Sk)]
        L.orq_hash.argtypes = [C.POINTER(_Sk), u64]
        L.orq_hash.restype = u64
        L.orq_insert.argtypes = [C.POINTER(_Sk), vp, sz]
        L.orq_insert.restype = C.c_int
        L.orq_insert_mt.argtypes = [C.POINTER(_Sk), vp, sz, C.c_int]
        L.orq_insert_mt.restype = C.c_int
        L.orq_query.argtypes = [C.POINTER(_Sk), vp, sz, C.POINTER(u32)]
        L.orq_query_mt.argtypes = [C.POINTER(_Sk), vp, sz, C.POINTER(u32), C.c_int]
        L.orq_exact_counts.argtypes = [vp, sz, u32, vp, C.POINTER(u32)]
        L.orq_sl_create.argtypes = [C.POINTER(_Sl), u64, u32, u32, u32, u32, u64]
        L.orq_sl_create.restype = C.c_int
        L.orq_sl_destroy.argtypes = [C.POINTER(_Sl)]
        L.orq_sl_insert.argtypes = [C.POINTER(_Sl), vp, sz]
        L.orq_sl_query.argtypes = [C.POINTER(_Sl), vp, sz, C.POINTER(u32)]
        L.orq_exact_counts.restype = sz
        L.orq_sf_create.argtypes = [C.POINTER(_Sf), u64, u32, u32, u32, u32, u32, u64]
        L.orq_sf_create.restype = C.c_int
        L.orq_sf_destroy.argtypes = [C.POINTER(_Sf)]
        L.orq_sf_insert.argtypes = [C.POINTER(_Sf), vp, sz]
        L.orq_sf_query.argtypes = [C.POINTER(_Sf), vp, sz, C.POINTER(u32)]
        L.orq_generate_hashmask_idx.argtypes = [C.POINTER(_Minstd), u32, u32, u32, u32, C.POINTER(C.c_uint8),
                                                C.POINTER(u32)]
        _lib = L
    return _lib


def _p32(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


class Minstd:
    """std::minstd_rand0 restatement."""

    def __init__(self, seed=None, seed_bytes=None):
        self.e = _Minstd()
        if seed_bytes is not None:
            b = np.asarray(seed_bytes, dtype=np.int8)
            lib().orq_minstd_seed_bytes(C.byref(self.e), b.ctypes.data_as(C.POINTER(C.c_int8)), b.size)
        else:
            lib().orq_minstd_seed(C.byref(self.e), 1 if seed is None else seed)

    def next(self) -> int:
        return lib().orq_minstd_next(C.byref(self.e))

    def uniform(self, a: int, b: int) -> int:
        return lib().orq_uniform_u32(C.byref(self.e), a, b)

    def hashmask(self, n, m_bits=32, ones=3, padding=4) -> np.ndarray:
        sz = (m_bits + 7) // 8
        out = np.zeros(n * sz, dtype=np.uint8)
        lib().orq_generate_hashmask(C.byref(self.e), n, m_bits, ones, padding,
                                    out.ctypes.data_as(C.POINTER(C.c_uint8)))
        return out


def seed_seq_generate(v, n) -> np.ndarray:
    v = np.asarray(v, dtype=np.uint32)
    out = np.zeros(n, dtype=np.uint32)
    lib().orq_seed_seq_generate(_p32(v), v.size, _p32(out), n)
    return out


def seed_buffer(rng_seed: int, stream: int) -> np.ndarray:
    out = np.zeros(128, dtype=np.int8)
    lib().orq_seed_buffer(rng_seed, stream, out.ctypes.data_as(C.POINTER(C.c_int8)))
    return out


def find_greatest_prime(x: int) -> int:
    return lib().orq_find_greatest_prime(x)


def number_of_buckets(n: int, m: int = 8, factor: float = 0.5) -> int:
    return lib().orq_number_of_buckets(n, m, factor)


class OracleSketch:
    """Sequential stream-order MSWS (the restated reference semantics)."""

    def __init__(self, n_low, n_high=1024, n_masks=1021, seed_sz=1, key_bytes=4,
                 semantics=LITERAL, rng_seed=1, seed=None, masks=None):
        self.s = _Sk()
        self.key_bytes = key_bytes
        if masks is not None:
            m = np.ascontiguousarray(masks, dtype=np.uint32)
            rc = lib().orq_create_explicit(C.byref(self.s), n_low, n_high, m.size, int(seed) & 0xffffffff,
                                           _p32(m), key_bytes, semantics)
        else:
            rc = lib().orq_create(C.byref(self.s), n_low, n_high, n_masks, seed_sz, key_bytes,
                                  semantics, rng_seed)
        if rc:
            raise RuntimeError(f"orq_create failed: {rc}")

    def __del__(self):
        if getattr(self, "s", None) is not None and self.s.low:
            lib().orq_destroy(C.byref(self.s))

    def _keys(self, keys):
        dt = np.uint32 if self.key_bytes == 4 else np.uint64
        return np.ascontiguousarray(keys, dtype=dt)

    def insert(self, keys, nthreads: int = 1) -> int:
        k = self._keys(keys)
        if nthreads > 1:
            return lib().orq_insert_mt(C.byref(self.s), k.ctypes.data, k.size, nthreads)
        return lib().orq_insert(C.byref(self.s), k.ctypes.data, k.size)

    def query(self, keys, nthreads: int = 1) -> np.ndarray:
        k = self._keys(keys)
        out = np.zeros(k.size, dtype=np.uint32)
        if nthreads > 1:
            lib().orq_query_mt(C.byref(self.s), k.ctypes.data, k.size, _p32(out), nthreads)
        else:
            lib().orq_query(C.byref(self.s), k.ctypes.data, k.size, _p32(out))
        return out

    def clear(self):
        lib().orq_clear(C.byref(self.s))

    def hash(self, key: int) -> int:
        return lib().orq_hash(C.byref(self.s), key)

    @property
    def seed(self) -> int:
        return self.s.seed

    @property
    def next(self) -> int:
        return self.s.next

    def masks(self) -> np.ndarray:
        return np.ctypeslib.as_array(self.s.masks, shape=(self.s.n_masks,)).copy()

    def low(self, copy: bool = True) -> np.ndarray:
        a = np.ctypeslib.as_array(self.s.low, shape=(self.s.n_low * 8,))
        return a.copy() if copy else a

    def high(self, copy: bool = True) -> np.ndarray:
        a = np.ctypeslib.as_array(self.s.high, shape=(self.s.n_high * 32,))
        return a.copy() if copy else a


class _Sl(C.Structure):
    _fields_ = [("n", C.c_uint64), ("m", C.c_uint32), ("n_masks", C.c_uint32), ("key_bytes", C.c_uint32),
                ("seed", C.c_uint32), ("masks", C.POINTER(C.c_uint32)), ("table", C.POINTER(C.c_uint32)),
                ("seed64", C.c_uint64)]


class OracleSingleLevel:
    """Single-level sub-warp sketch (device_sketch_v2.h): one bucket of m u32
    counters per key, 3 selected counters, count-min query."""

    def __init__(self, n, m=8, n_masks=1021, seed_sz=1, key_bytes=4, rng_seed=1):
        self.s = _Sl()
        self.key_bytes = key_bytes
        rc = lib().orq_sl_create(C.byref(self.s), int(n), int(m), int(n_masks), int(seed_sz), int(key_bytes),
                                 int(rng_seed))
        if rc:
            raise RuntimeError(f"orq_sl_create failed: {rc}")

    def __del__(self):
        if getattr(self, "s", None) is not None and self.s.table:
            lib().orq_sl_destroy(C.byref(self.s))

    def _keys(self, keys):
        return np.ascontiguousarray(keys, dtype=np.uint32 if self.key_bytes == 4 else np.uint64)

    def insert(self, keys) -> None:
        k = self._keys(keys)
        lib().orq_sl_insert(C.byref(self.s), k.ctypes.data, k.size)

    def query(self, keys) -> np.ndarray:
        k = self._keys(keys)
        out = np.zeros(k.size, dtype=np.uint32)
        lib().orq_sl_query(C.byref(self.s), k.ctypes.data, k.size, _p32(out))
        return out

    @property
    def seed(self) -> int:
        return self.s.seed

    def masks(self) -> np.ndarray:
        return np.ctypeslib.as_array(self.s.masks, shape=(self.s.n_masks,)).copy()

    def table(self) -> np.ndarray:
        return np.ctypeslib.as_array(self.s.table, shape=(self.s.n * self.s.m,)).copy()


class _Sf(C.Structure):
    _fields_ = [("n", C.c_uint64), ("m", C.c_uint32), ("times", C.c_uint32), ("n_masks", C.c_uint32),
                ("key_bytes", C.c_uint32), ("seed", C.c_uint64 * 3), ("masks", C.POINTER(C.c_uint32)),
                ("index", C.POINTER(C.c_uint8)), ("split", C.POINTER(C.c_uint32)),
                ("search", C.POINTER(C.c_uint32))]


class OracleSF:
    """SF sketch (device_sketch_v42.h:207-480): n buckets of m counters, each
    split into `times` sub-counters; the counter a query reads is the largest
    of its splits, and the estimate the smallest of the mask's counters."""

    def __init__(self, n, m=8, times=3, n_masks=1021, seed_sz=1, key_bytes=4, rng_seed=1):
        self.s = _Sf()
        self.key_bytes = key_bytes
        rc = lib().orq_sf_create(C.byref(self.s), int(n), int(m), int(times), int(n_masks), int(seed_sz),
                                 int(key_bytes), int(rng_seed))
        if rc:
            raise RuntimeError(f"orq_sf_create failed: {rc}")

    def __del__(self):
        if getattr(self, "s", None) is not None and self.s.split:
            lib().orq_sf_destroy(C.byref(self.s))

    def _keys(self, keys):
        return np.ascontiguousarray(keys, dtype=np.uint32 if self.key_bytes == 4 else np.uint64)

    def insert(self, keys) -> None:
        k = self._keys(keys)
        lib().orq_sf_insert(C.byref(self.s), k.ctypes.data, k.size)

    def query(self, keys) -> np.ndarray:
        k = self._keys(keys)
        out = np.zeros(k.size, dtype=np.uint32)
        lib().orq_sf_query(C.byref(self.s), k.ctypes.data, k.size, _p32(out))
        return out

    def seeds(self):
        return [int(x) for x in self.s.seed]

    def masks(self) -> np.ndarray:
        return np.ctypeslib.as_array(self.s.masks, shape=(self.s.n_masks,)).copy()

    def index(self) -> np.ndarray:
        return np.ctypeslib.as_array(self.s.index, shape=(self.s.n_masks * 3,)).copy()

    def split(self) -> np.ndarray:
        return np.ctypeslib.as_array(self.s.split, shape=(self.s.n * self.s.m * self.s.times,)).copy()

    def search(self) -> np.ndarray:
        return np.ctypeslib.as_array(self.s.search, shape=(self.s.n * self.s.m,)).copy()


def exact_counts(keys, key_bytes=4):
    """Sort + reduce-by-key: (distinct keys ascending, counts)."""
    dt = np.uint32 if key_bytes == 4 else np.uint64
    k = np.ascontiguousarray(keys, dtype=dt)
    uniq = np.zeros(k.size, dtype=dt)
    cnt = np.zeros(k.size, dtype=np.uint32)
    u = lib().orq_exact_counts(k.ctypes.data, k.size, key_bytes, uniq.ctypes.data, _p32(cnt))
    return uniq[:u].copy(), cnt[:u].copy()
