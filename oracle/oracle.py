"""TEST INFRASTRUCTURE ONLY -- ctypes access to the checker libraries.

* ``liboracle.so``: the C++ restatement of the reference's CPU path
  (oracle/cpu_ref.cpp, each function citing the reference file:line).
* ``_ref/libref.so``: the reference's own src/ntt_cpu.cpp and src/bbp_cpu.cpp
  compiled from source by oracle/Makefile (present only where it was built).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg
use this module, and only as the checker / CPU baseline; the product library
never loads it.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libref.so")

_sz = ctypes.c_size_t
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_o = None
_r = None


def _ptr(a, t):
    return a.ctypes.data_as(ctypes.POINTER(t))


def olib() -> ctypes.CDLL:
    global _o
    if _o is None:
        L = ctypes.CDLL(ORACLE_PATH)
        u64 = ctypes.c_uint64
        for f in ("orc_gl_add", "orc_gl_sub", "orc_gl_mul", "orc_gl_pow", "orc_gl_reduce128"):
            getattr(L, f).argtypes = [u64, u64]
            getattr(L, f).restype = u64
        for f in ("orc_gl_inv", "orc_gl_root_of_unity"):
            getattr(L, f).argtypes = [u64]
            getattr(L, f).restype = u64
        L.orc_ntt_length.argtypes = [_sz, _sz]
        L.orc_ntt_length.restype = _sz
        L.orc_repack_to16.argtypes = [_u32p, _sz, _u64p, _sz]
        L.orc_carry_propagate.argtypes = [_u64p, _sz, _u32p, _sz, ctypes.POINTER(_sz)]
        L.orc_ntt_transform.argtypes = [_u64p, _sz, ctypes.c_int]
        for f in ("orc_mul_ntt", "orc_mul_schoolbook"):
            getattr(L, f).argtypes = [_u32p, _sz, _u32p, _sz, _u32p, _sz, ctypes.POINTER(_sz)]
            getattr(L, f).restype = ctypes.c_int
        L.orc_mul_ntt_batch.argtypes = [_u32p, _u64p, _u64p, _u32p, _u64p, _u64p, _u32p, _u64p,
                                        _u64p, _sz, ctypes.c_int]
        L.orc_mul_ntt_batch.restype = ctypes.c_int
        L.orc_bbp_digit_value.argtypes = [u64]
        L.orc_bbp_series.argtypes = [ctypes.c_uint32, u64]
        L.orc_bbp_series.restype = ctypes.c_double
        L.orc_bbp_modpow16.argtypes = [u64, u64]
        L.orc_bbp_modpow16.restype = u64
        L.orc_bbp_hex_range.argtypes = [u64, ctypes.c_uint32, ctypes.c_char_p, ctypes.c_int]
        L.orc_bbp_positions.argtypes = [_u64p, _sz, _u8p, ctypes.c_int]
        L.orc_bbp_mulmod_count.argtypes = [u64]
        L.orc_bbp_mulmod_count.restype = u64
        L.orc_bbp_mulmod_count_brute.argtypes = [u64]
        L.orc_bbp_mulmod_count_brute.restype = u64
        _o = L
    return _o


def ref_available() -> bool:
    return os.path.exists(REF_PATH)


def rlib() -> ctypes.CDLL:
    global _r
    if _r is None:
        L = ctypes.CDLL(REF_PATH)
        L.ref_mul_ntt_cpu.argtypes = [_u32p, _sz, _u32p, _sz, _u32p, _sz, ctypes.POINTER(_sz)]
        L.ref_mul_ntt_cpu.restype = ctypes.c_int
        L.ref_mul_ntt_cpu_batch.argtypes = [_u32p, _u64p, _u64p, _u32p, _u64p, _u64p, _u32p,
                                            _u64p, _u64p, _sz, ctypes.c_int]
        L.ref_mul_ntt_cpu_batch.restype = ctypes.c_int
        L.ref_bbp_hex_range.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_char_p, ctypes.c_int]
        L.ref_bbp_positions.argtypes = [_u64p, _sz, ctypes.c_char_p, ctypes.c_int]
        L.ref_bbp_modpow16.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
        L.ref_bbp_modpow16.restype = ctypes.c_uint64
        _r = L
    return _r


def _mul(fn, a, b) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.uint32)
    b = np.ascontiguousarray(b, dtype=np.uint32)
    out = np.zeros(max(len(a) + len(b), 1), dtype=np.uint32)
    n = _sz(0)
    rc = fn(_ptr(a, ctypes.c_uint32), len(a), _ptr(b, ctypes.c_uint32), len(b),
            _ptr(out, ctypes.c_uint32), len(out), ctypes.byref(n))
    assert rc == 0
    return out[: n.value].copy()


def mul_ntt(a, b) -> np.ndarray:
    """ntt-cpu product (restatement of src/ntt_cpu.cpp:111-129)."""
    return _mul(olib().orc_mul_ntt, a, b)


def mul_schoolbook(a, b) -> np.ndarray:
    """schoolbook product (restatement of src/bignum.cpp:125-143)."""
    return _mul(olib().orc_mul_schoolbook, a, b)


def ref_mul_ntt(a, b) -> np.ndarray:
    """The reference's own mul_ntt_cpu, compiled from source."""
    return _mul(rlib().ref_mul_ntt_cpu, a, b)


def _batch(fn, pairs, threads):
    na = np.array([len(a) for a, _ in pairs], dtype=np.uint64)
    nb = np.array([len(b) for _, b in pairs], dtype=np.uint64)
    a_off = np.concatenate([[0], np.cumsum(na)[:-1]]).astype(np.uint64)
    b_off = np.concatenate([[0], np.cumsum(nb)[:-1]]).astype(np.uint64)
    o_off = np.concatenate([[0], np.cumsum(na + nb)[:-1]]).astype(np.uint64)
    A = np.ascontiguousarray(np.concatenate([a for a, _ in pairs]), dtype=np.uint32)
    B = np.ascontiguousarray(np.concatenate([b for _, b in pairs]), dtype=np.uint32)
    out = np.zeros(int((na + nb).sum()), dtype=np.uint32)
    olen = np.zeros(len(pairs), dtype=np.uint64)
    rc = fn(_ptr(A, ctypes.c_uint32), _ptr(a_off, ctypes.c_uint64), _ptr(na, ctypes.c_uint64),
            _ptr(B, ctypes.c_uint32), _ptr(b_off, ctypes.c_uint64), _ptr(nb, ctypes.c_uint64),
            _ptr(out, ctypes.c_uint32), _ptr(o_off, ctypes.c_uint64), _ptr(olen, ctypes.c_uint64),
            len(pairs), threads)
    assert rc == 0
    return [out[int(o): int(o) + int(l)].copy() for o, l in zip(o_off, olen)]


def mul_ntt_batch(pairs, threads: int = 1):
    return _batch(olib().orc_mul_ntt_batch, pairs, threads)


def ref_mul_ntt_batch(pairs, threads: int = 1):
    return _batch(rlib().ref_mul_ntt_cpu_batch, pairs, threads)


def bbp_hex_range(start: int, count: int, threads: int = 1) -> str:
    buf = ctypes.create_string_buffer(count + 1)
    olib().orc_bbp_hex_range(start, count, buf, threads)
    return buf.raw[:count].decode()


def ref_bbp_hex_range(start: int, count: int, threads: int = 1) -> str:
    buf = ctypes.create_string_buffer(count + 1)
    rlib().ref_bbp_hex_range(start, count, buf, threads)
    return buf.raw[:count].decode()


def bbp_values_positions(positions, threads: int = 1) -> np.ndarray:
    pos = np.ascontiguousarray(positions, dtype=np.uint64)
    out = np.zeros(max(len(pos), 1), dtype=np.uint8)
    olib().orc_bbp_positions(_ptr(pos, ctypes.c_uint64), len(pos), _ptr(out, ctypes.c_uint8), threads)
    return out[: len(pos)]


def ref_bbp_hex_positions(positions, threads: int = 1) -> str:
    pos = np.ascontiguousarray(positions, dtype=np.uint64)
    buf = ctypes.create_string_buffer(len(pos) + 1)
    rlib().ref_bbp_positions(_ptr(pos, ctypes.c_uint64), len(pos), buf, threads)
    return buf.raw[: len(pos)].decode()


def bbp_mulmod_count(position: int) -> int:
    """Exact number of bbp_mulmod calls the reference makes for one position."""
    return int(olib().orc_bbp_mulmod_count(position))


def bbp_mulmod_count_brute(position: int) -> int:
    return int(olib().orc_bbp_mulmod_count_brute(position))


def limbs_to_int(x) -> int:
    return int.from_bytes(np.ascontiguousarray(x, dtype=np.uint32).tobytes(), "little")


def int_to_limbs(v: int) -> np.ndarray:
    if v == 0:
        return np.zeros(0, dtype=np.uint32)
    nbytes = (v.bit_length() + 31) // 32 * 4
    return np.frombuffer(v.to_bytes(nbytes, "little"), dtype=np.uint32).copy()
