"""ctypes binding of the ntt-hip engine's C ABI (include/pdh_capi.h).

Python mirror of the reference's multiply and BBP entry points:

* ``mul_ntt_hip(a, b)`` -- as ``mul_ntt_cuda(a, b)`` (reference
  include/bignum.h:200-201): little-endian uint32 limbs in, normalized
  product limbs out.
* ``bbp_hex_digits_hip(start, count)`` -- as ``bbp_hex_digits_cuda``
  (reference include/bbp.h:34): uppercase hex string for 1-based positions.
* ``bbp_hex_digit(pos)`` -- as ``bbp_hex_digit`` (reference include/bbp.h:27).

Unlike the reference, nothing falls back to the CPU: a missing library or
device raises ``PdhError``. The library is loaded from this directory and
must have been built (``make -C`` this directory, or
``__graft_entry__.build()``).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpdh.so")

PDH_OK = 0
PDH_E_NODEVICE = -1
PDH_E_HIP = -2
PDH_E_ARG = -3
PDH_E_CAPACITY = -4
PDH_E_TOO_LARGE = -5

EXPORTED = (
    "pdh_last_error", "pdh_device_available", "pdh_ntt_length", "pdh_mul", "pdh_sqr",
    "pdh_mul_dev", "pdh_sqr_dev", "pdh_bbp_range", "pdh_bbp_range_dev", "pdh_bbp_positions",
    "pdh_bbp_positions_dev", "pdh_launch_count", "pdh_timing_enable", "pdh_timing_read", "pdh_probe_mad_rate",
    "pdh_batch_create", "pdh_batch_run", "pdh_batch_capture", "pdh_batch_destroy", "pdh_mul_batch_dev",
    "pdh_tune_get", "pdh_tune_set", "pdh_autotune", "pdh_tune_save", "pdh_tune_load", "pdh_dev_alloc", "pdh_dev_free", "pdh_copy_htod",
    "pdh_copy_dtoh", "pdh_stream_create", "pdh_stream_destroy", "pdh_stream_sync",
    "pdh_bbp_set_exact", "pdh_bbp_set_layout", "pdh_bbp_exact_count", "pdh_tune_engine", "pdh_bs_groups", "pdh_bs_merge_dev",
    "pdh_sadd_batch_dev", "pdh_shr_batch_dev", "pdh_decimal_leaves_dev", "pdh_sadd_dev", "pdh_shift_dev",
    "pdh_dev_free_async", "pdh_stream_query", "pdh_mul_dev_noalloc",
)


class MulDesc(ctypes.Structure):
    """pdh_mul_desc: device pointers of one multiply in a batch."""
    _fields_ = [("a", ctypes.c_void_p), ("na", ctypes.c_uint64), ("b", ctypes.c_void_p),
                ("nb", ctypes.c_uint64), ("out", ctypes.c_void_p), ("out_len", ctypes.c_void_p)]


class BsNode(ctypes.Structure):
    """pdh_bs_node: device magnitudes P, Q, |T| with host lengths and the sign of T."""
    _fields_ = [("P", ctypes.c_void_p), ("nP", ctypes.c_uint64), ("Q", ctypes.c_void_p),
                ("nQ", ctypes.c_uint64), ("T", ctypes.c_void_p), ("nT", ctypes.c_uint64),
                ("sT", ctypes.c_int32)]


class SaddDesc(ctypes.Structure):
    """pdh_sadd_desc: out = sa*a + sb*b on device magnitudes with device lengths."""
    _fields_ = [("a", ctypes.c_void_p), ("a_len", ctypes.c_void_p), ("sa", ctypes.c_int32),
                ("b", ctypes.c_void_p), ("b_len", ctypes.c_void_p), ("sb", ctypes.c_int32),
                ("out", ctypes.c_void_p), ("out_cap", ctypes.c_uint64),
                ("out_len", ctypes.c_void_p), ("out_sign", ctypes.c_void_p)]


class ShrDesc(ctypes.Structure):
    """pdh_shr_desc: out = a >> bits on device magnitudes with device lengths."""
    _fields_ = [("a", ctypes.c_void_p), ("a_len", ctypes.c_void_p), ("out", ctypes.c_void_p),
                ("out_cap", ctypes.c_uint64), ("out_len", ctypes.c_void_p), ("bits", ctypes.c_uint64)]


class PdhError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"pdh error {code}: {msg}")
        self.code = code


_lib = None
_sz = ctypes.c_size_t
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_vp = ctypes.c_void_p


def lib() -> ctypes.CDLL:
    """Load libpdh.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise PdhError(PDH_E_NODEVICE, f"{LIB_PATH} not built")
        # PyTorch-ROCm ships its own libamdhip64.so.7 / libhsa-runtime64.so.1.
        # Loading torch first makes libpdh.so bind to that same runtime (same
        # sonames), so device pointers and streams are shared and the process
        # never holds two HIP runtimes.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        L.pdh_last_error.restype = ctypes.c_char_p
        L.pdh_device_available.restype = ctypes.c_int
        L.pdh_ntt_length.argtypes = [_sz, _sz]
        L.pdh_ntt_length.restype = _sz
        L.pdh_mul.argtypes = [_u32p, _sz, _u32p, _sz, _u32p, _sz, ctypes.POINTER(_sz)]
        L.pdh_sqr.argtypes = [_u32p, _sz, _u32p, _sz, ctypes.POINTER(_sz)]
        L.pdh_mul_dev.argtypes = [_vp, _sz, _vp, _sz, _vp, _vp, _vp]
        L.pdh_sqr_dev.argtypes = [_vp, _sz, _vp, _vp, _vp]
        L.pdh_mul_dev_noalloc.argtypes = [_vp, _sz, _vp, _sz, _vp, _vp, _vp]
        L.pdh_mul_dev_noalloc.restype = ctypes.c_int
        L.pdh_bbp_range.argtypes = [ctypes.c_uint64, ctypes.c_uint32, _u8p]
        L.pdh_bbp_positions.argtypes = [_u64p, _sz, _u8p]
        L.pdh_bbp_positions_dev.argtypes = [_vp, _sz, _vp, _vp]
        L.pdh_bbp_range_dev.argtypes = [ctypes.c_uint64, ctypes.c_uint32, _vp, _vp]
        L.pdh_launch_count.restype = ctypes.c_uint64
        L.pdh_bs_groups.argtypes = [ctypes.c_uint64, ctypes.c_int, _vp, _vp, _vp, _vp, _vp, _vp]
        L.pdh_bs_groups.restype = ctypes.c_int
        L.pdh_bs_merge_dev.argtypes = [ctypes.POINTER(BsNode), _sz, ctypes.POINTER(BsNode), ctypes.c_int, _vp]
        L.pdh_bs_merge_dev.restype = ctypes.c_int
        L.pdh_sadd_batch_dev.argtypes = [ctypes.POINTER(SaddDesc), _sz, _vp]
        L.pdh_shr_batch_dev.argtypes = [ctypes.POINTER(ShrDesc), _sz, _vp]
        L.pdh_decimal_leaves_dev.argtypes = [_vp, ctypes.c_uint64, _vp, ctypes.c_uint64, ctypes.c_uint32, _vp, _vp]
        L.pdh_sadd_dev.argtypes = [_vp, _vp, ctypes.c_int32, _vp, _vp, ctypes.c_int32, _vp, ctypes.c_uint64, _vp, _vp,
                                   _vp]
        L.pdh_shift_dev.argtypes = [_vp, _vp, ctypes.c_int64, _vp, ctypes.c_uint64, _vp, _vp]
        for name in ("pdh_sadd_batch_dev", "pdh_shr_batch_dev", "pdh_decimal_leaves_dev", "pdh_sadd_dev",
                     "pdh_shift_dev"):
            getattr(L, name).restype = ctypes.c_int
        L.pdh_bbp_set_exact.argtypes = [ctypes.c_int]
        L.pdh_bbp_set_exact.restype = ctypes.c_int
        L.pdh_bbp_set_layout.argtypes = [ctypes.c_int]
        L.pdh_bbp_set_layout.restype = ctypes.c_int
        L.pdh_bbp_exact_count.restype = ctypes.c_uint64
        L.pdh_timing_enable.argtypes = [ctypes.c_int]
        L.pdh_timing_enable.restype = ctypes.c_int
        L.pdh_timing_read.argtypes = [ctypes.c_char_p, _sz, ctypes.POINTER(ctypes.c_double), _u64p,
                                      ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        L.pdh_timing_read.restype = ctypes.c_int
        _ip = ctypes.POINTER(ctypes.c_int)
        L.pdh_tune_get.argtypes = [ctypes.c_int, _ip, _ip, _ip, _ip]
        L.pdh_tune_set.argtypes = [ctypes.c_int] * 5
        L.pdh_autotune.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp]
        L.pdh_tune_engine.argtypes = [ctypes.c_int, ctypes.c_int, _ip]
        L.pdh_tune_save.argtypes = [ctypes.c_char_p]
        L.pdh_tune_load.argtypes = [ctypes.c_char_p]
        for name in ("pdh_tune_get", "pdh_tune_set", "pdh_autotune", "pdh_tune_save", "pdh_tune_load", "pdh_tune_engine"):
            getattr(L, name).restype = ctypes.c_int
        L.pdh_dev_alloc.argtypes = [_sz, ctypes.POINTER(_vp)]
        L.pdh_dev_free.argtypes = [_vp]
        L.pdh_dev_free_async.argtypes = [_vp, _vp]
        L.pdh_copy_htod.argtypes = [_vp, _vp, _sz, _vp]
        L.pdh_copy_dtoh.argtypes = [_vp, _vp, _sz, _vp]
        L.pdh_stream_create.argtypes = [ctypes.POINTER(_vp)]
        L.pdh_stream_destroy.argtypes = [_vp]
        L.pdh_stream_sync.argtypes = [_vp]
        L.pdh_stream_query.argtypes = [_vp]
        for name in ("pdh_dev_alloc", "pdh_dev_free", "pdh_dev_free_async", "pdh_copy_htod", "pdh_copy_dtoh",
                     "pdh_stream_create", "pdh_stream_destroy", "pdh_stream_sync", "pdh_stream_query"):
            getattr(L, name).restype = ctypes.c_int
        L.pdh_batch_create.argtypes = [ctypes.POINTER(MulDesc), _sz, ctypes.POINTER(_vp)]
        L.pdh_batch_run.argtypes = [_vp, _vp]
        L.pdh_batch_capture.argtypes = [_vp, _vp]
        L.pdh_batch_destroy.argtypes = [_vp]
        L.pdh_mul_batch_dev.argtypes = [ctypes.POINTER(MulDesc), _sz, _vp]
        for name in ("pdh_mul", "pdh_sqr", "pdh_mul_dev", "pdh_sqr_dev", "pdh_bbp_range", "pdh_bbp_range_dev",
                     "pdh_bbp_positions", "pdh_bbp_positions_dev", "pdh_batch_create", "pdh_batch_run",
                     "pdh_batch_capture", "pdh_batch_destroy", "pdh_mul_batch_dev"):
            getattr(L, name).restype = ctypes.c_int
        _lib = L
    return _lib


def _check(rc: int) -> None:
    if rc != PDH_OK:
        raise PdhError(rc, lib().pdh_last_error().decode(errors="replace"))


def device_available() -> bool:
    return bool(lib().pdh_device_available())


def launch_count() -> int:
    return int(lib().pdh_launch_count())


def timing_enable(on: bool) -> None:
    _check(lib().pdh_timing_enable(1 if on else 0))


def timing_read() -> dict:
    """{kernel name: (total ms, launches)} recorded since timing_enable(True)."""
    names = ctypes.create_string_buffer(4096)
    ms = (ctypes.c_double * 64)()
    cnt = (ctypes.c_uint64 * 64)()
    k = ctypes.c_int(0)
    _check(lib().pdh_timing_read(names, 4096, ms, cnt, 64, ctypes.byref(k)))
    keys = names.value.decode().split("\n")[: k.value]
    return {name: (ms[i], int(cnt[i])) for i, name in enumerate(keys)}


def tune_get(n: int) -> tuple:
    """(L1, T, lr_col, lr_row) of the shape used for a 2^n-point transform."""
    v = [ctypes.c_int() for _ in range(4)]
    _check(lib().pdh_tune_get(n, *[ctypes.byref(x) for x in v]))
    return tuple(x.value for x in v)


def tune_engine(n: int, engine: int = -1) -> int:
    """Engine for products with N = 2^n > 2^13 (0 Goldilocks, 1 three primes);
    engine = -1 only reads. Returns the engine in effect."""
    cur = ctypes.c_int()
    _check(lib().pdh_tune_engine(n, engine, ctypes.byref(cur)))
    return cur.value


def tune_set(n: int, L1: int, T: int, lr_col: int, lr_row: int) -> None:
    _check(lib().pdh_tune_set(n, L1, T, lr_col, lr_row))


def autotune(n_lo: int, n_hi: int, reps: int = 3, stream: int = 0) -> None:
    _check(lib().pdh_autotune(n_lo, n_hi, reps, stream or None))


def tune_save(path: str) -> None:
    """Write every transform size's shape to `path` (pdh_tune_save)."""
    _check(lib().pdh_tune_save(os.fsencode(path)))


def tune_load(path: str) -> None:
    """Read shapes written by tune_save; a bad file raises and changes nothing."""
    _check(lib().pdh_tune_load(os.fsencode(path)))


def ntt_length(na: int, nb: int) -> int:
    return int(lib().pdh_ntt_length(na, nb))


def _limbs(x) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x, dtype=np.uint32))


def _ptr(arr: np.ndarray, ctype):
    return arr.ctypes.data_as(ctypes.POINTER(ctype))


def mul_ntt_hip(a, b) -> np.ndarray:
    """Normalized product of two little-endian uint32 limb arrays."""
    a, b = _limbs(a), _limbs(b)
    out = np.zeros(max(len(a) + len(b), 1), dtype=np.uint32)
    n = _sz(0)
    _check(lib().pdh_mul(_ptr(a, ctypes.c_uint32), len(a), _ptr(b, ctypes.c_uint32), len(b),
                         _ptr(out, ctypes.c_uint32), len(out), ctypes.byref(n)))
    return out[: n.value].copy()


def sqr_ntt_hip(a) -> np.ndarray:
    a = _limbs(a)
    out = np.zeros(max(2 * len(a), 1), dtype=np.uint32)
    n = _sz(0)
    _check(lib().pdh_sqr(_ptr(a, ctypes.c_uint32), len(a), _ptr(out, ctypes.c_uint32), len(out),
                         ctypes.byref(n)))
    return out[: n.value].copy()


def mul_dev(d_a: int, na: int, d_b: int, nb: int, d_out: int, d_len: int, stream: int = 0) -> None:
    """Device-resident multiply on raw device pointers (e.g. torch data_ptr())."""
    _check(lib().pdh_mul_dev(d_a, na, d_b, nb, d_out, d_len, stream or None))


NOALLOC_MAX_LIMBS = 1 << 20


def mul_dev_noalloc(d_a: int, na: int, d_b: int, nb: int, d_out: int, d_len: int, stream: int = 0) -> None:
    """pdh_mul_dev without any allocation (na + nb <= 2^20; call device_available() first)."""
    _check(lib().pdh_mul_dev_noalloc(d_a, na, d_b, nb, d_out, d_len, stream or None))


def sqr_dev(d_a: int, na: int, d_out: int, d_len: int, stream: int = 0) -> None:
    _check(lib().pdh_sqr_dev(d_a, na, d_out, d_len, stream or None))


class Batch:
    """A planned batch of independent device-resident multiplies.

    ``descs``: iterable of (a_ptr, na, b_ptr, nb, out_ptr, out_len_ptr) device
    pointers. ``run(stream)`` enqueues every multiply; after ``capture`` each
    run replays one hipGraph."""

    def __init__(self, descs):
        arr = (MulDesc * max(len(descs), 1))()
        for i, d in enumerate(descs):
            arr[i] = MulDesc(*d)
        h = _vp()
        _check(lib().pdh_batch_create(arr, len(descs), ctypes.byref(h)))
        self.h = h

    def run(self, stream: int = 0) -> None:
        _check(lib().pdh_batch_run(self.h, stream or None))

    def capture(self, stream: int) -> None:
        _check(lib().pdh_batch_capture(self.h, stream))

    def close(self) -> None:
        if self.h:
            _check(lib().pdh_batch_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def mul_batch_dev(descs, stream: int = 0) -> None:
    arr = (MulDesc * max(len(descs), 1))()
    for i, d in enumerate(descs):
        arr[i] = MulDesc(*d)
    _check(lib().pdh_mul_batch_dev(arr, len(descs), stream or None))


def bbp_values_range(start: int, count: int) -> np.ndarray:
    out = np.zeros(max(count, 1), dtype=np.uint8)
    _check(lib().pdh_bbp_range(start, count, _ptr(out, ctypes.c_uint8)))
    return out[:count]


def bbp_values_positions(positions) -> np.ndarray:
    pos = np.ascontiguousarray(np.asarray(positions, dtype=np.uint64))
    out = np.zeros(max(len(pos), 1), dtype=np.uint8)
    _check(lib().pdh_bbp_positions(_ptr(pos, ctypes.c_uint64), len(pos), _ptr(out, ctypes.c_uint8)))
    return out[: len(pos)]


def bbp_positions_dev(d_pos: int, n: int, d_out: int, stream: int = 0) -> None:
    _check(lib().pdh_bbp_positions_dev(d_pos, n, d_out, stream or None))


def bbp_range_dev(start: int, count: int, d_out: int, stream: int = 0) -> None:
    """Digit values of positions start .. start + count - 1 into device memory
    (digit blocks: one fixed-point evaluation per few positions)."""
    _check(lib().pdh_bbp_range_dev(start, count, d_out, stream or None))


def bs_groups(n_terms: int, group_log: int, d_P: int, d_Q: int, d_T: int, d_len: int, d_sign: int,
              stream: int = 0) -> None:
    """Device binary-splitting groups (pdh_bs_groups); all pointers are device addresses."""
    _check(lib().pdh_bs_groups(n_terms, group_log, d_P, d_Q, d_T, d_len, d_sign, stream or None))


def bs_merge_dev(nodes, outs, skip_p: bool = False, stream: int = 0) -> list:
    """One binary-splitting merge level in device memory (pdh_bs_merge_dev).

    nodes: BsNode inputs; outs: BsNode outputs whose P, Q, T device buffers
    the caller sized as the C header states. Returns the filled outputs."""
    n_in = len(nodes)
    arr_in = (BsNode * max(n_in, 1))(*nodes)
    arr_out = (BsNode * max(len(outs), 1))(*outs)
    _check(lib().pdh_bs_merge_dev(arr_in, n_in, arr_out, 1 if skip_p else 0, stream or None))
    return [arr_out[i] for i in range(len(outs))]


def sadd_batch_dev(descs, stream: int = 0) -> None:
    """Signed sums of device magnitudes (pdh_sadd_batch_dev)."""
    arr = (SaddDesc * max(len(descs), 1))(*descs)
    _check(lib().pdh_sadd_batch_dev(arr, len(descs), stream or None))


def shr_batch_dev(descs, stream: int = 0) -> None:
    """Right shifts of device magnitudes (pdh_shr_batch_dev)."""
    arr = (ShrDesc * max(len(descs), 1))(*descs)
    _check(lib().pdh_shr_batch_dev(arr, len(descs), stream or None))


def decimal_leaves_dev(d_x: int, stride: int, d_len: int, n: int, width: int, d_out: int, stream: int = 0) -> None:
    """Decimal digits of n device values below 10^width (pdh_decimal_leaves_dev)."""
    _check(lib().pdh_decimal_leaves_dev(d_x, stride, d_len, n, width, d_out, stream or None))


def bbp_set_exact(mode: int) -> None:
    """0 fixed point with fallback (default), 1 chunked FP64-order path
    only, 2 fixed point with odd indices sent to the fallback too (testing),
    3 one lane per series in FP64 order (work-queue kernel)."""
    _check(lib().pdh_bbp_set_exact(int(mode)))


def bbp_set_layout(mode: int) -> None:
    """Term kernel of the fixed-point path: 0 chosen per call (default),
    1 terms in lanes, 2 positions in lanes, 3 moduli in lanes (testing; same
    digits)."""
    _check(lib().pdh_bbp_set_layout(int(mode)))


def probe_mad_rate() -> float:
    """Measured v_mad_u64_u32 issue rate of the device (lane ops per second)."""
    v = ctypes.c_double(0.0)
    _check(lib().pdh_probe_mad_rate(ctypes.byref(v)))
    return v.value


def bbp_exact_count() -> int:
    """Positions computed so far by the FP64-order path (forced or fallback)."""
    return int(lib().pdh_bbp_exact_count())


_HEX = np.frombuffer(b"0123456789ABCDEF", dtype=np.uint8)


def to_hex(values: np.ndarray) -> str:
    """Digit values 0..15 -> uppercase hex (reference src/bbp_cpu.cpp:20-22)."""
    return _HEX[np.asarray(values, dtype=np.uint8)].tobytes().decode()


def bbp_hex_digits_hip(start: int, count: int) -> str:
    return to_hex(bbp_values_range(start, count)) if count else ""


def bbp_hex_digit(position: int) -> str:
    return bbp_hex_digits_hip(position, 1)


# ---- libpidigits_hip.so: Chudnovsky and the serving loop (include/pdh_pi.h) ----
PI_LIB_PATH = os.path.join(HERE, "libpidigits_hip.so")
PI_EXPORTED = ("pdh_pi_chudnovsky", "pdh_pi_chudnovsky_cb", "pdh_chudnovsky_terms", "pdh_server_create",
               "pdh_server_submit_mul", "pdh_server_submit_pi", "pdh_server_submit_bbp", "pdh_server_wait",
               "pdh_server_result", "pdh_server_destroy", "pdh_divmod", "pdh_isqrt", "pdh_divmod_dev",
               "pdh_isqrt_dev", "pdh_divmod_cb", "pdh_isqrt_cb", "pdh_newton_set_perturb",
               "pdh_newton_corrections", "pdh_server_stats")
_pilib = None


def pilib() -> ctypes.CDLL:
    global _pilib
    if _pilib is None:
        lib()                                   # torch first, then libpdh.so, then this
        if not os.path.exists(PI_LIB_PATH):
            raise PdhError(PDH_E_NODEVICE, f"{PI_LIB_PATH} not built")
        L = ctypes.CDLL(PI_LIB_PATH)
        L.pdh_pi_chudnovsky.argtypes = [ctypes.c_uint64, ctypes.c_char_p, _sz, ctypes.POINTER(ctypes.c_double)]
        L.pdh_chudnovsky_terms.argtypes = [ctypes.c_uint64]
        L.pdh_chudnovsky_terms.restype = ctypes.c_uint64
        L.pdh_server_create.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.POINTER(_vp)]
        L.pdh_server_submit_mul.argtypes = [_vp, _u32p, _sz, _u32p, _sz, _u64p]
        L.pdh_server_submit_pi.argtypes = [_vp, ctypes.c_uint64, _u64p]
        L.pdh_server_submit_bbp.argtypes = [_vp, ctypes.c_uint64, ctypes.c_uint32, _u64p]
        L.pdh_server_wait.argtypes = [_vp, ctypes.c_uint64, ctypes.POINTER(ctypes.c_double)]
        L.pdh_server_result.argtypes = [_vp, ctypes.c_uint64, _vp, _sz, ctypes.POINTER(_sz)]
        L.pdh_server_destroy.argtypes = [_vp]
        L.pdh_server_stats.argtypes = [_vp, _u64p, ctypes.c_int]
        L.pdh_server_stats.restype = ctypes.c_int
        _szp = ctypes.POINTER(_sz)
        L.pdh_divmod.argtypes = [_u32p, _sz, _u32p, _sz, _u32p, _sz, _szp, _u32p, _sz, _szp]
        L.pdh_isqrt.argtypes = [_u32p, _sz, _u32p, _sz, _szp]
        L.pdh_divmod_cb.argtypes = [_vp, _u32p, _sz, _u32p, _sz, _u32p, _sz, _szp, _u32p, _sz, _szp]
        L.pdh_isqrt_cb.argtypes = [_vp, _u32p, _sz, _u32p, _sz, _szp]
        L.pdh_divmod_dev.argtypes = [_vp, _sz, _vp, _sz, _vp, _sz, _vp, _vp, _sz, _vp, _vp]
        L.pdh_isqrt_dev.argtypes = [_vp, _sz, _vp, _sz, _vp, _vp]
        L.pdh_newton_set_perturb.argtypes = [ctypes.c_int]
        L.pdh_newton_corrections.restype = ctypes.c_uint64
        for name in ("pdh_pi_chudnovsky", "pdh_server_create", "pdh_server_submit_mul", "pdh_server_submit_pi",
                     "pdh_server_submit_bbp", "pdh_server_wait", "pdh_server_result", "pdh_server_destroy",
                     "pdh_divmod", "pdh_isqrt", "pdh_divmod_cb", "pdh_isqrt_cb", "pdh_divmod_dev", "pdh_isqrt_dev",
                     "pdh_newton_set_perturb"):
            getattr(L, name).restype = ctypes.c_int
        _pilib = L
    return _pilib


def int_to_limbs(x: int) -> np.ndarray:
    """Little-endian uint32 limbs of a non-negative Python int (empty for 0)."""
    n = (x.bit_length() + 31) // 32
    return np.frombuffer(x.to_bytes(4 * n, "little"), dtype=np.uint32).copy() if n else np.zeros(0, np.uint32)


def limbs_to_int(a) -> int:
    return int.from_bytes(np.ascontiguousarray(np.asarray(a, dtype=np.uint32)).tobytes(), "little")


def divmod_hip(a, d, mul_fn=None):
    """(floor(a / d), a mod d) of limb arrays, by the Newton division on the
    device (pdh_divmod), or with every multiply done by the C function
    pointer ``mul_fn`` (pdh_divmod_cb)."""
    a, d = _limbs(a), _limbs(d)
    q = np.zeros(max(len(a) + 1, 1), np.uint32)
    r = np.zeros(max(len(d), 1), np.uint32)
    ql, rl = _sz(), _sz()
    args = (_ptr(a, ctypes.c_uint32), len(a), _ptr(d, ctypes.c_uint32), len(d), _ptr(q, ctypes.c_uint32), len(q),
            ctypes.byref(ql), _ptr(r, ctypes.c_uint32), len(r), ctypes.byref(rl))
    rc = pilib().pdh_divmod(*args) if mul_fn is None else pilib().pdh_divmod_cb(mul_fn, *args)
    _check(rc)
    return q[: ql.value].copy(), r[: rl.value].copy()


def isqrt_hip(s, mul_fn=None):
    """floor(sqrt(s)) of a limb array (pdh_isqrt, or pdh_isqrt_cb with ``mul_fn``)."""
    s = _limbs(s)
    x = np.zeros(max(len(s) // 2 + 2, 1), np.uint32)
    xl = _sz()
    args = (_ptr(s, ctypes.c_uint32), len(s), _ptr(x, ctypes.c_uint32), len(x), ctypes.byref(xl))
    rc = pilib().pdh_isqrt(*args) if mul_fn is None else pilib().pdh_isqrt_cb(mul_fn, *args)
    _check(rc)
    return x[: xl.value].copy()


def newton_set_perturb(k: int) -> None:
    """Test switch: move every Newton estimate by k units before its exact correction."""
    _check(pilib().pdh_newton_set_perturb(int(k)))


def newton_corrections() -> int:
    return int(pilib().pdh_newton_corrections())


def pi_chudnovsky(digits: int):
    """("3.1415...", phase ms [split, 10^prec, sqrt, divide, format, terms])."""
    n = digits + 8
    raw = bytearray(n)                          # decoded in place below (no bytes copy of 10^8 digits)
    buf = (ctypes.c_char * n).from_buffer(raw)
    ph = (ctypes.c_double * 6)()
    _check(pilib().pdh_pi_chudnovsky(digits, buf, n, ph))
    del buf
    return str(memoryview(raw)[: raw.index(0)], "ascii"), list(ph)


class Server:
    """Continuous-batching request loop on the device (pdh_server_*)."""

    def __init__(self, hbm_budget_bytes: int = 200 << 30, max_inflight: int = 64):
        h = _vp()
        _check(pilib().pdh_server_create(hbm_budget_bytes, max_inflight, ctypes.byref(h)))
        self.h = h
        self.kind = {}

    def submit_mul(self, a, b) -> int:
        a, b = _limbs(a), _limbs(b)
        t = ctypes.c_uint64()
        _check(pilib().pdh_server_submit_mul(self.h, _ptr(a, ctypes.c_uint32), len(a), _ptr(b, ctypes.c_uint32),
                                             len(b), ctypes.byref(t)))
        self.kind[t.value] = ("mul", len(a) + len(b))
        return t.value

    def submit_pi(self, digits: int) -> int:
        t = ctypes.c_uint64()
        _check(pilib().pdh_server_submit_pi(self.h, digits, ctypes.byref(t)))
        self.kind[t.value] = ("pi", digits + 2)
        return t.value

    def submit_bbp(self, start: int, count: int) -> int:
        t = ctypes.c_uint64()
        _check(pilib().pdh_server_submit_bbp(self.h, start, count, ctypes.byref(t)))
        self.kind[t.value] = ("bbp", count)
        return t.value

    def wait(self, ticket: int) -> float:
        ms = ctypes.c_double()
        _check(pilib().pdh_server_wait(self.h, ticket, ctypes.byref(ms)))
        return ms.value

    def result(self, ticket: int):
        kind, n = self.kind.pop(ticket)
        ln = _sz()
        if kind == "mul":
            out = np.zeros(max(n, 1), dtype=np.uint32)
            _check(pilib().pdh_server_result(self.h, ticket, out.ctypes.data, n, ctypes.byref(ln)))
            return out[: ln.value].copy()
        buf = ctypes.create_string_buffer(max(n, 1))
        _check(pilib().pdh_server_result(self.h, ticket, ctypes.cast(buf, _vp), n, ctypes.byref(ln)))
        raw = buf.raw[: ln.value]
        return raw.decode() if kind == "pi" else np.frombuffer(raw, dtype=np.uint8).copy()

    def stats(self) -> dict:
        v = (ctypes.c_uint64 * 7)()
        _check(pilib().pdh_server_stats(self.h, v, 7))
        keys = ("submitted", "completed", "waited_for_memory", "backfilled", "peak_bytes", "held_bytes", "budget")
        return dict(zip(keys, (int(x) for x in v)))

    def close(self) -> None:
        if self.h:
            _check(pilib().pdh_server_destroy(self.h))
            self.h = None
