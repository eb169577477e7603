"""ctypes binding of the CPU oracle (oracle/wilson_oracle.c).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg. The clwilson_amd package never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


class U64x2(C.Structure):
    _fields_ = [("s0", C.c_uint64), ("s1", C.c_uint64)]


class OrcStats(C.Structure):
    _fields_ = [
        ("checksum", C.c_uint64),
        ("totalcount", C.c_uint64),
        ("typeTarget", C.c_uint64 * 3),
        ("maxtarget", C.c_uint64),
        ("lastp", C.c_uint64),
        ("lastw", C.c_int64),
        ("tpcount", C.c_uint32),
        ("tpcnt", C.c_uint32 * 3),
        ("resultcount", C.c_uint32),
        ("pad", C.c_uint32),
    ]


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-B" if force else "-s", "-C", _HERE], check=True,
                       stdout=subprocess.DEVNULL)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        u64, u32, i64, p64 = C.c_uint64, C.c_uint32, C.c_int64, C.POINTER(C.c_uint64)
        L.orc_invert.restype = u64
        L.orc_invert.argtypes = [u64]
        for name in ("orc_m2p_dup",):
            getattr(L, name).restype = U64x2
            getattr(L, name).argtypes = [U64x2, u64]
        L.orc_m2p_one.restype = U64x2
        L.orc_m2p_one.argtypes = [u64]
        L.orc_m2p_add.restype = U64x2
        L.orc_m2p_add.argtypes = [U64x2, U64x2, u64]
        L.orc_m2p_square.restype = U64x2
        L.orc_m2p_square.argtypes = [U64x2, u64, u64]
        L.orc_m2p_mul.restype = U64x2
        L.orc_m2p_mul.argtypes = [U64x2, U64x2, u64, u64]
        L.orc_m2p_mul_r2.restype = U64x2
        L.orc_m2p_mul_r2.argtypes = [u64, U64x2, u64, u64]
        L.orc_m2p_get.restype = U64x2
        L.orc_m2p_get.argtypes = [U64x2, u64, u64]
        L.orc_setup_constants.restype = None
        L.orc_setup_constants.argtypes = [u64, p64]
        L.orc_strong_prp_two.restype = C.c_int
        L.orc_strong_prp_two.argtypes = [u64]
        L.orc_primes_range.restype = u64
        L.orc_primes_range.argtypes = [u64, u64, p64, u64]
        L.orc_legendre.restype = None
        L.orc_legendre.argtypes = [u64, u64, p64]
        L.orc_isqrt_bitwise.restype = u64
        L.orc_isqrt_bitwise.argtypes = [u64]
        L.orc_find_a.restype = i64
        L.orc_find_a.argtypes = [u64]
        L.orc_find_c.restype = i64
        L.orc_find_c.argtypes = [u64]
        L.orc_find_u.restype = i64
        L.orc_find_u.argtypes = [u64]
        L.orc_find_u_nth.restype = i64
        L.orc_find_u_nth.argtypes = [u64, C.c_int]
        L.orc_finalize.restype = None
        L.orc_finalize.argtypes = [u64, u32, u64, u64, p64]
        L.orc_finalize_ex.restype = None
        L.orc_finalize_ex.argtypes = [u64, u32, u64, u64, i64, i64, p64]
        L.orc_factorial_mod_p2_direct.restype = None
        L.orc_factorial_mod_p2_direct.argtypes = [u64, u64, p64]
        L.orc_direct_quotients.restype = None
        L.orc_direct_quotients.argtypes = [p64, u64, p64, p64, C.c_int]
        L.orc_window_product.restype = None
        L.orc_window_product.argtypes = [p64, p64, u32, u64, u64, C.c_int, p64, p64]
        L.orc_iterate.restype = None
        L.orc_iterate.argtypes = [p64, p64, p64, u32, C.c_int, p64, p64]
        L.orc_wilson_range.restype = C.c_int
        L.orc_wilson_range.argtypes = [u64, u64, C.c_int, p64, C.POINTER(C.c_uint32), p64, p64,
                                       C.POINTER(OrcStats)]
        _lib = L
    return _lib


def _p64(a: np.ndarray):
    assert a.dtype == np.uint64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(C.c_uint64))


def nthreads_default() -> int:
    return max(1, min(8, os.cpu_count() or 1))


def primes_range(lo: int, hi: int) -> np.ndarray:
    L = lib()
    n = L.orc_primes_range(lo, hi, None, 0)
    out = np.zeros(max(n, 1), dtype=np.uint64)
    L.orc_primes_range(lo, hi, _p64(out), n)
    return out[:n]


def legendre(q: int, T: int) -> tuple[int, int]:
    o = (C.c_uint64 * 2)()
    lib().orc_legendre(q, T, o)
    return int(o[0]), int(o[1])


def finalize(p: int, typ: int, value: int, acu1: int = 0, acu2: int = 0) -> tuple[int, int]:
    o = (C.c_uint64 * 2)()
    lib().orc_finalize_ex(p, typ, value & (2**64 - 1), value >> 64, acu1, acu2, o)
    return int(o[0]), int(o[1])


def window_product(tp: np.ndarray, targets: np.ndarray, qlo: int, qhi: int,
                   nthreads: int | None = None) -> list[int]:
    """Canonical prod_{q prime in [qlo,qhi), q<=T_i} q^{e_q(T_i)} mod p_i^2."""
    tp = np.ascontiguousarray(tp, dtype=np.uint64)
    targets = np.ascontiguousarray(targets, dtype=np.uint64)
    lo = np.zeros(len(tp), dtype=np.uint64)
    hi = np.zeros(len(tp), dtype=np.uint64)
    lib().orc_window_product(_p64(tp), _p64(targets), len(tp), qlo, qhi,
                             nthreads or nthreads_default(), _p64(lo), _p64(hi))
    return [int(a) | (int(b) << 64) for a, b in zip(lo, hi)]


def iterate(tp: np.ndarray, frm: np.ndarray, to: np.ndarray, values: list[int],
            nthreads: int | None = None) -> list[int]:
    tp = np.ascontiguousarray(tp, dtype=np.uint64)
    frm = np.ascontiguousarray(frm, dtype=np.uint64)
    to = np.ascontiguousarray(to, dtype=np.uint64)
    lo = np.array([v & (2**64 - 1) for v in values], dtype=np.uint64)
    hi = np.array([v >> 64 for v in values], dtype=np.uint64)
    lib().orc_iterate(_p64(tp), _p64(frm), _p64(to), len(tp), nthreads or nthreads_default(),
                      _p64(lo), _p64(hi))
    return [int(a) | (int(b) << 64) for a, b in zip(lo, hi)]


def wilson_range(pmin: int, pmax: int, nthreads: int | None = None):
    """Run the plain method on [pmin, pmax). Returns (arrays dict, stats dict)."""
    L = lib()
    n = L.orc_primes_range(pmin, pmax, None, 0)
    p = np.zeros(max(n, 1), dtype=np.uint64)
    ty = np.zeros(max(n, 1), dtype=np.uint32)
    quot = np.zeros(max(n, 1), dtype=np.uint64)
    rem = np.zeros(max(n, 1), dtype=np.uint64)
    st = OrcStats()
    rc = L.orc_wilson_range(pmin, pmax, nthreads or nthreads_default(), _p64(p),
                            ty.ctypes.data_as(C.POINTER(C.c_uint32)), _p64(quot), _p64(rem),
                            C.byref(st))
    stats = dict(rc=rc, checksum=st.checksum, totalcount=st.totalcount,
                 typeTarget=list(st.typeTarget), maxtarget=st.maxtarget, lastp=st.lastp,
                 lastw=st.lastw, tpcount=st.tpcount, tpcnt=list(st.tpcnt),
                 resultcount=st.resultcount)
    return dict(p=p[:n], type=ty[:n], quot=quot[:n], rem=rem[:n]), stats


def direct_quotients(ps: np.ndarray, nthreads: int | None = None) -> tuple[np.ndarray, np.ndarray]:
    """(quot, rem) of (p-1)!+1 = rem + p quot mod p^2 from the definition
    (the product 2 * 3 * ... * (p-1)), for p < 2^31."""
    ps = np.ascontiguousarray(ps, dtype=np.uint64)
    assert len(ps) == 0 or int(ps.max()) < 2**31
    q = np.zeros(max(len(ps), 1), dtype=np.uint64)
    r = np.zeros(max(len(ps), 1), dtype=np.uint64)
    lib().orc_direct_quotients(_p64(ps), len(ps), _p64(q), _p64(r), nthreads or nthreads_default())
    return q[: len(ps)], r[: len(ps)]


def wilson_quotient(p: int, quot: int) -> int:
    """Signed w_p in (-p/2, p/2] as the reference prints it (cl_wilson.cpp:872-898)."""
    neg = p - quot
    return quot if quot <= neg else -neg
