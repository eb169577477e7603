"""clwilson_amd -- MI355X Wilson-prime engine (Python host binding).

A ctypes binding of libclwilson.so, whose C ABI (include/clwilson.h) replaces
the reference's OpenCL layer (simpleCL.c). The compute runs in hand-written
gfx950 HIP kernels; this module never computes anything itself and raises if
the native library is missing or no GPU is present.

Typical use:
    import clwilson_amd as cw
    with cw.Device(0) as dev:
        res = dev.wilson_range(1_000_000, 2_000_000)
        print(res.checksum_hex, res.results)
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libclwilson.so")
_lib = None

CLW_OK = 0
CLW_ERR_HIP = -1
CLW_ERR_ARG = -2
CLW_ERR_OVERFLOW = -3
CLW_ERR_NOMEM = -4
CLW_ERR_CHECK = -5
CLW_ERR_STATE = -6
CLW_ERR_IO = -7

BLOCK_SPAN = 61440
REC_CAP = 4096
REC_WORDS = 8
MAXP = 0xFFFFFFFFFFFFFFFF // 4


class ClwError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class DeviceInfo(C.Structure):
    _fields_ = [("name", C.c_char * 256), ("arch", C.c_char * 64), ("compute_units", C.c_int32),
                ("clock_khz", C.c_int32), ("global_mem", C.c_uint64), ("max_alloc", C.c_uint64)]


class TpTable(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("p", "type", "tp_target", "n_lo", "n_hi", "one_lo", "one_hi",
                                          "r2_lo", "r2_hi", "ninv", "acc_lo", "acc_hi")] + \
               [("count", C.c_uint32), ("limbs", C.c_uint32)]


class SegBuffers(C.Structure):
    _fields_ = [("rec", C.c_void_p), ("nrec", C.c_void_p), ("nprimes", C.c_void_p),
                ("side", C.c_void_p * 3), ("side_count", C.c_void_p), ("flags", C.c_void_p),
                ("total", C.c_void_p), ("era", C.c_void_p), ("max_blocks", C.c_uint32), ("side_cap", C.c_uint32)]


class KernelStats(C.Structure):
    _fields_ = [("segments", C.c_uint64), ("sieve_launches", C.c_uint64), ("sieve_integers", C.c_uint64),
                ("mul_launches", C.c_uint64), ("mul_record_apps", C.c_uint64), ("mul_prime_apps", C.c_uint64),
                ("side_launches", C.c_uint64), ("reduce_launches", C.c_uint64), ("sieve_ms", C.c_double),
                ("mul_ms", C.c_double), ("side_ms", C.c_double), ("reduce_ms", C.c_double),
                ("mul_partial_products", C.c_uint64), ("mfma_int8_macs", C.c_uint64),
                ("mfma_ms", C.c_double), ("mfma_int8_issued", C.c_uint64),
                ("mfma_partial_products", C.c_uint64)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class BandTune(C.Structure):
    _fields_ = [("limbs", C.c_uint32), ("per_lane", C.c_uint32), ("items_per_block", C.c_uint32),
                ("trials", C.c_uint32), ("from_cache", C.c_uint32), ("active_test_primes", C.c_uint64),
                ("best_apps_per_s", C.c_double), ("trial_per_lane", C.c_uint32 * 16),
                ("trial_items_per_block", C.c_uint32 * 16), ("trial_apps_per_s", C.c_double * 16),
                ("mfma_slices", C.c_uint32), ("mfma_min_rows", C.c_uint32), ("mfma_trials", C.c_uint32),
                ("mfma_trial_slices", C.c_uint32 * 8), ("mfma_trial_rows", C.c_uint64 * 8),
                ("mfma_trial_ms", C.c_double * 8), ("trial_ms", C.c_double * 16)]

    def as_dict(self) -> dict:
        n = min(self.trials, 16)
        m = min(self.mfma_trials, 8)
        return dict(limbs=self.limbs, per_lane=self.per_lane, items_per_block=self.items_per_block,
                    trials=self.trials, from_cache=bool(self.from_cache),
                    active_test_primes=self.active_test_primes, best_apps_per_s=self.best_apps_per_s,
                    candidates=[(self.trial_per_lane[k], self.trial_items_per_block[k], self.trial_apps_per_s[k])
                                for k in range(n)],
                    candidate_ms=[self.trial_ms[k] for k in range(n)],
                    mfma_slices=self.mfma_slices, mfma_min_rows=self.mfma_min_rows,
                    mfma_candidates=[(self.mfma_trial_slices[k], self.mfma_trial_rows[k], self.mfma_trial_ms[k])
                                     for k in range(m)])


class WorkStatus(C.Structure):
    _fields_ = [("pmin", C.c_uint64), ("pmax", C.c_uint64), ("currp", C.c_uint64), ("trickle", C.c_uint64),
                ("state_sum", C.c_uint64), ("totalcount", C.c_uint64), ("tpcount", C.c_uint32),
                ("done", C.c_uint32)]


class Options(C.Structure):
    _fields_ = [("segment_span", C.c_uint64), ("target_waves", C.c_uint32), ("verbose", C.c_uint32),
                ("mode", C.c_uint32), ("per_lane", C.c_uint32), ("graph", C.c_uint32),
                ("segment_apps", C.c_uint64), ("host_primes", C.c_uint32), ("max_chunks", C.c_uint32),
                ("external_scratch", C.c_uint32), ("run_products", C.c_uint32), ("mfma_slices", C.c_uint32),
                ("external_workspace", C.c_uint32)]


class RangeStats(C.Structure):
    _fields_ = [("pmin", C.c_uint64), ("pmax", C.c_uint64), ("checksum", C.c_uint64),
                ("totalcount", C.c_uint64), ("typeTarget", C.c_uint64 * 3), ("maxtarget", C.c_uint64),
                ("lastp", C.c_uint64), ("lastw", C.c_int64), ("tpcount", C.c_uint32),
                ("tpcnt", C.c_uint32 * 3), ("resultcount", C.c_uint32), ("limbs", C.c_uint32),
                ("seconds", C.c_double)]


def lib():
    """Loads libclwilson.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ClwError(CLW_ERR_STATE, f"{LIB_PATH} is missing: run python -m clwilson_amd.build")
    L = C.CDLL(LIB_PATH)
    vp, u32, u64, i32 = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int
    P = C.POINTER
    sig = {
        "clw_init": (i32, [i32, P(vp)]),
        "clw_shutdown": (i32, [vp]),
        "clw_get_device_info": (i32, [vp, P(DeviceInfo)]),
        "clw_last_error": (C.c_char_p, []),
        "clw_alloc": (i32, [vp, C.c_size_t, P(vp)]),
        "clw_free": (i32, [vp, vp]),
        "clw_h2d": (i32, [vp, vp, vp, C.c_size_t, i32]),
        "clw_d2h": (i32, [vp, vp, vp, C.c_size_t, i32]),
        "clw_memset": (i32, [vp, vp, i32, C.c_size_t]),
        "clw_finish": (i32, [vp]),
        "clw_event_create": (i32, [vp, P(vp)]),
        "clw_event_record": (i32, [vp, vp]),
        "clw_event_query": (i32, [vp]),
        "clw_event_wait": (i32, [vp]),
        "clw_event_elapsed_ms": (i32, [vp, vp, P(C.c_float)]),
        "clw_event_destroy": (i32, [vp]),
        "clw_limbs_for": (i32, [u64]),
        "clw_tp_table_alloc": (i32, [vp, u32, u32, P(TpTable)]),
        "clw_tp_table_free": (i32, [vp, P(TpTable)]),
        "clw_setup": (i32, [vp, P(TpTable), i32]),
        "clw_seg_alloc": (i32, [vp, u32, u32, P(SegBuffers)]),
        "clw_seg_free": (i32, [vp, P(SegBuffers)]),
        "clw_seg_clear": (i32, [vp, P(SegBuffers)]),
        "clw_sieve_segment": (i32, [vp, P(SegBuffers), u64, u64, P(u64), P(u32)]),
        "clw_blocks_upto": (u32, [u64, u64, u64]),
        "clw_mul_records": (i32, [vp, P(TpTable), P(SegBuffers), u32, u32, u32, u32, u32, vp, vp, u32, u32]),
        "clw_default_per_lane": (u32, [u32]),
        "clw_batch_default_per_lane": (u32, [u32]),
        "clw_mul_side": (i32, [vp, P(TpTable), P(SegBuffers), u32, u32, u32, u32, vp, vp, u32]),
        "clw_reduce": (i32, [vp, P(TpTable), u32, u32, u32, vp, vp, u32]),
        "clw_iterate": (i32, [vp, P(TpTable), P(u64)]),
        "clw_finalize": (i32, [vp, P(TpTable), vp, vp, vp]),
        "clw_wilson_range": (i32, [vp, u64, u64, P(Options), P(RangeStats), vp, vp, vp, u64]),
        "clw_format_result": (i32, [u64, u64, C.c_char_p, C.c_size_t]),
        "clw_verify_results": (i32, [C.c_char_p, u64, u64, u64, vp, vp, u64, P(u32), P(u32)]),
        "clw_range_types": (i32, [u64, u64, P(u32), P(u64)]),
        "clw_host_primes": (u64, [u64, u64, vp, u64]),
        "clw_mul_records_lanes": (i32, [vp, P(TpTable), P(SegBuffers), u32, u32, u32, u32, u32, vp, vp, u32]),
        "clw_session_create": (i32, [vp, u64, u64, P(Options), P(vp)]),
        "clw_session_destroy": (i32, [vp]),
        "clw_session_info": (i32, [vp, P(RangeStats)]),
        "clw_session_run": (i32, [vp, u64, u64]),
        "clw_session_residues": (i32, [vp, vp, vp, u64]),
        "clw_session_primes": (i32, [vp, vp, u64]),
        "clw_session_iterate": (i32, [vp]),
        "clw_session_finalize": (i32, [vp, P(RangeStats), vp, vp, vp, u64]),
        "clw_session_kernel_stats": (i32, [vp, P(KernelStats)]),
        "clw_session_reset_stats": (i32, [vp]),
        "clw_ckpt_write": (i32, [C.c_char_p, P(WorkStatus), vp, vp]),
        "clw_ckpt_read": (i32, [C.c_char_p, C.c_char_p, u64, u64, u32, P(WorkStatus), vp, vp, P(i32), P(i32)]),
        "clw_session_set_residues": (i32, [vp, vp, vp, u64]),
        "clw_plain_mul": (i32, [vp, P(TpTable), P(SegBuffers), u32, u32, u32, u32, vp, vp]),
        "clw_plain_reduce": (i32, [vp, P(TpTable), u32, u32, vp, vp]),
        "clw_iterate_from": (i32, [vp, P(TpTable), vp]),
        "clw_device_primes": (i32, [vp, u64, u64, vp, u64, P(u64)]),
        "clw_device_primes_noalloc": (i32, [vp, u64, u64, vp, u64, P(u64)]),
        "clw_batch_create": (i32, [vp, vp, vp, u32, P(Options), P(vp)]),
        "clw_batch_destroy": (i32, [vp]),
        "clw_batch_size": (u32, [vp]),
        "clw_batch_info": (i32, [vp, u32, P(RangeStats)]),
        "clw_batch_run": (i32, [vp, u64, u64]),
        "clw_batch_applied": (u64, [vp]),
        "clw_batch_residues": (i32, [vp, u32, vp, vp, u64]),
        "clw_batch_set_residues": (i32, [vp, u32, vp, vp, u64, u64, u64]),
        "clw_batch_iterate": (i32, [vp]),
        "clw_batch_finalize": (i32, [vp]),
        "clw_batch_results": (i32, [vp, u32, P(RangeStats), vp, vp, vp, u64]),
        "clw_batch_kernel_stats": (i32, [vp, P(KernelStats)]),
        "clw_batch_autotune": (i32, [vp, u64, C.c_char_p, P(BandTune), u32]),
        "clw_batch_launch": (i32, [vp, u64, u64]),
        "clw_batch_wait": (i32, [vp]),
        "clw_batch_query": (i32, [vp]),
        "clw_batch_bind": (i32, [vp, vp]),
        "clw_batch_set_scratch": (i32, [vp, vp, vp, u64]),
        "clw_batch_ckpt_write": (i32, [vp, C.c_char_p]),
        "clw_batch_ckpt_read": (i32, [vp, C.c_char_p, C.c_char_p, P(i32), P(i32)]),
        "clw_batch_covered": (u64, [vp]),
        "clw_batch_device_bytes": (u64, [vp]),
        "clw_batch_finish_range": (i32, [vp, u32, P(RangeStats), vp, vp, vp, u64]),
        "clw_batch_finish_launch": (i32, [vp, u32]),
        "clw_workspace_create": (i32, [vp, u64, P(vp)]),
        "clw_workspace_destroy": (i32, [vp]),
        "clw_workspace_bytes": (u64, [vp]),
        "clw_batch_set_workspace": (i32, [vp, vp]),
        "clw_batch_finish_query": (i32, [vp, u32]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    global _bound
    _bound = sorted(sig)
    _lib = L
    return L


_bound: list = []


def bound_symbols() -> list[str]:
    """Names of the C ABI functions this module has ctypes signatures for."""
    lib()
    return list(_bound)


def last_error() -> str:
    return lib().clw_last_error().decode(errors="replace")


def check(rc: int, what: str = "") -> None:
    if rc != CLW_OK:
        raise ClwError(rc, f"{what}: {last_error()}" if what else last_error())


def host_primes(lo: int, hi: int) -> np.ndarray:
    L = lib()
    n = L.clw_host_primes(lo, hi, None, 0)
    out = np.zeros(max(n, 1), dtype=np.uint64)
    L.clw_host_primes(lo, hi, out.ctypes.data, n)
    return out[:n]


def format_result(p: int, quot: int) -> str | None:
    buf = C.create_string_buffer(128)
    return buf.value.decode() if lib().clw_format_result(p, quot, buf, 128) else None


def result_lines(p: np.ndarray, quot: np.ndarray) -> list[str]:
    """results.txt lines of the Wilson and near-Wilson primes among p (in
    order): clw_format_result on the entries with |w_p| < 1000, the only ones
    it formats."""
    near = np.nonzero(np.minimum(quot, p - quot) < 1000)[0]
    out = []
    for i in near.tolist():
        line = format_result(int(p[i]), int(quot[i]))
        if line:
            out.append(line)
    return out


GOOD_RESULTS_SUM = 0x6659912234F44428  # the reference's goodWilsonResults.txt (cl_wilson.cpp:970)


def verify_results(path: str, pmin: int, pmax: int, p, quot, expected_sum: int = GOOD_RESULTS_SUM) -> tuple[int, int]:
    """The -r check (clw_verify_results): (matched, listed) on success;
    raises ClwError (CLW_ERR_CHECK / CLW_ERR_IO) with the reference's message."""
    pa = np.ascontiguousarray(p, dtype=np.uint64)
    qa = np.ascontiguousarray(quot, dtype=np.uint64)
    m, n = C.c_uint32(0), C.c_uint32(0)
    check(lib().clw_verify_results(path.encode(), expected_sum, pmin, pmax, pa.ctypes.data, qa.ctypes.data, len(pa),
                                   C.byref(m), C.byref(n)), "verify_results")
    return m.value, n.value


def range_types(pmin: int, pmax: int) -> tuple[list[int], list[int]]:
    cnt = (C.c_uint32 * 3)()
    tg = (C.c_uint64 * 3)()
    check(lib().clw_range_types(pmin, pmax, cnt, tg), "range_types")
    return list(cnt), list(tg)


@dataclass
class RangeResult:
    pmin: int
    pmax: int
    p: np.ndarray
    quot: np.ndarray
    rem: np.ndarray
    checksum: int
    totalcount: int
    type_targets: list
    tpcnt: list
    resultcount: int
    lastp: int
    lastw: int
    limbs: int
    seconds: float
    results: list = field(default_factory=list)

    @property
    def checksum_hex(self) -> str:
        return f"{self.checksum:016X}"

    def results_txt(self) -> str:
        """results.txt contents (cl_wilson.cpp:657-688, 1093-1143)."""
        body = "".join(self.results)
        return body + ("no results\n" if not self.results else "") + f"{self.checksum:016X}\n"


class Device:
    """One GPU context (clw_init / clw_shutdown)."""

    def __init__(self, index: int = 0):
        self.ctx = C.c_void_p()
        check(lib().clw_init(index, C.byref(self.ctx)), "clw_init")

    def close(self) -> None:
        if self.ctx:
            lib().clw_shutdown(self.ctx)
            self.ctx = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def info(self) -> dict:
        d = DeviceInfo()
        check(lib().clw_get_device_info(self.ctx, C.byref(d)), "device_info")
        return dict(name=d.name.decode(), arch=d.arch.decode(), compute_units=d.compute_units,
                    clock_khz=d.clock_khz, global_mem=d.global_mem)

    def primes(self, lo: int, hi: int) -> np.ndarray:
        """Primes of [lo, hi) sieved on the device (the test-prime list)."""
        cap = (hi - lo) // 2 + 4
        out = np.zeros(cap, dtype=np.uint64)
        n = C.c_uint64(0)
        check(lib().clw_device_primes(self.ctx, lo, hi, out.ctypes.data, cap, C.byref(n)), "device_primes")
        return out[: n.value]

    def primes_noalloc(self, lo: int, hi: int, cap: int | None = None) -> tuple[np.ndarray, int]:
        """The same list through clw_device_primes_noalloc (no device
        allocation; hi - lo <= 10^7): (primes written, total count)."""
        cap = (hi - lo) // 2 + 4 if cap is None else cap
        out = np.zeros(max(cap, 1), dtype=np.uint64)
        n = C.c_uint64(0)
        check(lib().clw_device_primes_noalloc(self.ctx, lo, hi, out.ctypes.data, cap, C.byref(n)),
              "device_primes_noalloc")
        return out[: min(cap, n.value)], n.value

    def wilson_range(self, pmin: int, pmax: int, segment_span: int = 0, target_waves: int = 0,
                     verbose: bool = False, allow_check_failure: bool = False, mode: int = 0,
                     per_lane: int = 0) -> RangeResult:
        n = int(lib().clw_host_primes(pmin, pmax, None, 0))
        p = np.zeros(max(n, 1), dtype=np.uint64)
        q = np.zeros(max(n, 1), dtype=np.uint64)
        r = np.zeros(max(n, 1), dtype=np.uint64)
        st = RangeStats()
        opt = Options(segment_span, target_waves, 1 if verbose else 0, mode, per_lane)
        rc = lib().clw_wilson_range(self.ctx, pmin, pmax, C.byref(opt), C.byref(st), p.ctypes.data,
                                    q.ctypes.data, r.ctypes.data, n)
        if rc != CLW_OK and not (allow_check_failure and rc == CLW_ERR_CHECK):
            check(rc, f"wilson_range [{pmin}, {pmax})")
        res = RangeResult(pmin, pmax, p[:n], q[:n], r[:n], st.checksum, st.totalcount, list(st.typeTarget),
                          list(st.tpcnt), st.resultcount, st.lastp, st.lastw, st.limbs, st.seconds)
        res.results = result_lines(res.p, res.quot)
        return res


class Session:
    """One range in phases, as cl_wilson() runs it: create (setup), run
    windows of the sieving-prime axis, iterate, finalize."""

    def __init__(self, dev: "Device", pmin: int, pmax: int, segment_span: int = 0, target_waves: int = 0,
                 mode: int = 0, per_lane: int = 0):
        self.dev = dev
        self.h = C.c_void_p()
        opt = Options(segment_span, target_waves, 0, mode, per_lane)
        check(lib().clw_session_create(dev.ctx, pmin, pmax, C.byref(opt), C.byref(self.h)), "session_create")
        st = RangeStats()
        check(lib().clw_session_info(self.h, C.byref(st)))
        self.pmin, self.pmax = pmin, pmax
        self.n = st.tpcount
        self.type_targets = list(st.typeTarget)
        self.tpcnt = list(st.tpcnt)
        self.maxtarget = st.maxtarget
        self.limbs = st.limbs

    def close(self) -> None:
        if self.h:
            lib().clw_session_destroy(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def primes(self) -> np.ndarray:
        out = np.zeros(max(self.n, 1), dtype=np.uint64)
        check(lib().clw_session_primes(self.h, out.ctypes.data, self.n))
        return out[: self.n]

    def run(self, qlo: int, qhi: int) -> None:
        check(lib().clw_session_run(self.h, qlo, qhi), f"session_run [{qlo}, {qhi})")

    def totalcount(self) -> int:
        st = RangeStats()
        check(lib().clw_session_info(self.h, C.byref(st)))
        return st.totalcount

    def residues(self) -> list[int]:
        lo = np.zeros(max(self.n, 1), dtype=np.uint64)
        hi = np.zeros(max(self.n, 1), dtype=np.uint64)
        check(lib().clw_session_residues(self.h, lo.ctypes.data, hi.ctypes.data, self.n))
        return [int(a) | (int(b) << 64) for a, b in zip(lo[: self.n], hi[: self.n])]

    def iterate(self) -> None:
        check(lib().clw_session_iterate(self.h), "session_iterate")

    def set_residues(self, values: list[int], totalcount: int) -> None:
        lo = np.array([v & (2**64 - 1) for v in values], dtype=np.uint64)
        hi = np.array([v >> 64 for v in values], dtype=np.uint64)
        check(lib().clw_session_set_residues(self.h, lo.ctypes.data, hi.ctypes.data, totalcount),
              "session_set_residues")

    def checkpoint(self, path: str, currp: int, done: int = 0, trickle: int = 0) -> WorkStatus:
        """Writes the reference's checkpoint layout for this session."""
        vals = self.residues()
        st = WorkStatus(self.pmin, self.pmax, currp, trickle, 0, self.totalcount(), self.n, done)
        ckpt_write(path, st, vals)
        return st

    def kernel_stats(self) -> dict:
        ks = KernelStats()
        check(lib().clw_session_kernel_stats(self.h, C.byref(ks)))
        return ks.as_dict()

    def reset_stats(self) -> None:
        check(lib().clw_session_reset_stats(self.h))

    def finalize(self, allow_check_failure: bool = False) -> RangeResult:
        n = self.n
        p = np.zeros(max(n, 1), dtype=np.uint64)
        q = np.zeros(max(n, 1), dtype=np.uint64)
        r = np.zeros(max(n, 1), dtype=np.uint64)
        st = RangeStats()
        rc = lib().clw_session_finalize(self.h, C.byref(st), p.ctypes.data, q.ctypes.data, r.ctypes.data, n)
        if rc != CLW_OK and not (allow_check_failure and rc == CLW_ERR_CHECK):
            check(rc, "session_finalize")
        res = RangeResult(self.pmin, self.pmax, p[:n], q[:n], r[:n], st.checksum, st.totalcount,
                          list(st.typeTarget), list(st.tpcnt), st.resultcount, st.lastp, st.lastw, st.limbs,
                          st.seconds)
        res.results = result_lines(res.p, res.quot)
        return res


class Scratch:
    """Device buffers for chunk products (2 x cap 64-bit slots) shared by
    batches that run one after another on one stream."""

    def __init__(self, dev: "Device", cap: int):
        self.dev, self.cap = dev, cap
        lo, hi = C.c_void_p(), C.c_void_p()
        check(lib().clw_alloc(dev.ctx, cap * 8, C.byref(lo)), "scratch alloc")
        check(lib().clw_alloc(dev.ctx, cap * 8, C.byref(hi)), "scratch alloc")
        self.lo, self.hi = lo.value, hi.value

    def close(self) -> None:
        if self.lo:
            lib().clw_free(self.dev.ctx, self.lo)
            lib().clw_free(self.dev.ctx, self.hi)
            self.lo = self.hi = None


class Workspace:
    """Window buffers (record buffer, sieve bitmap, run products) shared by
    batches that run their windows one after another on one stream
    (clw_workspace_*), sized for batches of the given segment span."""

    def __init__(self, dev: "Device", segment_span: int = 0):
        self.dev = dev
        self.h = C.c_void_p()
        check(lib().clw_workspace_create(dev.ctx, segment_span, C.byref(self.h)), "workspace_create")
        self.bytes = lib().clw_workspace_bytes(self.h)

    def close(self) -> None:
        if self.h:
            lib().clw_workspace_destroy(self.h)
            self.h = C.c_void_p()


class Batch:
    """Many ranges searched together over one shared sieve (clw_batch_*).
    Each range's results equal clw_wilson_range's on it, bit for bit."""

    def __init__(self, dev: "Device", ranges: list[tuple[int, int]], segment_span: int = 0,
                 target_waves: int = 0, per_lane: int = 0, graph: bool = True, segment_apps: int = 0,
                 host_primes: bool = False, max_chunks: int = 0, external_scratch: bool = False,
                 run_products: int = 0, mfma_slices: int = 0, external_workspace: bool = False):
        if not ranges:
            raise ValueError("a batch needs at least one range")
        self.dev = dev
        self.h = C.c_void_p()
        self.ranges = [(int(a), int(b)) for a, b in ranges]
        lo = np.array([a for a, _ in self.ranges], dtype=np.uint64)
        hi = np.array([b for _, b in self.ranges], dtype=np.uint64)
        opt = Options(segment_span, target_waves, 0, 0, per_lane, 1 if graph else 0, segment_apps,
                      1 if host_primes else 0, max_chunks, 1 if external_scratch else 0,
                      run_products, mfma_slices, 1 if external_workspace else 0)
        check(lib().clw_batch_create(dev.ctx, lo.ctypes.data, hi.ctypes.data, len(self.ranges), C.byref(opt),
                                     C.byref(self.h)), "batch_create")
        self.info = []
        for r in range(len(self.ranges)):
            st = RangeStats()
            check(lib().clw_batch_info(self.h, r, C.byref(st)))
            self.info.append(st)
        self.maxtarget = max(st.maxtarget for st in self.info)

    def close(self) -> None:
        if self.h:
            lib().clw_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __len__(self) -> int:
        return len(self.ranges)

    def run(self, qlo: int, qhi: int) -> None:
        check(lib().clw_batch_run(self.h, qlo, qhi), f"batch_run [{qlo}, {qhi})")

    def launch(self, qlo: int, qhi: int) -> None:
        """Enqueue a window on this batch's stream (returns at once)."""
        check(lib().clw_batch_launch(self.h, qlo, qhi), f"batch_launch [{qlo}, {qhi})")

    def wait(self) -> None:
        check(lib().clw_batch_wait(self.h), "batch_wait")

    def checkpoint(self, path: str) -> None:
        """Write every range's running state and the covered prefix."""
        check(lib().clw_batch_ckpt_write(self.h, path.encode()), "batch_ckpt_write")

    def resume(self, path_a: str, path_b: str) -> tuple[int, int]:
        """(result, which): 1 if a valid checkpoint was loaded (continue from
        covered()), else 0; which = 0 for path_a, 1 for path_b."""
        which, result = C.c_int(-1), C.c_int(0)
        check(lib().clw_batch_ckpt_read(self.h, path_a.encode(), path_b.encode(), C.byref(which),
                                        C.byref(result)), "batch_ckpt_read")
        return result.value, which.value

    def set_scratch(self, scratch: "Scratch") -> None:
        """Keep chunk products in a shared device scratch (see Scratch)."""
        check(lib().clw_batch_set_scratch(self.h, scratch.lo, scratch.hi, scratch.cap), "batch_set_scratch")

    def set_workspace(self, ws: Workspace) -> None:
        """Run this batch's windows in ws's buffers (ws shared by batches
        whose windows run one after another on one stream)."""
        check(lib().clw_batch_set_workspace(self.h, ws.h), "batch_set_workspace")

    def bind(self, dev: "Device") -> None:
        """Run this batch on another device context (stream) from now on."""
        check(lib().clw_batch_bind(self.h, dev.ctx), "batch_bind")
        self.dev = dev

    def done(self) -> bool:
        """True when the launched window has finished (never blocks)."""
        rc = lib().clw_batch_query(self.h)
        if rc < 0:
            check(rc, "batch_query")
        return rc == 1

    def covered(self) -> int:
        """Sieving primes [2, covered) have been applied to every range."""
        return lib().clw_batch_covered(self.h)

    def device_bytes(self) -> int:
        return lib().clw_batch_device_bytes(self.h)

    def finish_range(self, r: int, allow_check_failure: bool = False) -> RangeResult:
        """Iterate and check range r alone once its sieving primes are all
        applied (covered() > its largest target); the rest keep running.
        Collects the results of finish_launch(r) if that was called first."""
        return self._results(r, lib().clw_batch_finish_range, allow_check_failure)

    def finish_launch(self, r: int) -> None:
        """finish_range in the background: enqueues range r's iterate and
        check (and the copy of its results) on the batch's stream and returns;
        finish_ready tells when finish_range will collect without waiting."""
        check(lib().clw_batch_finish_launch(self.h, r), f"batch_finish_launch {r}")

    def finish_ready(self, r: int) -> bool:
        rc = lib().clw_batch_finish_query(self.h, r)
        if rc < 0:
            check(rc, f"batch_finish_query {r}")
        return rc == 1

    def applied(self) -> int:
        return lib().clw_batch_applied(self.h)

    def residues(self, r: int) -> list[int]:
        n = self.info[r].tpcount
        lo = np.zeros(max(n, 1), dtype=np.uint64)
        hi = np.zeros(max(n, 1), dtype=np.uint64)
        check(lib().clw_batch_residues(self.h, r, lo.ctypes.data, hi.ctypes.data, n))
        return [int(a) | (int(b) << 64) for a, b in zip(lo[:n], hi[:n])]

    def set_residues(self, r: int, values: list[int], covered: int, applied: int) -> None:
        """Resume a one-range batch from residues that hold the sieving
        primes [2, covered) (a checkpoint of the command line's format)."""
        lo = np.array([v & (2**64 - 1) for v in values], dtype=np.uint64)
        hi = np.array([v >> 64 for v in values], dtype=np.uint64)
        check(lib().clw_batch_set_residues(self.h, r, lo.ctypes.data, hi.ctypes.data, len(values), covered,
                                           applied), "batch_set_residues")

    def iterate(self) -> None:
        check(lib().clw_batch_iterate(self.h), "batch_iterate")

    def finalize(self) -> None:
        check(lib().clw_batch_finalize(self.h), "batch_finalize")

    def results(self, r: int, allow_check_failure: bool = False) -> RangeResult:
        return self._results(r, lib().clw_batch_results, allow_check_failure)

    def _results(self, r, fn, allow_check_failure) -> RangeResult:
        n = self.info[r].tpcount
        p = np.zeros(max(n, 1), dtype=np.uint64)
        q = np.zeros(max(n, 1), dtype=np.uint64)
        m = np.zeros(max(n, 1), dtype=np.uint64)
        st = RangeStats()
        rc = fn(self.h, r, C.byref(st), p.ctypes.data, q.ctypes.data, m.ctypes.data, n)
        if rc != CLW_OK and not (allow_check_failure and rc == CLW_ERR_CHECK):
            check(rc, f"batch_results {r}")
        pmin, pmax = self.ranges[r]
        res = RangeResult(pmin, pmax, p[:n], q[:n], m[:n], st.checksum, st.totalcount, list(st.typeTarget),
                          list(st.tpcnt), st.resultcount, st.lastp, st.lastw, st.limbs, st.seconds)
        res.results = result_lines(res.p, res.quot)
        return res

    def search(self) -> list[RangeResult]:
        """The whole plain method for every range: all sieving primes up to
        the largest target, iterate, finalize."""
        self.run(2, self.maxtarget + 1)
        self.iterate()
        self.finalize()
        return [self.results(r) for r in range(len(self.ranges))]

    def kernel_stats(self) -> dict:
        ks = KernelStats()
        check(lib().clw_batch_kernel_stats(self.h, C.byref(ks)))
        return ks.as_dict()

    def autotune(self, q: int, cache_path: str | None = None) -> list[dict]:
        """Per limb band: the matrix-core row slices and switch (bands with
        matrix-core run products), then test primes per lane and items per
        resident block, timed on trial segments at magnitude q (no residue
        changes)."""
        rep = (BandTune * 4)()
        check(lib().clw_batch_autotune(self.h, q, cache_path.encode() if cache_path else None, rep, 4),
              "batch_autotune")
        return [rep[k].as_dict() for k in range(4) if rep[k].limbs]


def ckpt_write(path: str, st: WorkStatus, values: list[int]) -> None:
    lo = np.array([v & (2**64 - 1) for v in values] or [0], dtype=np.uint64)
    hi = np.array([v >> 64 for v in values] or [0], dtype=np.uint64)
    check(lib().clw_ckpt_write(path.encode(), C.byref(st), lo.ctypes.data, hi.ctypes.data), "ckpt_write")


def ckpt_read(path_a: str, path_b: str, pmin: int, pmax: int, tpcount: int):
    """(result, which, WorkStatus, residues): result 0 none, 1 resume, 2 done."""
    st = WorkStatus()
    lo = np.zeros(max(tpcount, 1), dtype=np.uint64)
    hi = np.zeros(max(tpcount, 1), dtype=np.uint64)
    which, result = C.c_int(-1), C.c_int(0)
    check(lib().clw_ckpt_read(path_a.encode(), path_b.encode(), pmin, pmax, tpcount, C.byref(st), lo.ctypes.data,
                              hi.ctypes.data, C.byref(which), C.byref(result)), "ckpt_read")
    vals = [int(a) | (int(b) << 64) for a, b in zip(lo[:tpcount], hi[:tpcount])]
    return result.value, which.value, st, vals


def wilson_quotient(p: int, quot: int) -> int:
    """Signed w_p as the result lines print it (cl_wilson.cpp:872-898)."""
    neg = p - quot
    return quot if quot <= neg else -neg
