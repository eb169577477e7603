"""The C ABI library: loads, exports every function include/clwilson.h
declares, and behaves on host-only paths (no GPU compute here)."""
import ctypes as C
import os
import re
import subprocess

import pytest
import sympy

import clwilson_amd as cw

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "clwilson.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(clw_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(cw.LIB_PATH)
    names = _declared()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", cw.LIB_PATH], capture_output=True, text=True).stdout
    for n in names:
        assert re.search(rf"\bT {n}$", out, flags=re.M), n


def test_binding_covers_header():
    """The Python host binding has a signature for every declared function."""
    missing = set(_declared()) - set(cw.bound_symbols())
    assert not missing, missing


def test_host_primes_matches_sympy():
    assert list(cw.host_primes(0, 5000)) == list(sympy.primerange(0, 5000))
    lo = 2**61
    assert list(cw.host_primes(lo, lo + 3000)) == list(sympy.primerange(lo, lo + 3000))
    assert len(cw.host_primes(1000000, 2000000)) == 70435
    assert len(cw.host_primes(20000000000000, 20000010000000)) == 327107


def test_range_types_match_survey_counts():
    cnt, tg = cw.range_types(1000000, 2000000)
    assert cnt == [35181, 17625, 17629] and tg == [166667, 250009, 500075]
    cnt, tg = cw.range_types(20000000000000, 20000010000000)
    assert cnt == [163479, 81891, 81737]
    assert tg == [3333333333340, 5000000000005, 10000000000055]


def test_limbs_for_bands():
    L = cw.lib()
    # L limbs of 30 bits serve p while 2 p^2 < 2^(30 L)
    assert L.clw_limbs_for(1000000) == 2
    assert L.clw_limbs_for(759250124) == 2 and L.clw_limbs_for(759250125) == 3
    assert L.clw_limbs_for(20000010000000) == 3
    assert L.clw_limbs_for(24879108095803) == 3 and L.clw_limbs_for(24879108095804) == 4
    assert L.clw_limbs_for(815238614083298888) == 4 and L.clw_limbs_for(815238614083298889) == 5
    assert L.clw_limbs_for(2**62 - 1) == 5


def test_format_result_lines():
    assert cw.format_result(563, 0) == "563 is a Wilson prime\n"
    assert cw.format_result(1239053554603, 1239053554599) == "1239053554603 is a Near-Wilson prime -4\n"
    assert cw.format_result(17524177394617, 256) == "17524177394617 is a Near-Wilson prime +256\n"
    assert cw.format_result(1000003, 500000) is None


def test_blocks_upto():
    L = cw.lib()
    assert L.clw_blocks_upto(7, 1000, 500) == 1
    assert L.clw_blocks_upto(7, 10**6, 5) == 0
    assert L.clw_blocks_upto(30, 30 + 2 * cw.BLOCK_SPAN, 10**9) == 2


def test_errors_are_returned_not_fatal():
    L = cw.lib()
    st = cw.RangeStats()
    assert L.clw_wilson_range(None, 10, 20, None, C.byref(st), None, None, None, 0) == cw.CLW_ERR_ARG
    assert "null" in cw.last_error()
    ctx = C.c_void_p()
    rc = L.clw_init(999, C.byref(ctx))
    assert rc in (cw.CLW_ERR_ARG, cw.CLW_ERR_HIP) and not ctx.value
    assert cw.last_error()


def test_cli_help_and_argument_parsing():
    exe = os.path.join(ROOT, "clwilson_amd", "clwilson")
    r = subprocess.run([exe, "-h"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "-p #" in r.stdout
    r = subprocess.run([exe, "-p", "12x"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "invalid argument" in r.stdout
    r = subprocess.run([exe, "-p", "2"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "out of range" in r.stdout


# ctypes mirror of each struct the header declares for the binding
_STRUCTS = {"DeviceInfo": "clw_device_info", "TpTable": "clw_tp_table", "SegBuffers": "clw_seg_buffers",
            "KernelStats": "clw_kernel_stats", "BandTune": "clw_band_tune", "WorkStatus": "clw_work_status",
            "Options": "clw_options", "RangeStats": "clw_range_stats"}


def test_binding_structs_match_header_layout(tmp_path):
    """Every ctypes struct of the binding has the size and field offsets that
    a C compiler gives the header's struct (a field added on one side only
    shifts everything after it)."""
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "clwilson.h"', "int main(void) {"]
    want = []
    for py, c in _STRUCTS.items():
        cls = getattr(cw, py)
        lines.append(f'  printf("%zu\\n", sizeof({c}));')
        want.append(C.sizeof(cls))
        for name, *_ in cls._fields_:
            lines.append(f'  printf("%zu\\n", offsetof({c}, {name}));')
            want.append(getattr(cls, name).offset)
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", f"-I{os.path.join(ROOT, 'include')}", str(src), "-o", str(exe)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    assert got == want
