"""The -r result verification (clw_verify_results; cl_wilson.cpp:902-930,
936-1014, 1073-1087) on synthetic known-good files built from the
Costa-Gerbicz-Harvey tables: host-only, no GPU."""
import numpy as np
import pytest

import clwilson_amd as cw


def _quot(p, w):
    """The quotient whose printed w_p is w (cl_wilson.cpp:872-898)."""
    return w if w >= 0 else p + w


def _write(path, rows):
    with open(path, "w") as f:
        for p, w in rows:
            f.write(f"{p} {w}\n")
    return sum(p + w for p, w in rows) % 2**64


def _cgh(golden):
    g = golden("cgh_tables.json")
    seen = {}
    for p, w in g["table1"] + g["table2"]:
        if p >= 5:
            seen[p] = w
    return sorted(seen.items())


def test_all_listed_results_match(golden, tmp_path):
    rows = _cgh(golden)
    path = str(tmp_path / "good.txt")
    s = _write(path, rows)
    # the tested primes: every listed one plus ordinary primes (|w_p/p| large)
    p = [r[0] for r in rows] + [1000003, 1000033]
    q = [_quot(a, w) for a, w in rows] + [500000, 123456]
    m, n = cw.verify_results(path, 5, 2 * 10**13, p, q, expected_sum=s)
    assert (m, n) == (len(rows), len(rows))
    # a sub-range: only the file's lines inside it count
    sub = [(a, w) for a, w in rows if 10**6 <= a < 10**9]
    m, n = cw.verify_results(path, 10**6, 10**9, [a for a, _ in sub], [_quot(a, w) for a, w in sub], expected_sum=s)
    assert (m, n) == (len(sub), len(sub))


def test_reference_file_checksum_constant(golden, tmp_path):
    """A file whose sum is not the reference's constant is refused."""
    path = str(tmp_path / "good.txt")
    _write(path, _cgh(golden))
    with pytest.raises(cw.ClwError, match="Checksum error"):
        cw.verify_results(path, 5, 2 * 10**13, [], [])


def test_mismatches(golden, tmp_path):
    rows = _cgh(golden)
    path = str(tmp_path / "good.txt")
    s = _write(path, rows)
    p = [r[0] for r in rows]
    q = [_quot(a, w) for a, w in rows]
    # a different w_p
    bad = list(q)
    bad[3] = _quot(p[3], rows[3][1] + 1)
    with pytest.raises(cw.ClwError, match=f"p: {p[3]} "):
        cw.verify_results(path, 5, 2 * 10**13, p, bad, expected_sum=s)
    # a tested prime below the threshold that the file does not list
    with pytest.raises(cw.ClwError, match="is not in result file"):
        cw.verify_results(path, 5, 2 * 10**13, p + [1000003], q + [7], expected_sum=s)
    # a listed prime of the range that was not tested
    with pytest.raises(cw.ClwError, match="was not found in results"):
        cw.verify_results(path, 5, 2 * 10**13, p[1:], q[1:], expected_sum=s)
    # an unreadable line
    with open(path, "a") as f:
        f.write("not a result\n")
    with pytest.raises(cw.ClwError, match="Error reading"):
        cw.verify_results(path, 5, 2 * 10**13, p, q, expected_sum=s)
    with pytest.raises(cw.ClwError):
        cw.verify_results(str(tmp_path / "missing.txt"), 5, 2 * 10**13, p, q, expected_sum=s)


def test_threshold_rounds_the_ratio_to_double(tmp_path):
    """cl_wilson.cpp:906-907 rounds min(quot, p - quot) / p (computed in long
    double) to double before comparing with 1/50000: a ratio just below
    1/50000 in long double that rounds up to the double threshold is not
    checked (a long-double compare would demand it in the file)."""
    thr = np.float64(1.0) / np.float64(50000.0)
    # double(1/50000) lies 1.636e-21 above 1/50000 and half an ulp is
    # 1.694e-21, so ratios in the 5.8e-23 below 1/50000 round up to it:
    # s / (50000 s + 1) gets there for p above ~3.45e17
    found = None
    for s in range(8 * 10**12, 8 * 10**12 + 5000):
        p = 50000 * s + 1
        r = np.longdouble(s) / np.longdouble(p)
        if r < np.longdouble(1) / np.longdouble(50000) and not (np.float64(r) < thr):
            found = (p, s)
            break
    assert found, "no ratio straddles the double rounding in this span"
    p, s = found
    path = str(tmp_path / "empty.txt")
    open(path, "w").close()
    assert cw.verify_results(path, p, p + 1, [p], [s], expected_sum=0) == (0, 0)
    # one ulp further below is checked, and fails since the file is empty
    with pytest.raises(cw.ClwError, match="is not in result file"):
        cw.verify_results(path, p, p + 1, [p], [s - 1], expected_sum=0)
