"""Checkpoint / resume on the GPU and the command line end to end."""
import os
import subprocess

import numpy as np
import pytest

import clwilson_amd as cw

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_resume_equals_uninterrupted(dev, tmp_path):
    lo, hi = 10**9, 10**9 + 20000
    ref = dev.wilson_range(lo, hi)
    a, b = str(tmp_path / "stateA.ckp"), str(tmp_path / "stateB.ckp")
    cut = 123_456_789
    with cw.Session(dev, lo, hi) as s:
        s.run(2, cut)
        s.checkpoint(a, cut)
        n = s.n
    res, which, st, vals = cw.ckpt_read(a, b, lo, hi, n)
    assert res == 1 and which == 0 and st.currp == cut
    with cw.Session(dev, lo, hi) as s:
        s.set_residues(vals, st.totalcount)
        s.run(st.currp, s.maxtarget + 1)
        s.iterate()
        out = s.finalize()
    assert np.array_equal(out.quot, ref.quot) and out.checksum == ref.checksum
    assert out.totalcount == ref.totalcount


def test_batch_resume_from_reference_checkpoint(dev, tmp_path):
    """The command line runs a one-range batch and keeps the reference's
    checkpoint files: residues written from a batch after [2, cut) and read
    back into a new batch (clw_batch_set_residues) finish bit for bit like
    an uninterrupted run, at 3 limbs through the matrix-core run products."""
    lo, hi = 10**9, 10**9 + 200000
    ref = dev.wilson_range(lo, hi)
    a, b = str(tmp_path / "stateA.ckp"), str(tmp_path / "stateB.ckp")
    cut = 123_456_789
    with cw.Batch(dev, [(lo, hi)]) as bt:
        bt.run(2, cut)
        vals, applied = bt.residues(0), bt.applied()
    with cw.Session(dev, lo, hi) as s:  # the file layout of the reference (cl_wilson.cpp:1779-1810)
        s.set_residues(vals, applied)
        s.checkpoint(a, cut)
        n = s.n
    res, which, st, back = cw.ckpt_read(a, b, lo, hi, n)
    assert res == 1 and st.currp == cut and back == vals
    with cw.Batch(dev, [(lo, hi)]) as bt:
        bt.set_residues(0, back, st.currp, st.totalcount)
        bt.run(st.currp, bt.maxtarget + 1)
        bt.iterate()
        bt.finalize()
        out = bt.results(0)
        total = bt.applied()
    assert np.array_equal(out.quot, ref.quot) and out.checksum == ref.checksum
    assert total == ref.totalcount


def test_cli_large_range_through_the_batch(dev, tmp_path):
    """A range of >= 65,536 test primes goes through a one-range batch on the
    command line (matrix-core run products at 3 limbs): results.txt ends
    with the session path's checksum."""
    lo, hi = 10**9, 10**9 + 2_000_000
    ref = dev.wilson_range(lo, hi, mode=1)
    assert len(ref.p) >= 65536
    exe = os.path.join(ROOT, "clwilson_amd", "clwilson")
    r = subprocess.run([exe, "-p", str(lo), "-P", str(hi)], cwd=tmp_path, capture_output=True, text=True, timeout=280)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = open(tmp_path / "results.txt").read().splitlines()
    assert lines[-1] == ref.checksum_hex


def test_cli_range_results_file_and_done_state(tmp_path):
    exe = os.path.join(ROOT, "clwilson_amd", "clwilson")
    r = subprocess.run([exe, "-p", "200", "-P", "564"], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = open(tmp_path / "results.txt").read().splitlines()
    assert len(lines) == 58 and lines[-1] == "00000000000080A3"
    assert lines[-2] == "563 is a Wilson prime"
    assert os.path.exists(tmp_path / "stateA.ckp")
    r = subprocess.run([exe, "-p", "200", "-P", "564"], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and "Workunit complete." in r.stdout


def test_cli_self_test(tmp_path):
    """The -s path of the command line: the reference's seven self-test
    ranges (cl_wilson.cpp:1998-2139) through the binary, each checked
    against its result count, checksum, pi(maxT), last prime and last w_p."""
    exe = os.path.join(ROOT, "clwilson_amd", "clwilson")
    print("clwilson -s", flush=True)
    r = subprocess.run([exe, "-s"], cwd=tmp_path, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-2000:]
    assert "All test cases completed successfully!" in r.stdout
    assert sum(f"test case {k} passed." in r.stdout for k in range(1, 8)) == 7


def test_cli_suffix_arguments(tmp_path):
    exe = os.path.join(ROOT, "clwilson_amd", "clwilson")
    r = subprocess.run([exe, "-p", "86M", "-P", "86010000"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    import oracle as O
    arr, st = O.wilson_range(86000000, 86010000)
    assert open(tmp_path / "results.txt").read().splitlines()[-1] == f"{st['checksum']:016X}"


def test_cli_failed_results_write_is_not_complete(tmp_path):
    """results.txt that cannot be written (a directory of that name): the run
    fails and writes no finished checkpoint, so the rerun searches again and
    writes the results instead of reporting 'Workunit complete.'."""
    exe = os.path.join(ROOT, "clwilson_amd", "clwilson")
    os.mkdir(tmp_path / "results.txt")
    r = subprocess.run([exe, "-p", "200", "-P", "564"], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode != 0, r.stdout + r.stderr
    os.rmdir(tmp_path / "results.txt")
    r = subprocess.run([exe, "-p", "200", "-P", "564"], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Workunit complete." not in r.stdout
    assert open(tmp_path / "results.txt").read().splitlines()[-1] == "00000000000080A3"


def _good_file(path, rows):
    """A known-good list whose p + w sum equals the reference file's constant
    (one balancing line far outside any searched range)."""
    s = sum(p + w for p, w in rows) % 2**64
    rows = rows + [((cw.GOOD_RESULTS_SUM - s) % 2**64, 0)]
    with open(path, "w") as f:
        for p, w in rows:
            f.write(f"{p} {w}\n")


def test_cli_result_verification(tmp_path):
    """-r (cl_wilson.cpp:902-1014): 563 is the only prime of [200, 564) with
    |w_p/p| < 1/50000 (w = 0); a file listing it passes, one with another
    w_p fails and leaves the workunit unfinished."""
    exe = os.path.join(ROOT, "clwilson_amd", "clwilson")
    _good_file(tmp_path / "goodWilsonResults.txt", [(5, 0), (13, 0), (563, 0), (8315831, 3)])
    r = subprocess.run([exe, "-p", "200", "-P", "564", "-r"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "All results matched result file!" in r.stdout
    d = tmp_path / "bad"
    d.mkdir()
    _good_file(d / "goodWilsonResults.txt", [(563, 1)])
    r = subprocess.run([exe, "-p", "200", "-P", "564", "-r"], cwd=d, capture_output=True, text=True, timeout=300)
    assert r.returncode != 0 and "result check failed! p: 563" in r.stdout
    assert not os.path.exists(d / "results.txt")
