"""Checkpoint files (cl_wilson.cpp:110-319): layout, checksum, A/B choice,
rejection rules. Host-only; the GPU resume path is in test_gpu_resume.py."""
import os
import struct

import clwilson_amd as cw


def _vals(n, seed=1):
    import random
    r = random.Random(seed)
    return [r.getrandbits(100) for _ in range(n)]


def test_layout_and_checksum(tmp_path):
    p = str(tmp_path / "stateA.ckp")
    vals = _vals(5)
    st = cw.WorkStatus(100, 200, 150, 7, 0, 42, 5, 0)
    cw.ckpt_write(p, st, vals)
    raw = open(p, "rb").read()
    assert len(raw) == 56 + 16 * 5
    pmin, pmax, currp, trickle, ssum, total, tpc, done = struct.unpack("<6Q2I", raw[:56])
    assert (pmin, pmax, currp, trickle, total, tpc, done) == (100, 200, 150, 7, 42, 5, 0)
    words = struct.unpack("<10Q", raw[56:])
    # the reference's sum (cl_wilson.cpp:116-119) plus the format constant "CLWAMD01"
    expect = (0x434C57414D443031 + 100 + 200 + 150 + 7 + 42 + 5 + 0 + sum(words)) % 2**64
    assert ssum == expect
    assert [words[2 * i] | (words[2 * i + 1] << 64) for i in range(5)] == vals


def test_pick_newest_and_fallbacks(tmp_path):
    a, b = str(tmp_path / "stateA.ckp"), str(tmp_path / "stateB.ckp")
    va, vb = _vals(4, 1), _vals(4, 2)
    cw.ckpt_write(a, cw.WorkStatus(10, 20, 300, 0, 0, 9, 4, 0), va)
    cw.ckpt_write(b, cw.WorkStatus(10, 20, 500, 0, 0, 11, 4, 0), vb)
    res, which, st, vals = cw.ckpt_read(a, b, 10, 20, 4)
    assert (res, which, st.currp, st.totalcount, vals) == (1, 1, 500, 11, vb)
    # corrupt B's residues: checksum mismatch -> A
    raw = bytearray(open(b, "rb").read())
    raw[60] ^= 1
    open(b, "wb").write(bytes(raw))
    res, which, st, vals = cw.ckpt_read(a, b, 10, 20, 4)
    assert (res, which, st.currp, vals) == (1, 0, 300, va)
    # truncated A -> nothing valid
    open(a, "wb").write(open(a, "rb").read()[:70])
    assert cw.ckpt_read(a, b, 10, 20, 4)[0] == 0
    # another range or count -> rejected
    cw.ckpt_write(a, cw.WorkStatus(10, 21, 300, 0, 0, 9, 4, 0), va)
    assert cw.ckpt_read(a, b, 10, 20, 4)[0] == 0
    assert cw.ckpt_read(a, b, 10, 21, 3)[0] == 0
    # missing files
    assert cw.ckpt_read(str(tmp_path / "x"), str(tmp_path / "y"), 10, 20, 4)[0] == 0


def test_done_flag(tmp_path):
    a, b = str(tmp_path / "stateA.ckp"), str(tmp_path / "stateB.ckp")
    cw.ckpt_write(a, cw.WorkStatus(10, 20, 300, 0, 0, 9, 4, 0), _vals(4))
    cw.ckpt_write(b, cw.WorkStatus(10, 20, 300, 0, 0, 9, 4, 1), _vals(4))
    assert cw.ckpt_read(a, b, 10, 20, 4)[0] == 2


def test_reference_written_files_are_rejected(tmp_path):
    """A file with the reference's own checksum (cl_wilson.cpp:116-119: the
    plain sum, over Montgomery-form residues) for the same range and count
    never validates here, done or not: its residues are in another form."""
    a, b = str(tmp_path / "stateA.ckp"), str(tmp_path / "stateB.ckp")
    vals = _vals(4, 3)
    for done in (0, 1):
        words = [w for v in vals for w in (v & (2**64 - 1), v >> 64)]
        ssum = (10 + 20 + 300 + 0 + 9 + 4 + done + sum(words)) % 2**64
        raw = struct.pack("<6Q2I", 10, 20, 300, 0, ssum, 9, 4, done) + struct.pack(f"<{len(words)}Q", *words)
        open(a, "wb").write(raw)
        assert cw.ckpt_read(a, b, 10, 20, 4)[0] == 0
    # this library's own file of the same content is accepted
    cw.ckpt_write(a, cw.WorkStatus(10, 20, 300, 0, 0, 9, 4, 0), vals)
    res, which, st, got = cw.ckpt_read(a, b, 10, 20, 4)
    assert (res, which, got) == (1, 0, vals)


def test_done_file_with_a_bad_sum_is_not_done(tmp_path):
    a, b = str(tmp_path / "stateA.ckp"), str(tmp_path / "stateB.ckp")
    cw.ckpt_write(a, cw.WorkStatus(10, 20, 300, 0, 0, 9, 4, 1), _vals(4))
    raw = bytearray(open(a, "rb").read())
    raw[-1] ^= 0x40
    open(a, "wb").write(bytes(raw))
    assert cw.ckpt_read(a, b, 10, 20, 4)[0] == 0
