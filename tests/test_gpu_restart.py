"""Native restart and snapshot files (imhd_write_restart / imhd_read_restart /
imhd_write_cube; io_mod.f90:249-517) against the restatement in oracle/restart_format.py
and the digests of files the reference's own io_mod wrote (tests/golden/restart_vectors.json).
FP64 contexts must reproduce the reference's bytes exactly; FP32 contexts write their
(widened) FP32 state, checked against the restatement applied to get_state."""
import hashlib
import json
import os
import sys

import numpy as np
import pytest

from refutil import io_state

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "oracle"))
import restart_format as RF  # noqa: E402

pytestmark = pytest.mark.gpu
GOLD = json.load(open(os.path.join(HERE, "golden", "restart_vectors.json")))


def say(*a):
    print(*a, flush=True)


def sha(b):
    return hashlib.sha256(b).hexdigest()


def solver(dims, npass, precision=64):
    from isothermal_mhd_amd import VanLeerSolver
    return VanLeerSolver(dims, npass=npass, riemann="hlld", bc="periodic", precision=precision)


@pytest.mark.parametrize("name", sorted(GOLD))
def test_native_files_match_reference_digests(name, tmp_path):
    g = GOLD[name]
    dims, npass, dp = tuple(g["dims"]), g["npass"], g["double_prec"]
    say(f"[gpu] restart/cube digests {name}")
    q, bf, qp = io_state(*dims, npass=npass, seed=g["seed"])
    S = solver(dims, npass)
    S.set_state(q, bf, qp)
    S.write_restart(str(tmp_path / "r.dat"), g["t_sim"], g["dump_count"], g["seeds"])
    assert sha((tmp_path / "r.dat").read_bytes()) == g["restart_sha256"]
    S.write_cube(str(tmp_path / "c.dat"), double_prec=dp)
    assert sha((tmp_path / "c.dat").read_bytes()) == g["cube_sha256"]
    S.write_cube(str(tmp_path / "p.dat"), double_prec=dp, npass_out=npass)
    assert sha((tmp_path / "p.dat").read_bytes()) == g["parallel_sha256"]
    # the other precision of the cube against the restatement
    with np.errstate(over="ignore"):
        want = RF.cube_bytes(q, qp, dims, 6, not dp, min(npass, 1))
    S.write_cube(str(tmp_path / "c2.dat"), double_prec=not dp)
    assert (tmp_path / "c2.dat").read_bytes() == want
    S.close()


def test_read_restart_sets_state_fp64(tmp_path):
    g = GOLD["vl_hlld_p12x10x8"]
    dims, npass = tuple(g["dims"]), g["npass"]
    say("[gpu] read_restart fp64 (plain and subrecord framing)")
    q, bf, qp = io_state(*dims, npass=npass, seed=11)
    S = solver(dims, npass)
    for max_sub in (RF.MAX_SUBRECORD, 1000):
        p = str(tmp_path / f"r{max_sub}.dat")
        RF.write_restart(p, q, bf, qp, 0.75, 12, (3, 2, 1), max_sub=max_sub)
        t, d, s = S.read_restart(p)
        assert (t, d, s) == (0.75, 12, (3, 2, 1))
        q2, bf2, qp2 = S.get_state()
        assert np.array_equal(q2, q) and np.array_equal(bf2, bf) and np.array_equal(qp2, qp)
    # a write after the read reproduces the file
    S.write_restart(str(tmp_path / "w.dat"), 0.75, 12, (3, 2, 1))
    assert (tmp_path / "w.dat").read_bytes() == (tmp_path / f"r{RF.MAX_SUBRECORD}.dat").read_bytes()
    S.close()


def test_subrecord_writer(tmp_path, monkeypatch):
    """IMHD_MAX_SUBRECORD lowers the split point: the native writer's subrecords equal the
    restatement's at the same limit (a 1 KB limit splits every array record)."""
    g = GOLD["tvd_roe_o12x10x8"]
    dims, npass = tuple(g["dims"]), g["npass"]
    say("[gpu] subrecord writer")
    q, bf, qp = io_state(*dims, npass=npass, seed=g["seed"])
    S = solver(dims, npass)
    S.set_state(q, bf, qp)
    for lim in (1000, 8, 4097):
        monkeypatch.setenv("IMHD_MAX_SUBRECORD", str(lim))
        S.write_restart(str(tmp_path / "n.dat"), 1.5, 2, (4, 5, 6))
        RF.write_restart(str(tmp_path / "o.dat"), q, bf, qp, 1.5, 2, (4, 5, 6), max_sub=lim)
        assert (tmp_path / "n.dat").read_bytes() == (tmp_path / "o.dat").read_bytes(), lim
    S.close()


def test_read_restart_rejects_other_sizes(tmp_path):
    """A file for another box size or passive count fails with IMHD_EIO before any state
    changes; so do truncated files and trailing data."""
    from isothermal_mhd_amd import ImhdError
    dims, npass = (12, 10, 8), 2
    say("[gpu] read_restart rejects")
    q, bf, qp = io_state(*dims, npass=npass, seed=5)
    S = solver(dims, npass)
    S.set_state(q, bf, qp)
    qo, bo, po = io_state(12, 10, 9, npass=2, seed=6)
    RF.write_restart(str(tmp_path / "big.dat"), qo, bo, po, 1.0, 1, (1, 2, 3))
    qn, bn, _ = io_state(*dims, npass=0, seed=7)
    RF.write_restart(str(tmp_path / "np0.dat"), qn, bn, None, 1.0, 1, (1, 2, 3))
    q3, b3, p3 = io_state(*dims, npass=npass, seed=8)
    RF.write_restart(str(tmp_path / "ok.dat"), q3, b3, p3, 1.0, 1, (1, 2, 3))
    good = (tmp_path / "ok.dat").read_bytes()
    (tmp_path / "short.dat").write_bytes(good[:-3])
    (tmp_path / "long.dat").write_bytes(good + b"\0")
    bad = bytearray(good)
    bad[-4] ^= 1    # last trailing marker
    (tmp_path / "marker.dat").write_bytes(bytes(bad))
    for f in ("big.dat", "np0.dat", "short.dat", "long.dat", "marker.dat", "missing.dat"):
        with pytest.raises(ImhdError, match="IMHD_EIO"):
            S.read_restart(str(tmp_path / f))
        q2, bf2, qp2 = S.get_state()
        assert np.array_equal(q2, q) and np.array_equal(bf2, bf) and np.array_equal(qp2, qp), f
    S.close()


def test_fp32_context_files(tmp_path):
    g = GOLD["vl_hlld_p16"]
    dims, npass = tuple(g["dims"]), g["npass"]
    say("[gpu] fp32 restart/cube")
    q, bf, qp = io_state(*dims, npass=npass, seed=g["seed"])
    S = solver(dims, npass, precision=32)
    S.set_state(q, bf, qp)
    qs, bs, ps = S.get_state()
    with np.errstate(over="ignore"):
        assert np.array_equal(qs, q.astype(np.float32).astype(np.float64))
    S.write_restart(str(tmp_path / "r.dat"), 2.0, 5, (1, 1, 1))
    RF.write_restart(str(tmp_path / "o.dat"), qs, bs, ps[..., :npass], 2.0, 5, (1, 1, 1))
    assert (tmp_path / "r.dat").read_bytes() == (tmp_path / "o.dat").read_bytes()
    for dp in (False, True):
        S.write_cube(str(tmp_path / "c.dat"), double_prec=dp)
        assert (tmp_path / "c.dat").read_bytes() == RF.cube_bytes(qs, ps, dims, 6, dp, 1)
    # read back into a fresh FP32 context: same state
    T = solver(dims, npass, precision=32)
    assert T.read_restart(str(tmp_path / "r.dat")) == (2.0, 5, (1, 1, 1))
    qt, bt, pt = T.get_state()
    assert np.array_equal(qt, qs) and np.array_equal(bt, bs) and np.array_equal(pt, ps)
    S.close()
    T.close()


def test_ensemble_box_and_restart_continues_run(tmp_path):
    """Box 1 of an ensemble round-trips; a run resumed from a restart continues bit for
    bit like the uninterrupted run (FP64, native loop)."""
    from refutil import random_state
    dims = [(8, 8, 8), (12, 10, 6)]
    say("[gpu] ensemble restart + resume")
    S = solver(dims, 1)
    states = [random_state(*d, npass=1, seed=20 + i) for i, d in enumerate(dims)]
    for i, st in enumerate(states):
        S.set_state(*st, box=i)
    S.boundary_conditions()
    S.run(3, graph=False)
    t, dt = S.get_time()
    S.write_restart(str(tmp_path / "b1.dat"), t[1], 3, (1, 2, 3), box=1)
    S.write_restart(str(tmp_path / "b0.dat"), t[0], 3, (1, 2, 3), box=0)
    S.run(4, graph=False)
    ref = [S.get_state(box=i) for i in range(2)]
    T = solver(dims, 1)
    for i in range(2):
        T.read_restart(str(tmp_path / f"b{i}.dat"), box=i)
    T.set_time(t, dt)
    T.run(4, graph=False)
    for i in range(2):
        got = T.get_state(box=i)
        for x, y in zip(ref[i], got):
            assert np.array_equal(x, y), f"box {i}"
    S.close()
    T.close()
