"""Restart and snapshot file formats (io_mod.f90:249-517): the numpy restatement in
oracle/restart_format.py against files written by the reference's own io_mod.

* Golden digests (tests/golden/restart_vectors.json, from make_restart_golden.py): the
  oracle's bytes for the recorded inputs hash to what the reference wrote.  Runs anywhere.
* Live reference (oracle/_ref builds, where /root/reference was present): byte equality of
  restart file, xyz cube and parallel-I/O buffer, and the reference's read_restart reading
  an oracle-written file with subrecord-free framing.
* Record framing: a hand-built gfortran subrecord sequence, split and re-joined.
"""
import hashlib
import io
import json
import os
import struct
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "oracle"))
import restart_format as RF  # noqa: E402
from refutil import Reference, io_state, ref_available  # noqa: E402

GOLD = json.load(open(os.path.join(HERE, "golden", "restart_vectors.json")))


def sha(b):
    return hashlib.sha256(b).hexdigest()


def oracle_bytes(g):
    """The case's inputs, regenerated from its seed."""
    dims, npass = tuple(g["dims"]), g["npass"]
    q, bf, qp = io_state(*dims, npass=npass, seed=g["seed"])
    return q, bf, qp, dims, npass


def restart_bytes(q, bf, qp, g, tmp_path, max_sub=RF.MAX_SUBRECORD):
    p = tmp_path / "r.dat"
    RF.write_restart(str(p), q, bf, qp, g["t_sim"], g["dump_count"], g["seeds"], max_sub=max_sub)
    return p.read_bytes()


@pytest.mark.parametrize("name", sorted(GOLD))
def test_oracle_matches_reference_digests(name, tmp_path):
    g = GOLD[name]
    q, bf, qp, dims, npass = oracle_bytes(g)
    rb = restart_bytes(q, bf, qp, g, tmp_path)
    assert len(rb) == g["restart_bytes"]
    assert sha(rb) == g["restart_sha256"]
    with np.errstate(over="ignore"):
        cube = RF.cube_bytes(q, qp, dims, 6, g["double_prec"], min(npass, 1))
        par = RF.cube_bytes(q, qp, dims, 6, g["double_prec"], npass)
    assert len(cube) == g["cube_bytes"]
    assert sha(cube) == g["cube_sha256"]
    assert sha(par) == g["parallel_sha256"]


def test_restart_round_trip_and_layout(tmp_path):
    g = GOLD["vl_hlld_p12x10x8"]
    q, bf, qp, dims, npass = oracle_bytes(g)
    rb = restart_bytes(q, bf, qp, g, tmp_path)
    # first record: 8-byte t_sim between two markers of 8
    assert struct.unpack("<i", rb[:4])[0] == 8 and struct.unpack("<d", rb[4:12])[0] == g["t_sim"]
    # q_3d record holds the array column-major: its first value is q(1-nb,1-nb,1-nb,1)
    n7 = struct.unpack("<i", rb[16:20])[0]
    assert n7 == 7 * 24 * 22 * 20 * 8
    assert struct.unpack("<d", rb[20:28])[0] == q[0, 0, 0, 0]
    assert struct.unpack("<d", rb[28:36])[0] == q[1, 0, 0, 0]
    t, q2, bf2, qp2, d, s = RF.read_restart(str(tmp_path / "r.dat"), dims, 6, npass)
    assert t == g["t_sim"] and d == g["dump_count"] and s == tuple(g["seeds"])
    assert np.array_equal(q2, q) and np.array_equal(bf2, bf) and np.array_equal(qp2, qp)


def test_subrecord_framing_known_answer():
    """A 10-byte record with a 4-byte subrecord limit: lead markers negated while another
    subrecord follows, trail markers negated once one precedes (gfortran's rule)."""
    f = io.BytesIO()
    RF.write_record(f, bytes(range(10)), max_sub=4)
    want = (struct.pack("<i", -4) + bytes([0, 1, 2, 3]) + struct.pack("<i", 4) +
            struct.pack("<i", -4) + bytes([4, 5, 6, 7]) + struct.pack("<i", -4) +
            struct.pack("<i", 2) + bytes([8, 9]) + struct.pack("<i", -2))
    assert f.getvalue() == want
    f.seek(0)
    assert RF.read_record(f, 10) == bytes(range(10))
    # a record that fills its subrecord exactly is not split
    f = io.BytesIO()
    RF.write_record(f, bytes(8), max_sub=8)
    assert f.getvalue() == struct.pack("<i", 8) + bytes(8) + struct.pack("<i", 8)


def test_subrecord_restart_reads_back(tmp_path):
    g = GOLD["tvd_roe_o12x10x8"]
    q, bf, qp, dims, npass = oracle_bytes(g)
    rb = restart_bytes(q, bf, qp, g, tmp_path, max_sub=1000)
    assert len(rb) > g["restart_bytes"]   # extra markers
    t, q2, bf2, qp2, d, s = RF.read_restart(str(tmp_path / "r.dat"), dims, 6, npass)
    assert np.array_equal(q2, q) and np.array_equal(qp2, qp) and s == tuple(g["seeds"])


def test_reader_rejects_mismatched_sizes(tmp_path):
    g = GOLD["vl_hlld_p12x10x8"]
    q, bf, qp, dims, npass = oracle_bytes(g)
    restart_bytes(q, bf, qp, g, tmp_path)
    with pytest.raises(ValueError):
        RF.read_restart(str(tmp_path / "r.dat"), (12, 10, 9), 6, npass)
    with pytest.raises(ValueError):
        RF.read_restart(str(tmp_path / "r.dat"), dims, 6, 1)


def test_dump_labels():
    assert RF.cube_name(7) == "Snapshots/xyzcube_rank=000_dump=7.dat"
    assert RF.cube_name(1234) == "Snapshots/xyzcube_rank=000_dump=1234.dat"
    assert RF.restart_name((1, 0, 2)) == "Restart/restart_mhd_rank=102.dat"


def _run_reference_writers(R, g, tmp_path, monkeypatch):
    (tmp_path / "Output" / "Restart").mkdir(parents=True)
    (tmp_path / "Output" / "Snapshots").mkdir(parents=True)
    monkeypatch.chdir(tmp_path)
    R.write_restart(g["t_sim"], g["dump_count"], g["seeds"])
    R.write_snapshot(g["dump_count"])
    rb = (tmp_path / "Output" / "Restart" / "restart_mhd_rank=000.dat").read_bytes()
    cb = (tmp_path / "Output" / "Snapshots" / f"xyzcube_rank=000_dump={g['dump_count']}.dat").read_bytes()
    return rb, cb


@pytest.mark.parametrize("name", sorted(GOLD))
def test_reference_io_mod_byte_equal(name, tmp_path, monkeypatch):
    if not ref_available(name):
        pytest.skip("reference build not present")
    g = GOLD[name]
    q, bf, qp, dims, npass = oracle_bytes(g)
    R = Reference(name)
    R.set_state(q, bf, qp)
    rb, cb = _run_reference_writers(R, g, tmp_path, monkeypatch)
    assert rb == restart_bytes(q, bf, qp, g, tmp_path)
    with np.errstate(over="ignore"):
        assert cb == RF.cube_bytes(q, qp, dims, 6, g["double_prec"], min(npass, 1))
        par = np.frombuffer(RF.cube_bytes(q, qp, dims, 6, g["double_prec"], npass),
                            "<f8" if g["double_prec"] else "<f4").astype(np.float64)
    assert np.array_equal(R.parallel_buffer(), par)


@pytest.mark.parametrize("name", ["vl_hlld_p12x10x8", "vl_hlle_o8x6x4_np0"])
def test_reference_reads_oracle_restart(name, tmp_path, monkeypatch):
    if not ref_available(name):
        pytest.skip("reference build not present")
    g = GOLD[name]
    dims, npass = tuple(g["dims"]), g["npass"]
    q, bf, qp = io_state(*dims, npass=npass, seed=g["seed"] + 100)
    (tmp_path / "Output" / "Restart").mkdir(parents=True)
    RF.write_restart(str(tmp_path / "Output" / "Restart" / "restart_mhd_rank=000.dat"), q, bf, qp,
                     -3.25, 4321, (9, 8, 7))
    monkeypatch.chdir(tmp_path)
    R = Reference(name)
    t, d, s = R.read_restart()
    assert (t, d, s) == (-3.25, 4321, (9, 8, 7))
    q2, bf2, qp2, _ = R.get_state()
    assert np.array_equal(q2, q) and np.array_equal(bf2, bf)
    if npass:
        assert np.array_equal(qp2[..., :npass], qp)
