"""Restart and snapshot file formats of the reference (3D_MPI/src/io_mod.f90), restated in
numpy.  TEST INFRASTRUCTURE ONLY: the parity tests use it as the checker of the library's
native writers and readers (imhd_write_restart / imhd_read_restart / imhd_write_cube);
nothing in the product imports it.

Formats
-------
* Restart (write_restart / read_restart, io_mod.f90:440-517): one file per rank, Fortran
  ``FORM='UNFORMATTED', ACCESS='SEQUENTIAL'`` records, in order
  t_sim (REAL*8), q_3d (7 padded fields), bface_3d (3 padded fields),
  qpass_3d (only when npass > 0), dump_count, s1, s2, s3 (default INTEGER, 4 bytes).
  Arrays are written whole, column-major, ghost layers included.
* Record framing: the reference builds with gfortran (3D_MPI/build/Makefile:3).  gfortran
  frames each record with 4-byte little-endian length markers before and after the data and
  splits a record longer than 2147483639 bytes into subrecords of at most that length; the
  leading marker of a subrecord is negated when another subrecord follows it, the trailing
  marker when one precedes it (gfortran manual, "File format of unformatted sequential
  files" and -fmax-subrecord-length).  AMD flang, which builds the reference here, writes
  one record whose 4-byte markers hold the length modulo 2**32; the reader accepts both.
* xyz cube (writetofile_unformatted, io_mod.f90:249-277): ``ACCESS='STREAM'``, no markers,
  interior cells i fastest, per cell q_3d(i,j,k,1:7) then qpass_3d(i,j,k,1) when npass > 0,
  REAL*8 when double_prec, else SNGL (round to nearest).
* Parallel dump (load_parallel_file_buffer + writetofile_unformatted_parallel,
  io_mod.f90:302-436): the same interleave with all npass passives per cell; on one rank
  the file is the buffer at offset 0.

Pinning: files written by the reference's own io_mod (compiled by oracle/ref/build_ref.py)
are compared byte for byte with this module's output (tests/test_restart_format.py), and
the reference reads this module's files back.  Records above 2 GiB cannot be written by a
reference build here (no gfortran in the image); their subrecord framing follows the
gfortran manual and is checked with a reduced subrecord length ("parity unpinned" against
a gfortran-written file).
"""
import struct

import numpy as np

MAX_SUBRECORD = 2147483639   # gfortran's default and largest subrecord length (2**31 - 9)


def record_markers(length, max_sub=MAX_SUBRECORD):
    """(lead, size, trail) of each subrecord of a record of ``length`` bytes."""
    out = []
    done = 0
    first = True
    while True:
        n = min(max_sub, length - done)
        last = done + n == length
        out.append((n if last else -n, n, n if first else -n))
        done += n
        first = False
        if last:
            return out


def write_record(f, payload, max_sub=MAX_SUBRECORD):
    """One sequential record of the bytes ``payload`` (bytes or a contiguous array)."""
    mv = memoryview(payload).cast("B")
    o = 0
    for lead, n, trail in record_markers(len(mv), max_sub):
        f.write(struct.pack("<i", lead))
        f.write(mv[o:o + n])
        f.write(struct.pack("<i", trail))
        o += n


def read_record(f, length):
    """The next record, which must hold ``length`` bytes, in either framing."""
    pos = f.tell()
    lead = struct.unpack("<i", f.read(4))[0]
    if lead == length or (lead & 0xFFFFFFFF) == (length & 0xFFFFFFFF):
        data = f.read(length)
        trail = struct.unpack("<i", f.read(4))[0]
        if len(data) == length and trail == lead:
            return data
        if lead >= 0:
            raise ValueError(f"record at {pos}: bad trailing marker")
    # gfortran subrecords
    f.seek(pos)
    parts = []
    first = True
    got = 0
    while True:
        lead = struct.unpack("<i", f.read(4))[0]
        n = abs(lead)
        parts.append(f.read(n))
        trail = struct.unpack("<i", f.read(4))[0]
        if trail != (n if first else -n) or len(parts[-1]) != n:
            raise ValueError(f"record at {pos}: bad subrecord markers")
        got += n
        first = False
        if lead >= 0:
            break
    if got != length:
        raise ValueError(f"record at {pos}: {got} bytes, expected {length}")
    return b"".join(parts)


def padded(n, nb):
    return tuple(x + 2 * nb for x in n)


def pass_padded(n, nb):
    return tuple(x + 2 * nb - 4 for x in n)


def fbytes(a):
    """The bytes of a Fortran array: column-major, whatever the numpy order."""
    return np.asarray(a, "<f8").tobytes(order="F")


def write_restart(path, q, bf, qp, t_sim, dump_count, seeds, max_sub=MAX_SUBRECORD):
    """write_restart (io_mod.f90:459-468).  q (Px,Py,Pz,7), bf (Px,Py,Pz,3) and qp
    (PPx,PPy,PPz,npass) in the reference's index order; qp None or npass 0 omits the
    passive record (io_mod.f90:464)."""
    with open(path, "wb") as f:
        write_record(f, np.array([t_sim], "<f8").tobytes(), max_sub)
        write_record(f, fbytes(q), max_sub)
        write_record(f, fbytes(bf), max_sub)
        if qp is not None and qp.shape[-1] > 0:
            write_record(f, fbytes(qp), max_sub)
        for v in (dump_count,) + tuple(seeds):
            write_record(f, np.array([v], "<i4").tobytes(), max_sub)


def read_restart(path, n, nb, npass):
    """read_restart (io_mod.f90:499-508) -> (t_sim, q, bf, qp, dump_count, seeds), arrays
    Fortran-ordered in the reference's index order."""
    P = padded(n, nb)
    PP = pass_padded(n, nb)
    nc = P[0] * P[1] * P[2]

    def arr(buf, shape):
        return np.frombuffer(buf, "<f8").reshape(shape, order="F")

    with open(path, "rb") as f:
        t = np.frombuffer(read_record(f, 8), "<f8")[0]
        q = arr(read_record(f, 7 * nc * 8), P + (7,))
        bf = arr(read_record(f, 3 * nc * 8), P + (3,))
        qp = None
        if npass > 0:
            qp = arr(read_record(f, npass * PP[0] * PP[1] * PP[2] * 8), PP + (npass,))
        ints = [int(np.frombuffer(read_record(f, 4), "<i4")[0]) for _ in range(4)]
        if f.read(1):
            raise ValueError("data after the last record")
    return float(t), q, bf, qp, ints[0], tuple(ints[1:])


def cube_bytes(q, qp, n, nb, double_prec, npass_out):
    """Per-cell interleave of the interior, i fastest: written by writetofile_unformatted
    with npass_out = min(npass, 1) (io_mod.f90:253-273) and by the parallel dump with
    npass_out = npass (io_mod.f90:378-431).  REAL*8 when double_prec, else SNGL."""
    nx, ny, nz = n
    cols = [q[nb:nb + nx, nb:nb + ny, nb:nb + nz, f] for f in range(7)]
    cols += [qp[nb - 2:nb - 2 + nx, nb - 2:nb - 2 + ny, nb - 2:nb - 2 + nz, p] for p in range(npass_out)]
    a = np.stack(cols, axis=0)          # (7 + npass_out, nx, ny, nz): value fastest in F order
    return a.astype("<f8" if double_prec else "<f4").tobytes(order="F")


def write_cube(path, q, qp, n, nb, double_prec, npass_out):
    with open(path, "wb") as f:
        f.write(cube_bytes(q, qp, n, nb, double_prec, npass_out))


def dump_label(t_dump):
    """The dump number as io_mod.f90:28-38 formats it (I1.1 .. I5.3: no padding below
    100000; larger numbers leave the label unset)."""
    if not 0 <= t_dump < 100000:
        raise ValueError("dump numbers 0..99999 only")
    return str(t_dump)


def restart_name(coord=(0, 0, 0)):
    return "Restart/restart_mhd_rank=%d%d%d.dat" % tuple(coord)


def cube_name(t_dump, coord=(0, 0, 0)):
    return "Snapshots/xyzcube_rank=%d%d%d_dump=%s.dat" % (tuple(coord) + (dump_label(t_dump),))
