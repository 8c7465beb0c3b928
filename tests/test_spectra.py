"""The restatement of the analysis program's spectra and density PDF (oracle/spectra.py;
3D_MPI/Analysis/src/analysis_main.f90:543-698), pinned by fields whose answer is known in
closed form: a single Fourier mode puts all its power into one shell, in the bin
INT(k / dk) that compute_pk's REAL*4 geometry gives (dk = sqrt(3) pi for unit box length),
with the weight 2 |F|^2 dx^3 = A^2 n^3 / 2 for amplitude A; the mean goes to bin 0 with the
zero mode's single weight; a two-valued density fills exactly two PDF bins."""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "oracle"))
import spectra as SP  # noqa: E402

NB = 6


def cube(n, rho, v=(0, 0, 0), b=(0, 0, 0)):
    P = n + 2 * NB
    q = np.zeros((P, P, P, 7), order="F")
    I = (slice(NB, NB + n),) * 3
    q[I + (0,)] = rho
    for d in range(3):
        q[I + (1 + d,)] = rho * v[d]
        q[I + (4 + d,)] = b[d]
    return q


def grid(n):
    x = (np.arange(n) + 0.5) / n
    return np.meshgrid(x, x, x, indexing="ij")


@pytest.mark.parametrize("n,kx0", [(16, 3), (32, 5), (24, 7)])
def test_single_density_mode(n, kx0):
    X, Y, Z = grid(n)
    A = 0.25
    q = cube(n, 1.0 + A * np.cos(2 * np.pi * kx0 * X))
    k, pk = SP.spectrum(q, (n, n, n), 0)
    ik = int(2 * kx0 / np.sqrt(3.0))
    assert ik == int(np.float32(np.float32(2 * np.pi) * np.float32(kx0)) / np.float32(np.sqrt(3.0) * np.float32(2 * np.pi) * (n // 2) / n))
    assert pk[0] == pytest.approx(n ** 3, rel=1e-6)                  # mean 1: |F(0)|^2 dx^3
    assert pk[ik] == pytest.approx(A ** 2 * n ** 3 / 2, rel=1e-5)
    rest = np.delete(pk, [0, ik])
    assert np.max(np.abs(rest)) <= 1e-6 * pk[ik]
    assert k[ik] == pytest.approx(ik * np.sqrt(3.0) * np.pi, rel=1e-6)


def test_velocity_and_b_modes():
    n = 16
    X, Y, Z = grid(n)
    A = 0.5
    # v_x(y) = A sin(2 pi 2 y), rho = 1: all power in the shell of (0, 2, 0); B along z is uniform
    q = cube(n, 1.0, v=(A * np.sin(2 * np.pi * 2 * Y), 0 * X, 0 * X), b=(0 * X, 0 * X, 1.0 + 0 * X))
    k, pv = SP.spectrum(q, (n, n, n), 1)
    ik = int(2 * 2 / np.sqrt(3.0))
    assert pv[ik] == pytest.approx(A ** 2 * n ** 3 / 2, rel=1e-5)
    assert np.max(np.abs(np.delete(pv, ik))) <= 1e-6 * pv[ik]
    _, pb = SP.spectrum(q, (n, n, n), 2)
    assert pb[0] == pytest.approx(n ** 3, rel=1e-6) and np.max(np.abs(pb[1:])) <= 1e-6 * pb[0]
    # v perpendicular to B = z: v_par = 0, v_perp = v
    _, ppar = SP.spectrum(q, (n, n, n), 3)
    _, pperp = SP.spectrum(q, (n, n, n), 4)
    assert np.max(np.abs(ppar)) <= 1e-9 * pv[ik]
    assert np.allclose(pperp, pv, rtol=1e-6, atol=1e-9 * pv[ik])


def test_density_pdf_two_values():
    n = 8
    X, Y, Z = grid(n)
    rho = np.where(X < 0.5, 0.5, 1.5)
    q = cube(n, rho)
    r, pdf = SP.density_pdf(q, (n, n, n))
    drho = 4.0 / n
    i1, i2 = int(0.5 / drho), int(1.5 / drho)
    assert pdf[i1] == pytest.approx(0.5 / drho) and pdf[i2] == pytest.approx(0.5 / drho)
    assert np.count_nonzero(pdf) == 2
    assert r[i1] == pytest.approx((i1 + 0.5) * drho)


def test_mode_decomposition_single_modes():
    """B0 along z, k along x (k_par = 0).  A velocity along y (normal to k and B0) is a pure
    Alfven mode: its whole power A^2 n^3 / 2 is in the Alfven velocity and field spectra
    of shell INT(2 kx0 / sqrt 3), nothing in fast or slow.  A velocity along x (along k)
    is a pure fast mode.  (The v_perp input is v itself here, v being normal to B0.)"""
    n, kx0, A = 16, 3, 0.2
    X, Y, Z = grid(n)
    s = A * np.sin(2 * np.pi * kx0 * X)
    ik = int(2 * kx0 / np.sqrt(3.0))
    b = (0 * X, 0 * X, 1.0 + 0 * X)
    q = cube(n, 1.0, v=(0 * X, s, 0 * X), b=b)
    k, P = SP.mode_spectra(q, (n, n, n))
    want = A ** 2 * n ** 3 / 2
    assert P[0, ik] == pytest.approx(want, rel=1e-5) and P[3, ik] == pytest.approx(want, rel=1e-5)
    for m in (1, 2, 4, 5, 6, 7):
        assert np.max(np.abs(P[m])) <= 1e-8 * want
    # fast mode: a sine on the index grid x = i/n has a purely imaginary coefficient
    # -i A n^3/2.  At k_par = 0, c_f = |B0| sqrt(1 + alpha) = sqrt(2) and zeta_f = khat, so the
    # density amplitude 2 Im(v_f) / c_f and the field amplitude 2 Im(v_f) |B0 x zeta_f| / c_f
    # both have doubled squares of 2 A^2 n^3 / 2 = 2 want.
    xi = np.arange(n)[:, None, None] / n + 0 * X
    s0 = A * np.sin(2 * np.pi * kx0 * xi)
    q = cube(n, 1.0, v=(s0, 0 * X, 0 * X), b=b)
    k, P = SP.mode_spectra(q, (n, n, n))
    assert P[1, ik] == pytest.approx(want, rel=1e-5)
    for m in (0, 2, 3, 5, 7):
        assert np.max(np.abs(P[m])) <= 1e-8 * want
    assert P[6, ik] == pytest.approx(2 * want, rel=1e-5)
    assert P[4, ik] == pytest.approx(2 * want, rel=1e-5)
