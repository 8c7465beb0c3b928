"""Power spectra and density PDF of the reference's analysis program
(3D_MPI/Analysis/src/analysis_main.f90), restated in numpy.  TEST INFRASTRUCTURE ONLY: the
GPU diagnostics (imhd_spectrum, imhd_density_pdf) are checked against it; nothing in the
product imports it.

The analysis program reads the REAL*4 snapshot cube and works in single precision:
* velocity v = (rho v) / rho in REAL*4 (analysis_main.f90:105-113);
* v_par = (v.B / |B|) B / |B| and v_perp = v - v_par, evaluated as written at :176-190 and
  :211-226, in REAL*4;
* compute_pk (:543-633): the field's 3-D DFT over the n^3 interior (unnormalised; the
  sign of the exponent does not matter for a real field), kbins = n bins of width
  dk = sqrt(3) k0 (n/2) / n with k0 = 2 pi (box length 1).  Only the octant
  0 <= kx, ky, kz < n/2 is binned, with weight 2 (the zero mode weight 1);
  k = k0 sqrt(SNGL(kx^2 + ky^2 + kz^2)) and the bin 1 + INT(k / dk) in REAL*4; the sum of
  |F|^2 over the components is formed in REAL*4 left to right, then doubled; Pk is
  multiplied by dx^3 (dx = 1/n).  The spectrum file holds ((ik-1) dk, Pk(ik)).
* compute_density_pdf (:637-698): n bins of width 4/n on [0, 4), bin 1 + INT(rho / drho),
  normalised by drho * (number of counted cells).  The reference writes out of bounds for
  rho >= 4; these cells are not counted here.

The DFT here is numpy's (FP64) of the REAL*4 field; the reference uses its own REAL*4 radix-2
FFT (fft_mod.f90:160-261), so parity is to FP32 rounding.  The binning rules are pinned by
single-mode fields whose whole power lands in one known bin (tests/test_spectra.py).
"""
import numpy as np

FIELDS = {"density": 0, "velocity": 1, "bfield": 2, "vpar": 3, "vperp": 4}


def interior_f32(q, n, nb=6):
    """The seven cell fields on the interior as REAL*4 (SNGL of the snapshot cube)."""
    nx, ny, nz = n
    return [np.asarray(q[nb:nb + nx, nb:nb + ny, nb:nb + nz, f], dtype=np.float32) for f in range(7)]


def field_components(q, n, which, nb=6):
    fx = interior_f32(q, n, nb)
    rho = fx[0]
    v = [fx[1] / rho, fx[2] / rho, fx[3] / rho]
    b = fx[4:7]
    if which == 0:
        return [rho]
    if which == 1:
        return v
    if which == 2:
        return list(b)
    bmag = np.sqrt(b[0] ** 2 + b[1] ** 2 + b[2] ** 2)
    bhat = [b[0] / bmag, b[1] / bmag, b[2] / bmag]
    vdotb = (v[0] * b[0] + v[1] * b[1] + v[2] * b[2]) / bmag
    if which == 3:
        return [vdotb * bhat[d] for d in range(3)]
    return [v[d] - vdotb * bhat[d] for d in range(3)]


def spectrum(q, n, which, nb=6):
    """compute_pk for one field kind -> (k, Pk), arrays of length n (FP64)."""
    nx, ny, nz = n
    assert nx == ny == nz, "the analysis program assumes a cubic box"
    nm = nx
    h = nm // 2
    s = np.zeros((h, h, h), dtype=np.float32)
    for comp in field_components(q, n, which, nb):
        F = np.fft.fftn(comp.astype(np.float64))[:h, :h, :h]     # [kx, ky, kz] of the x-fastest field
        re = F.real.astype(np.float32)
        im = F.imag.astype(np.float32)
        s = s + re * re
        s = s + im * im
    fk = 2.0 * s.astype(np.float64)
    fk[0, 0, 0] *= 0.5
    k0 = np.float32(2.0 * np.pi)                                 # REAL*4 TWOPI / Lx, Lx = 1
    kmax = np.float32(np.sqrt(3.0) * float(k0) * h)              # SQRT(3.d0)*k0*(nmax/2) -> REAL*4
    dk = np.float32(kmax / np.float32(nm))
    kk = np.arange(h)
    ksq = (kk[:, None, None] ** 2 + kk[None, :, None] ** 2 + kk[None, None, :] ** 2).astype(np.float32)
    k = k0 * np.sqrt(ksq)
    ik = (k / dk).astype(np.int64)        # INT truncates toward zero; k >= 0
    pk = np.bincount(ik.ravel(), weights=fk.astype(np.float32).astype(np.float64).ravel(), minlength=nm)[:nm]
    dx = 1.0 / nm
    return np.arange(nm) * float(dk), pk * dx ** 3


def density_pdf(q, n, nbins=None, rho_max=4.0, nb=6):
    """compute_density_pdf -> (bin centres, PDF)."""
    rho = interior_f32(q, n, nb)[0]
    nbins = nbins or min(n)
    drho = np.float32(np.float32(rho_max) / np.float32(nbins))
    ib = (rho / drho).astype(np.int64)
    ok = (ib >= 0) & (ib < nbins)
    cnt = np.bincount(ib[ok].ravel(), minlength=nbins)[:nbins].astype(np.float64)
    pdf = cnt / (float(drho) * cnt.sum())
    return (np.arange(nbins) + 0.5) * float(drho), pdf
