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


def mode_spectra(q, n, nb=6):
    """mode_decomposition (analysis_main.f90:305-540): Alfven / fast / slow spectra after
    Cho & Lazarian (2003).  As the analysis program calls it (:249), the velocity
    transforms it projects are those of v_perp (the last ones it computed, :238-247).
    B0 is the mean cell-centred field.  Per octant mode k != 0 (REAL*4, double constants
    promoted as written): displacement vectors zeta_a = khat_perp x khat_par,
    zeta_f,s ~ (-1 + alpha +- sqrt D) k_par + (1 + alpha +- sqrt D) k_perp with
    alpha = 1/|B0|^2, D = max(0, (1+alpha)^2 - 4 alpha |k_par|^2/|k|^2); velocity, field and
    density amplitudes and their doubled squares binned as in compute_pk.  Returns
    (k, P) with P of shape (8, n): v_a, v_f, v_s, b_a, b_f, b_s, rho_f, rho_s."""
    nx, ny, nz = n
    assert nx == ny == nz
    nm = nx
    h = nm // 2
    comps = field_components(q, n, 4, nb)
    F = [np.fft.fftn(c.astype(np.float64))[:h, :h, :h] for c in comps]
    vre = [f.real for f in F]
    vim = [f.imag for f in F]
    b = interior_f32(q, n, nb)[4:7]
    B0 = np.array([float(np.sum(b[d].astype(np.float64))) / nm ** 3 for d in range(3)])
    bmag = np.sqrt(np.sum(B0 ** 2))
    alpha = 1.0 / np.sum(B0 ** 2)
    k0 = float(np.float32(2.0 * np.pi))
    kk = np.arange(h)
    KX, KY, KZ = np.meshgrid(kk, kk, kk, indexing="ij")
    k = k0 * np.sqrt((KX ** 2 + KY ** 2 + KZ ** 2).astype(np.float64))
    kv = [k0 * KX, k0 * KY, k0 * KZ]
    ksafe = np.where(k > 0, k, 1.0)
    khat = [c / ksafe for c in kv]
    ph = B0 / bmag
    kdot = kv[0] * ph[0] + kv[1] * ph[1] + kv[2] * ph[2]
    kpar = [kdot * ph[d] for d in range(3)]
    kperp = [kv[d] - kpar[d] for d in range(3)]
    kpn = np.sqrt(kperp[0] ** 2 + kperp[1] ** 2 + kperp[2] ** 2)
    kph = [np.where(kpn > 0, kperp[d] / np.where(kpn > 0, kpn, 1.0), 0.0) for d in range(3)]
    D = np.maximum(0.0, (1.0 + alpha) ** 2 - 4.0 * alpha * (kpar[0] ** 2 + kpar[1] ** 2 + kpar[2] ** 2) / ksafe ** 2)
    sD = np.sqrt(D)
    za = [kph[1] * ph[2] - kph[2] * ph[1], kph[2] * ph[0] - kph[0] * ph[2], kph[0] * ph[1] - kph[1] * ph[0]]

    def unit(z):
        nz_ = np.sqrt(z[0] ** 2 + z[1] ** 2 + z[2] ** 2)
        return [np.where(nz_ > 0, z[d] / np.where(nz_ > 0, nz_, 1.0), z[d]) for d in range(3)]

    zf = unit([(-1.0 + alpha + sD) * kpar[d] + (1.0 + alpha + sD) * kperp[d] for d in range(3)])
    zs = unit([(-1.0 + alpha - sD) * kpar[d] + (1.0 + alpha - sD) * kperp[d] for d in range(3)])

    def proj(z):
        return (vre[0] * z[0] + vre[1] * z[1] + vre[2] * z[2], vim[0] * z[0] + vim[1] * z[1] + vim[2] * z[2])

    va, vf, vs = proj(za), proj(zf), proj(zs)
    cf = np.sqrt(0.5 * bmag ** 2 * (1.0 + alpha + sD))
    csl = np.sqrt(np.maximum(0.0, 0.5 * bmag ** 2 * (1.0 + alpha - sD)))
    t1 = khat[0] * zf[0] + khat[1] * zf[1] + khat[2] * zf[2]
    t2 = khat[0] * zs[0] + khat[1] * zs[1] + khat[2] * zs[2]
    dvf, dvs = 2.0 * vf[1], 2.0 * vs[1]
    with np.errstate(divide="ignore", invalid="ignore"):
        rhof = dvf * t1 / cf
        rhos = dvs * t2 / csl

        def cross_norm(z):
            c = [B0[1] * z[2] - B0[2] * z[1], B0[2] * z[0] - B0[0] * z[2], B0[0] * z[1] - B0[1] * z[0]]
            return np.sqrt(c[0] ** 2 + c[1] ** 2 + c[2] ** 2)

        bf_ = dvf * cross_norm(zf) / cf
        bs_ = dvs * cross_norm(zs) / csl
    amps = [2 * (va[0] ** 2 + va[1] ** 2), 2 * (vf[0] ** 2 + vf[1] ** 2), 2 * (vs[0] ** 2 + vs[1] ** 2),
            2 * (va[0] ** 2 + va[1] ** 2), 2 * bf_ ** 2, 2 * bs_ ** 2, 2 * rhof ** 2, 2 * rhos ** 2]
    kmax = np.float32(np.sqrt(3.0) * k0 * h)
    dk = float(np.float32(kmax / np.float32(nm)))
    kf = np.float32(k0) * np.sqrt((KX ** 2 + KY ** 2 + KZ ** 2).astype(np.float32))
    ik = (kf / np.float32(dk)).astype(np.int64)
    live = (KX + KY + KZ) > 0
    P = np.zeros((8, nm))
    for m, a in enumerate(amps):
        w = np.where(live, a, 0.0)
        w = np.where(np.isfinite(w), w, 0.0)
        P[m] = np.bincount(ik.ravel(), weights=w.ravel(), minlength=nm)[:nm]
    return np.arange(nm) * dk, P / nm ** 3
