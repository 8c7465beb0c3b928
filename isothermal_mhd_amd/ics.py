"""Seeded initial conditions for the benchmark workloads (the reference has only
hard-coded test problems, init_domain_mod.f90; these random-field ICs are new).

Velocity: Gaussian random Fourier modes with shell spectrum E(k) ~ k^-p on
kmin <= |k| <= kmax (per-mode power ~ k^-(p+2)), optionally projected solenoidal, scaled
to a target rms Mach number v_rms / c_s.  Random numbers come from numpy's PCG64 keyed by
the seed, so every host produces the same field.  Output is the reference host layout:
q_3d(7), bface_3d(3), qpass_3d with nb ghost layers (ghosts left for the ghost fill).
"""
import numpy as np

from .solver import padded_shapes


def random_velocity(n, kmin, kmax, slope, mach, cs=1.0, solenoidal=True, seed=0, dtype=np.float64):
    nx, ny, nz = n
    rng = np.random.default_rng(seed)
    kx = np.fft.fftfreq(nx) * nx
    ky = np.fft.fftfreq(ny) * ny
    kz = np.fft.rfftfreq(nz) * nz
    KX, KY, KZ = np.meshgrid(kx, ky, kz, indexing="ij")
    k2 = KX ** 2 + KY ** 2 + KZ ** 2
    k = np.sqrt(k2)
    shell = (k >= kmin) & (k <= kmax)
    amp = np.zeros_like(k)
    amp[shell] = k[shell] ** (-(slope + 2.0) / 2.0)
    vh = []
    for _ in range(3):
        re = rng.standard_normal(k.shape)
        im = rng.standard_normal(k.shape)
        vh.append((re + 1j * im) * amp)
    if solenoidal:
        kdotv = KX * vh[0] + KY * vh[1] + KZ * vh[2]
        with np.errstate(invalid="ignore", divide="ignore"):
            f = np.where(k2 > 0, kdotv / k2, 0.0)
        vh = [vh[0] - KX * f, vh[1] - KY * f, vh[2] - KZ * f]
    v = [np.fft.irfftn(c, s=(nx, ny, nz), axes=(0, 1, 2)) for c in vh]
    rms = np.sqrt(np.mean(v[0] ** 2 + v[1] ** 2 + v[2] ** 2))
    scale = mach * cs / rms if rms > 0 else 0.0
    return [(c * scale).astype(dtype) for c in v]


def make_state(n, nb=6, npass=0, rho=None, v=None, b0=(0.0, 0.0, 1.0), passive=None):
    nx, ny, nz = n
    sh, shp = padded_shapes(nx, ny, nz, nb, npass)
    q = np.zeros(sh + (7,), order="F")
    bf = np.zeros(sh + (3,), order="F")
    qp = np.zeros(shp, order="F")
    I = (slice(nb, nb + nx), slice(nb, nb + ny), slice(nb, nb + nz))
    r = np.ones((nx, ny, nz)) if rho is None else rho
    q[I + (0,)] = r
    if v is not None:
        for m in range(3):
            q[I + (1 + m,)] = r * v[m]
    for m in range(3):
        q[I + (4 + m,)] = b0[m]
        bf[I + (m,)] = b0[m]
    Ip = (slice(nb - 2, nb - 2 + nx), slice(nb - 2, nb - 2 + ny), slice(nb - 2, nb - 2 + nz))
    for p in range(npass):
        qp[Ip + (p,)] = passive[p]
    return q, bf, qp


def config_c2(n=256, seed=2002, nb=6):
    """256^3 periodic, rho = 1, B0 = (0,0,1), solenoidal k^-2 velocity on 1 <= |k| <= 8 at
    rms Mach 2, one passive scalar of uniform [0,1) seeded noise (HLLD + VL + CT)."""
    v = random_velocity((n, n, n), 1.0, 8.0, 2.0, 2.0, solenoidal=True, seed=seed)
    s = np.random.default_rng(seed + 1).random((n, n, n))
    return make_state((n, n, n), nb=nb, npass=1, v=v, b0=(0.0, 0.0, 1.0), passive=[s])


def config_c3(n=512, seed=3003, nb=6):
    """Forced-turbulence TVD configuration: rho = 1, B0 = (0.2,0,0), v = 0 (the OU forcing
    drives the flow), two passive scalars: sin(2 pi x) and a step at x = 1/2 whose position
    is jittered per (y, z) column by up to +-2 cells (seeded)."""
    nx = ny = nz = n
    x = (np.arange(nx) + 0.5) / nx
    s1 = np.broadcast_to(np.sin(2.0 * np.pi * x)[:, None, None], (nx, ny, nz))
    jit = np.random.default_rng(seed).integers(-2, 3, size=(ny, nz))
    edge = nx // 2 + jit
    s2 = (np.arange(nx)[:, None, None] >= edge[None, :, :]).astype(np.float64)
    return make_state((n, n, n), nb=nb, npass=2, v=None, b0=(0.2, 0.0, 0.0), passive=[s1, s2])


def config_c1(n=64, seed=1001, nb=6):
    """64^3 periodic, rho = 1, B0 = (0.5,0,0), k^-2 velocity on 1 <= |k| <= 4 at rms Mach 1."""
    v = random_velocity((n, n, n), 1.0, 4.0, 2.0, 1.0, solenoidal=False, seed=seed)
    return make_state((n, n, n), nb=nb, npass=0, v=v, b0=(0.5, 0.0, 0.0))


def ensemble_c4(nbox=48, seed=4004, lo=32, hi=192, nb=6):
    """Uneven boxes: sizes ~ U{lo..hi}^3, |B0| ~ U[0.1,1] with a random direction, rms Mach
    ~ U[0.5,4], k^-2 velocity (per box substream of the seed)."""
    rng = np.random.default_rng(seed)
    dims, states, machs = [], [], []
    for b in range(nbox):
        n = tuple(int(x) for x in rng.integers(lo, hi + 1, size=3))
        bm = rng.uniform(0.1, 1.0)
        d = rng.standard_normal(3)
        d /= np.linalg.norm(d)
        mach = rng.uniform(0.5, 4.0)
        v = random_velocity(n, 1.0, 4.0, 2.0, mach, solenoidal=False, seed=seed * 1000 + b)
        dims.append(n)
        machs.append(mach)
        states.append(make_state(n, nb=nb, npass=0, v=v, b0=tuple(bm * d)))
    return dims, states, machs


def config_c5(n=1024, seed=5005, nb=6):
    """rho = 1 with seeded +-10 % lognormal noise (rho = exp(0.1 g - 0.005), g standard
    normal per cell), B0 = (0.05, 0.05, 0), k^-5/3 velocity on 1 <= |k| <= 16 at rms Mach 3
    (the 1024^3 configuration; smaller n where an FP64 run must fit beside it)."""
    g = np.random.default_rng(seed + 7).standard_normal((n, n, n))
    rho = np.exp(0.1 * g - 0.005)
    del g
    v = random_velocity((n, n, n), 1.0, 16.0, 5.0 / 3.0, 3.0, solenoidal=False, seed=seed)
    return make_state((n, n, n), nb=nb, npass=0, rho=rho, v=v, b0=(0.05, 0.05, 0.0))


def separable_large(n, nb=6, nmodes=6, kmax=16, mach=2.0, b0=(0.05, 0.05, 0.0), seed=5005):
    """Large-box throughput IC built without 3-D FFTs: rho = 1, uniform B0, and a velocity
    whose x-component depends on (y, z) only, y on (z, x) and z on (x, y) -- divergence-free
    by construction -- each a sum of `nmodes` seeded products of sines with integer wave
    numbers 1..kmax, scaled to rms Mach `mach`.  Returns q_3d, bface_3d (no passives); the
    host arrays are filled by broadcasting the 2-D component fields along their third axis."""
    rng = np.random.default_rng(seed)
    sh = padded_shapes(n, n, n, nb, 0)[0]
    x = (np.arange(n) + 0.5) / n
    comps = []
    for c in range(3):
        f = np.zeros((n, n))
        for _ in range(nmodes):
            ka, kb = rng.integers(1, kmax + 1, size=2)
            pa, pb = rng.random(2)
            amp = rng.standard_normal() * (ka * ka + kb * kb) ** -0.5
            f += amp * np.outer(np.sin(2 * np.pi * (ka * x + pa)), np.sin(2 * np.pi * (kb * x + pb)))
        comps.append(f)
    rms = np.sqrt(sum(np.mean(f * f) for f in comps))
    comps = [f * (mach / rms) for f in comps]
    q = np.zeros(sh + (7,), order="F")
    bf = np.zeros(sh + (3,), order="F")
    I = slice(nb, nb + n)
    q[I, I, I, 0] = 1.0
    q[I, I, I, 1] = comps[0][None, :, :]          # v_x(y, z)
    q[I, I, I, 2] = comps[1].T[:, None, :]        # v_y(z, x), stored as (x, z)
    q[I, I, I, 3] = comps[2][:, :, None]          # v_z(x, y)
    for m in range(3):
        q[I, I, I, 4 + m] = b0[m]
        bf[I, I, I, m] = b0[m]
    return q, bf
