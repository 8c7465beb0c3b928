"""Test-side loaders for the two CPU checkers.

* the CPU restatement oracle (oracle/liboracle_imhd.so, built by oracle/Makefile), and
* the reference Fortran solver built from /root/reference by oracle/ref/build_ref.py
  (oracle/_ref/libimhd_ref_<cfg>.so; only present where the reference is).

Test infrastructure only: the product package never imports this module.
"""
import ctypes
import os

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(REPO, "oracle", "liboracle_imhd.so")
REF_DIR = os.path.join(REPO, "oracle", "_ref")

_dp = ctypes.POINTER(ctypes.c_double)


def P(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class OrcCfg(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("nb", ctypes.c_int), ("npass", ctypes.c_int), ("mhd", ctypes.c_int),
                ("riemann", ctypes.c_int), ("recon", ctypes.c_int), ("bc", ctypes.c_int),
                ("cs", ctypes.c_double), ("mc_beta", ctypes.c_double),
                ("protection", ctypes.c_int), ("_pad", ctypes.c_int), ("prot_dens", ctypes.c_double)]


class OrcDiag(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in
                ("ke", "vrms_sq", "momx", "momy", "momz", "divb_sum", "divb_max", "mass")]


_oracle = None


def oracle_lib():
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_SO):
            import subprocess
            subprocess.run(["make", "-C", os.path.join(REPO, "oracle")], check=True,
                           capture_output=True)
        L = ctypes.CDLL(ORACLE_SO)
        L.orc_create.restype = ctypes.c_void_p
        L.orc_create.argtypes = [ctypes.POINTER(OrcCfg)]
        for f in ("orc_destroy", "orc_bc"):
            getattr(L, f).argtypes = [ctypes.c_void_p]
        L.orc_set_state.argtypes = [ctypes.c_void_p] * 4
        L.orc_get_state.argtypes = [ctypes.c_void_p] * 5
        L.orc_step.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.orc_passive_step.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.orc_copy_velocity.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_max_speed.argtypes = [ctypes.c_void_p]
        L.orc_max_speed.restype = ctypes.c_double
        L.orc_smax.argtypes = [ctypes.c_void_p]
        L.orc_smax.restype = ctypes.c_double
        L.orc_floor_count.argtypes = [ctypes.c_void_p]
        L.orc_floor_count.restype = ctypes.c_int
        L.orc_diagnostics.argtypes = [ctypes.c_void_p, ctypes.POINTER(OrcDiag)]
        L.orc_update_test.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_int] + [ctypes.c_void_p] * 4
        L.orc_update_test.restype = ctypes.c_int
        L.orc_riemann_plane.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                        ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, _dp]
        L.orc_recon_plane.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_double, ctypes.c_double, ctypes.c_void_p,
                                      ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.orc_ou_setup.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                   ctypes.c_double, ctypes.c_ulonglong, ctypes.c_int, ctypes.c_int]
        L.orc_ou_setup.restype = ctypes.c_int
        L.orc_ou_step.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.orc_ou_modes_step.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.orc_ou_get.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.orc_ou_field.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        L.orc_philox4x32.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.orc_set_force.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_double]
        L.orc_step_drive.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.orc_step_drive.restype = ctypes.c_int
        L.orc_taus88.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
        L.orc_idft_1d.argtypes = [ctypes.c_void_p] * 4 + [ctypes.c_int] * 4
        L.orc_rand_force.argtypes = [ctypes.c_int] * 5 + [ctypes.c_void_p] * 3
        L.orc_tvd_flux_plane.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                         ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                         ctypes.c_void_p, ctypes.c_void_p]
        _oracle = L
    return _oracle


def padded_shapes(nx, ny, nz, nb, npass):
    sh = (nx + 2 * nb, ny + 2 * nb, nz + 2 * nb)
    shp = (nx + 2 * nb - 4, ny + 2 * nb - 4, nz + 2 * nb - 4, max(npass, 1))
    return sh, shp


class Oracle:
    """The CPU restatement with the reference's array layout."""

    def __init__(self, nx, ny, nz, npass=0, mhd=2, riemann=2, recon=0, bc=2, cs=1.0,
                 nb=6, mc_beta=1.01, protection=False, prot_dens=0.0):
        self.L = oracle_lib()
        self.dims = (nx, ny, nz, nb, npass)
        cfg = OrcCfg(nx, ny, nz, nb, npass, mhd, riemann, recon, bc, cs, mc_beta, int(bool(protection)), 0,
                     prot_dens)
        self.c = self.L.orc_create(ctypes.byref(cfg))
        self.sh, self.shp = padded_shapes(nx, ny, nz, nb, npass)
        self.npass = npass

    def __del__(self):
        try:
            self.L.orc_destroy(self.c)
        except Exception:
            pass

    def set_state(self, q, bf, qp=None):
        if qp is None:
            qp = np.zeros(self.shp, order="F")
        self._keep = (np.asfortranarray(q), np.asfortranarray(bf), np.asfortranarray(qp))
        self.L.orc_set_state(self.c, *[P(a) for a in self._keep])

    def get_state(self):
        q = np.zeros(self.sh + (7,), order="F")
        bf = np.zeros(self.sh + (3,), order="F")
        qp = np.zeros(self.shp, order="F")
        ec = np.zeros(self.sh + (3,), order="F")
        self.L.orc_get_state(self.c, P(q), P(bf), P(qp), P(ec))
        return q, bf, qp, ec

    def bc(self):
        self.L.orc_bc(self.c)

    def step(self, dt):
        self.L.orc_step(self.c, dt)

    def max_speed(self):
        return self.L.orc_max_speed(self.c)

    def smax(self):
        return self.L.orc_smax(self.c)

    def copy_velocity(self, stage):
        self.L.orc_copy_velocity(self.c, stage)

    def passive_step(self, dt):
        self.L.orc_passive_step(self.c, dt)

    def diagnostics(self):
        d = OrcDiag()
        self.L.orc_diagnostics(self.c, ctypes.byref(d))
        return {n: getattr(d, n) for n, _ in OrcDiag._fields_}

    def floor_count(self):
        return self.L.orc_floor_count(self.c)


def ref_available(name):
    return os.path.exists(os.path.join(REF_DIR, f"libimhd_ref_{name}.so"))


_refs = {}


def _oracle_force_methods():
    def set_force(self, delv, edot=1e-3, dt_turb=0.05):
        self._delv = np.asfortranarray(delv, dtype=np.float64)
        self.L.orc_set_force(self.c, P(self._delv), edot, dt_turb)

    def step_drive(self, dt):
        rc = self.L.orc_step_drive(self.c, dt)
        assert rc == 0, "bad forcing normalisation"
    Oracle.set_force = set_force
    Oracle.step_drive = step_drive

    def ou_setup(self, kmin=1.0, kmax=2.0, t_corr=0.5, sigma=1.0, seed=3003, solenoidal=True, box=0):
        self.ou_nmode = self.L.orc_ou_setup(self.c, kmin, kmax, t_corr, sigma, seed, int(solenoidal), box)
        return self.ou_nmode

    def ou_step(self, dt):
        self.L.orc_ou_step(self.c, dt)

    def ou_modes_step(self, dt):
        self.L.orc_ou_modes_step(self.c, dt)

    def ou_modes(self):
        k = np.zeros((self.ou_nmode, 3), dtype=np.int32)
        a = np.zeros((self.ou_nmode, 6))
        self.L.orc_ou_get(self.c, P(k), P(a))
        return k, a[:, :3] + 1j * a[:, 3:]

    def ou_field(self):
        nx, ny, nz = self.dims[:3]
        out = np.zeros((nx, ny, nz, 3), order="F")
        self.L.orc_ou_field(self.c, P(out))
        return out
    Oracle.ou_setup = ou_setup
    Oracle.ou_step = ou_step
    Oracle.ou_modes_step = ou_modes_step
    Oracle.ou_modes = ou_modes
    Oracle.ou_field = ou_field


_oracle_force_methods()


class Reference:
    """The reference Fortran modules (one process-wide instance per build config)."""

    def __init__(self, name):
        if name not in _refs:
            L = ctypes.CDLL(os.path.join(REF_DIR, f"libimhd_ref_{name}.so"))
            L.ref_step.argtypes = [ctypes.c_double]
            if hasattr(L, "ref_step_drive"):
                L.ref_step_drive.argtypes = [ctypes.c_double]
            L.ref_passive_step.argtypes = [ctypes.c_double]
            L.ref_copy_velocity.argtypes = [ctypes.c_int]
            L.ref_set_sound_speed.argtypes = [ctypes.c_double]
            L.ref_vl_riemann.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                         ctypes.c_void_p, ctypes.c_void_p, _dp]
            L.ref_vl_recon.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                       ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
            d = (ctypes.c_int * 8)()
            L.ref_dims(d)
            L.ref_setup()
            _refs[name] = (L, list(d))
        self.L, d = _refs[name]
        self.nx, self.ny, self.nz, self.nb, self.npass, self.mhd, self.riemann, self.bc_type = d
        self.sh, self.shp = padded_shapes(self.nx, self.ny, self.nz, self.nb, self.npass)

    def set_state(self, q, bf, qp=None):
        if qp is None:
            qp = np.zeros(self.shp, order="F")
        self._keep = (np.asfortranarray(q), np.asfortranarray(bf), np.asfortranarray(qp))
        self.L.ref_set_state(*[P(a) for a in self._keep])

    def get_state(self):
        q = np.zeros(self.sh + (7,), order="F")
        bf = np.zeros(self.sh + (3,), order="F")
        qp = np.zeros(self.shp, order="F")
        ec = np.zeros(self.sh + (3,), order="F")
        self.L.ref_get_state(P(q), P(bf), P(qp), P(ec))
        return q, bf, qp, ec

    def bc(self):
        self.L.ref_bc()

    def step(self, dt):
        self.L.ref_step(dt)

    def max_speed(self):
        c = ctypes.c_double()
        self.L.ref_max_speed(ctypes.byref(c))
        return c.value

    def smax(self):
        s = ctypes.c_double()
        self.L.ref_vl_smax(ctypes.byref(s))
        return s.value

    def copy_velocity(self, stage):
        self.L.ref_copy_velocity(stage)

    def passive_step(self, dt):
        self.L.ref_passive_step(dt)

    def set_force(self, delv, edot=None, dt_turb=None):
        """edot / dt_turb are the build's constants (1e-3, 0.05)."""
        self._delv = np.asfortranarray(delv, dtype=np.float64)
        self.L.ref_set_force(P(self._delv))

    def step_drive(self, dt):
        self.L.ref_step_drive(ctypes.c_double(dt))

    # io_mod.f90 through the harness; files go to Output/ under the working directory
    def write_restart(self, t_sim, dump_count, seeds):
        s = (ctypes.c_int * 3)(*seeds)
        self.L.ref_write_restart(ctypes.c_double(t_sim), ctypes.c_int(dump_count), s)

    def read_restart(self):
        t = ctypes.c_double()
        d = ctypes.c_int()
        s = (ctypes.c_int * 3)()
        self.L.ref_read_restart(ctypes.byref(t), ctypes.byref(d), s)
        return t.value, d.value, tuple(s)

    def rand_taus(self, seeds, n):
        out = np.zeros(n)
        s = (ctypes.c_int * 3)(*seeds)
        self.L.ref_rand_taus(s, ctypes.c_int(n), P(out))
        return out

    def idft_1d(self, in_re, in_im, ilow, ihi, nnx):
        m = ihi - ilow + 1
        ore, oim = np.zeros(m), np.zeros(m)
        a, b = np.ascontiguousarray(in_re, dtype=np.float64), np.ascontiguousarray(in_im, dtype=np.float64)
        self.L.ref_idft_1d(P(a), P(b), P(ore), P(oim), ctypes.c_int(ilow), ctypes.c_int(ihi), ctypes.c_int(nnx))
        return ore, oim

    def write_snapshot(self, dump):
        self.L.ref_write_snapshot(ctypes.c_int(dump))

    def parallel_buffer(self):
        n = self.nx * self.ny * self.nz * (7 + self.npass)
        out = np.zeros(n)
        self.L.ref_parallel_buffer(ctypes.c_int(n), P(out))
        return out


def oracle_rand_force(nx, ny, nz, K=2, seeds=(1234, 5678, 9123), nb=6):
    """The oracle's generate_rand_force (modes zeroed first): (delv (Px,Py,Pz,3) Fortran
    order, modes re/im (3, 2K+1, 2K+1, 2K+1) indexed [d, kz+K, ky+K, kx+K], new seeds)."""
    L = oracle_lib()
    sh = (nx + 2 * nb, ny + 2 * nb, nz + 2 * nb)
    delv = np.zeros(sh + (3,), order="F")
    W = 2 * K + 1
    modes = np.zeros(2 * 3 * W ** 3)
    s = (ctypes.c_int * 3)(*seeds)
    L.orc_rand_force(nx, ny, nz, nb, K, s, P(delv), P(modes))
    m = modes.reshape(2, 3, W, W, W)
    return delv, m[0] + 1j * m[1], tuple(s)


def oracle_taus88(seeds, n):
    L = oracle_lib()
    out = np.zeros(n)
    s = (ctypes.c_int * 3)(*seeds)
    L.orc_taus88(s, n, P(out))
    return out, tuple(s)


def io_state(nx, ny, nz, nb=6, npass=1, seed=0):
    """Arbitrary values in every cell, ghosts included, for the file-format tests: signed
    values over many decades plus a few exact specials (zeros of both signs, values that
    SNGL rounds up, overflows to infinity or flushes below the float range)."""
    rng = np.random.default_rng(seed)
    sh, shp = padded_shapes(nx, ny, nz, nb, npass)

    def arr(shape):
        a = rng.uniform(-1.0, 1.0, shape) * 10.0 ** rng.integers(-12, 12, shape)
        flat = a.reshape(-1, order="F")
        specials = [0.0, -0.0, 1.0 + 2.0 ** -24 + 2.0 ** -40, 3.5e38, -1e-46, 1.0e300, 2.0 ** -149]
        flat[:len(specials)] = specials
        return np.asfortranarray(a)

    q = arr(sh + (7,))
    bf = arr(sh + (3,))
    qp = arr(shp[:3] + (npass,)) if npass > 0 else None
    return q, bf, qp


def random_state(nx, ny, nz, nb=6, npass=1, seed=0, amp_v=0.5, amp_rho=0.2, b0=(0.5, 0.2, 0.1),
                 amp_b=0.2):
    """A smooth, divergence-free random periodic state in the reference layout.

    Face B comes from a vector potential on cell edges so that the discrete (raw)
    divergence vanishes to roundoff; cell-centred B is the face average.  Values are
    set on the interior and the ghosts are left for the boundary fill.
    """
    rng = np.random.default_rng(seed)
    sh, shp = padded_shapes(nx, ny, nz, nb, npass)
    q = np.zeros(sh + (7,), order="F")
    bf = np.zeros(sh + (3,), order="F")
    qp = np.zeros(shp, order="F")
    x = (np.arange(nx) + 0.5) / nx
    y = (np.arange(ny) + 0.5) / ny
    z = (np.arange(nz) + 0.5) / nz
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")

    def smooth():
        f = np.zeros((nx, ny, nz))
        for _ in range(3):
            kx, ky, kz = rng.integers(1, 3, size=3)
            ph = rng.random(3) * 2 * np.pi
            f += np.sin(2 * np.pi * kx * X + ph[0]) * np.cos(2 * np.pi * ky * Y + ph[1]) * \
                np.sin(2 * np.pi * kz * Z + ph[2])
        return f / 3.0

    I = (slice(nb, nb + nx), slice(nb, nb + ny), slice(nb, nb + nz))
    rho = 1.0 + amp_rho * smooth()
    q[I + (0,)] = rho
    for m in range(3):
        q[I + (1 + m,)] = rho * amp_v * smooth()
    # periodic edge vector potential A on the interior index grid
    A = [amp_b * smooth() / (2 * np.pi) for _ in range(3)]
    # face B = B0 + curl A (discrete, periodic), raw differences scaled by n
    def d(a, ax):
        return (a - np.roll(a, 1, axis=ax))
    bx = b0[0] + (d(A[2], 1) - d(A[1], 2))
    by = b0[1] + (d(A[0], 2) - d(A[2], 0))
    bz = b0[2] + (d(A[1], 0) - d(A[0], 1))
    bf[I + (0,)] = bx
    bf[I + (1,)] = by
    bf[I + (2,)] = bz
    q[I + (4,)] = 0.5 * (bx + np.roll(bx, 1, axis=0))
    q[I + (5,)] = 0.5 * (by + np.roll(by, 1, axis=1))
    q[I + (6,)] = 0.5 * (bz + np.roll(bz, 1, axis=2))
    Ip = (slice(nb - 2, nb - 2 + nx), slice(nb - 2, nb - 2 + ny), slice(nb - 2, nb - 2 + nz))
    for p in range(npass):
        qp[Ip + (p,)] = 0.5 + 0.4 * smooth()
    return q, bf, qp


def oracle_driven(dims, q, bf, qp, nsteps, riemann=2, recon=2, npass=1, dt_turb=0.02, edot=1e-3, k_max=2,
                  seeds=(1234, 5678, 9123), cour=0.35):
    """The reference driver's loop with drive_turbulence (driver.f90:46-160, 336-344) over
    the oracle: a force field drawn at start-up and after every kick, a kick when t_turb
    falls in (t_old, t_old + dt] with dt halved for that step, and COUR halved for the dt
    after it.  Returns ((q, bf, qp), t, dt, t_turb, seeds, kicks)."""
    O = Oracle(*dims, npass=npass, mhd=2, riemann=riemann, recon=recon, bc=2)
    O.set_state(q, bf, qp)
    O.bc()
    dt, t = 1e-15, 0.0
    dx = 1.0 / max(dims)
    delv, _, seeds = oracle_rand_force(*dims, K=k_max, seeds=seeds)
    drive_next, t_turb, kicks = False, t + dt_turb, 0
    for _ in range(nsteps):
        t_old, c = t, cour
        if npass:
            O.copy_velocity(0)
        if t_turb > t_old and t_turb <= t_old + dt:
            drive_next = True
            t_turb = max(t_turb + dt_turb, t + 1e-2 * dt)
            dt = 0.5 * dt
        if drive_next:
            O.set_force(delv, edot, dt_turb)
            O.step_drive(dt)
        else:
            O.step(dt)
        if npass:
            O.copy_velocity(1)
            O.passive_step(dt)
        t += dt
        if drive_next:
            delv, _, seeds = oracle_rand_force(*dims, K=k_max, seeds=seeds)
            drive_next, c = False, 0.5 * cour
            kicks += 1
        dt = c * dx / O.max_speed()
        O.bc()
    return O.get_state()[:3], t, dt, t_turb, tuple(seeds), kicks
