"""Host-side mirror of the reference's solver interface over the imhd C ABI.

Method names follow the reference's Fortran procedures (3D_MPI/src):
vanleer_solve_3d / get_max_speed_vanleer (VanLeerSolver, vanleer_mod.f90),
tvd_solve_3d / get_max_speed_tvd (TvdSolver, isotvd_mod.f90), boundary_conditions,
copy_velocity, passive_solve_3d, compute_KE/compute_divB, and the driver's time loop
(driver.f90:35-186).
Arrays are the reference's host layout (FP64, Fortran order, nb ghost layers).
"""
import ctypes
import os

import numpy as np

from . import _lib as L

_RIEMANN = {"roe": L.IMHD_ROE, "hlld": L.IMHD_HLLD, "hlle": L.IMHD_HLLE}
_BC = {"open": L.IMHD_BC_OPEN, "periodic": L.IMHD_BC_PERIODIC}
_RECON = {"plm": L.IMHD_PLM_PRIM, "ppm_prim": L.IMHD_PPM_PRIM, "ppm_char": L.IMHD_PPM_CHAR}


def padded_shapes(nx, ny, nz, nb, npass):
    """Host shapes of q_3d / bface_3d (without the field axis) and qpass_3d."""
    sh = (nx + 2 * nb, ny + 2 * nb, nz + 2 * nb)
    shp = (nx + 2 * nb - 4, ny + 2 * nb - 4, nz + 2 * nb - 4, max(npass, 1))
    return sh, shp


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class VanLeerSolver:
    """VL2 + CT isothermal MHD on MI355X for one box or an ensemble of boxes.

    `dims` is (nx, ny, nz) or a list of such triples (ensemble: one launch per kernel
    serves every box; each box keeps its own dt).
    """

    integrator = L.IMHD_VL

    def __init__(self, dims, nb=6, npass=0, riemann="hlld", recon="plm", bc="periodic",
                 precision=64, compensated=None, cs=1.0, mc_beta=1.01, cour=0.35, flux_protection=False):
        self.lib = L.lib()
        boxes = [tuple(dims)] if np.ndim(dims) == 1 else [tuple(d) for d in dims]
        if recon not in _RECON:
            raise ValueError("recon must be 'plm', 'ppm_prim' or 'ppm_char'")
        if compensated is None:
            compensated = precision == 32
        cfgs = (L.ImhdConfig * len(boxes))()
        for i, (nx, ny, nz) in enumerate(boxes):
            c = cfgs[i]
            c.nx, c.ny, c.nz, c.nb, c.npass = nx, ny, nz, nb, npass
            c.integrator, c.riemann, c.recon = self.integrator, _RIEMANN[riemann], _RECON[recon]
            c.bc, c.precision, c.compensated = _BC[bc], precision, int(bool(compensated))
            c.flux_protection = int(bool(flux_protection))
            c.cs, c.mc_beta = cs, mc_beta
            c.cour = cour[i] if np.ndim(cour) else cour
        self.boxes, self.nb, self.npass, self.precision = boxes, nb, npass, precision
        self._c = ctypes.c_void_p()
        rc = self.lib.imhd_create_ensemble(cfgs, len(boxes), ctypes.byref(self._c))
        self._check(rc)

    # -- plumbing -------------------------------------------------------------------
    def _check(self, rc):
        if rc != 0:
            msg = self.lib.imhd_last_error(self._c).decode() if self._c else ""
            raise L.ImhdError(f"{L.ERRORS.get(rc, rc)}: {msg}")

    def close(self):
        if self._c:
            self.lib.imhd_destroy(self._c)
            self._c = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def shapes(self, box=0):
        nx, ny, nz = self.boxes[box]
        return padded_shapes(nx, ny, nz, self.nb, self.npass)

    # -- state transfer -------------------------------------------------------------
    def set_state(self, q_3d, bface_3d, qpass_3d=None, box=0):
        q = np.asfortranarray(q_3d, dtype=np.float64)
        b = np.asfortranarray(bface_3d, dtype=np.float64)
        sh, shp = self.shapes(box)
        if q.shape != sh + (7,) or b.shape != sh + (3,):
            raise ValueError(f"expected q_3d {sh + (7,)} and bface_3d {sh + (3,)}")
        qp = None
        if self.npass > 0:
            if qpass_3d is None:
                raise ValueError("qpass_3d required when npass > 0")
            qp = np.asfortranarray(qpass_3d, dtype=np.float64)
            if qp.shape[:3] != shp[:3] or qp.shape[3] < self.npass:
                raise ValueError(f"expected qpass_3d {shp[:3] + (self.npass,)}")
            qp = np.asfortranarray(qp[..., :self.npass])
        self._check(self.lib.imhd_set_state(self._c, box, _p(q), _p(b), _p(qp) if qp is not None else None))

    def get_state(self, box=0):
        sh, shp = self.shapes(box)
        q = np.zeros(sh + (7,), order="F")
        b = np.zeros(sh + (3,), order="F")
        qp = np.zeros(shp[:3] + (max(self.npass, 1),), order="F")
        self._check(self.lib.imhd_get_state(self._c, box, _p(q), _p(b), _p(qp)))
        return q, b, qp

    def get_field(self, field, box=0, out=None):
        """One field of the device state on its padded grid: 0..6 q_3d(...,1..7), 7..9
        bface_3d(...,1..3), 10 + p qpass_3d(...,p+1).  `out` (FP64, Fortran order, that
        shape) is filled and returned instead of a new array."""
        sh, shp = self.shapes(box)
        shape = shp[:3] if field >= 10 else sh
        if out is None:
            out = np.zeros(shape, order="F")
        elif out.shape != shape or out.dtype != np.float64 or not out.flags.f_contiguous:
            raise ValueError(f"out must be a Fortran-ordered float64 array of shape {shape}")
        self._check(self.lib.imhd_get_field(self._c, box, int(field), _p(out)))
        return out

    def init_turbulence(self, rho0=1.0, sigma_ln=0.0, b0=(0.0, 0.0, 1.0), kmin=1.0, kmax=8.0, slope=2.0,
                        mach=1.0, seed=0, solenoidal=True, passive=False, box=0):
        """Seeded random-field start generated on the device (see imhd.h,
        imhd_init_turbulence): lognormal density of width sigma_ln about rho0, Gaussian
        Fourier-mode velocity with shell spectrum k^-slope on kmin <= |n| <= kmax scaled to
        rms Mach `mach`, uniform field b0, passive scalars U[0,1) when `passive`."""
        ic = L.ImhdTurbIC(float(rho0), float(sigma_ln), (ctypes.c_double * 3)(*[float(x) for x in b0]),
                          float(kmin), float(kmax), float(slope), float(mach), int(seed),
                          int(bool(solenoidal)), int(bool(passive)))
        self._check(self.lib.imhd_init_turbulence(self._c, box, ctypes.byref(ic)))

    # -- restart and snapshot files (io_mod.f90) ---------------------------------------
    def write_restart(self, path, t_sim, dump_count=0, seeds=(1234, 5678, 9123), box=0):
        """write_restart (io_mod.f90:440-473) of one box, straight from the device state:
        records t_sim, q_3d, bface_3d, qpass_3d (npass > 0), dump_count, s1, s2, s3.  The
        default seeds are the reference's s1_def, s2_def, s3_def (constants_mod.f90:23)."""
        s = (ctypes.c_int * 3)(*[int(v) for v in seeds])
        self._check(self.lib.imhd_write_restart(self._c, box, os.fsencode(path), float(t_sim), int(dump_count), s))

    def read_restart(self, path, box=0):
        """read_restart (io_mod.f90:477-517) into one box's device state; returns
        (t_sim, dump_count, (s1, s2, s3)).  Raises ImhdError (IMHD_EIO) when the file's
        records do not match this context's sizes; the state is then unchanged."""
        t = ctypes.c_double()
        d = ctypes.c_int()
        s = (ctypes.c_int * 3)()
        self._check(self.lib.imhd_read_restart(self._c, box, os.fsencode(path), ctypes.byref(t), ctypes.byref(d), s))
        return t.value, d.value, tuple(s)

    def write_cube(self, path, double_prec=False, npass_out=None, box=0):
        """Snapshot cube of the interior (io_mod.f90:249-277): per cell the 7 fields and
        passives 1..npass_out, REAL*8 or REAL*4.  npass_out defaults to min(npass, 1), the
        reference's xyz cube; npass_out = npass gives the parallel dump's layout
        (io_mod.f90:368-436)."""
        if npass_out is None:
            npass_out = min(self.npass, 1)
        self._check(self.lib.imhd_write_cube(self._c, box, os.fsencode(path), int(bool(double_prec)), int(npass_out)))

    # -- reference procedures ------------------------------------------------------
    def vanleer_solve_3d(self, dt, drive_next=False):
        """vanleer_mod.f90:279-426 on every box with the given dt; drive_next applies the
        force field of set_force (get_rho_turb before, add_random_force after the step)."""
        self._check(self.lib.imhd_step(self._c, float(dt), int(bool(drive_next))))

    def set_force(self, delv_re, edot, dt_turb, box=0):
        """The driver's force field delv_re (padded grid, 3 components; what
        generate_rand_force leaves in grid_data_mod) with Edot and dt_turb."""
        sh, _ = self.shapes(box)
        d = np.asfortranarray(delv_re, dtype=np.float64)
        if d.shape != sh + (3,):
            raise ValueError(f"expected delv {sh + (3,)}")
        self._check(self.lib.imhd_set_force(self._c, box, _p(d), float(edot), float(dt_turb)))

    def generate_rand_force(self, k_max=2, edot=1.0e-3, dt_turb=0.05, seeds=(1234, 5678, 9123), box=0):
        """generate_rand_force (turb_driver_mod.f90:197-391) on the device, with the modes
        the reference leaves uninitialised set to zero; the field is applied by the next
        vanleer_solve_3d(dt, drive_next=True).  seeds is the driver's taus88 state
        (s1, s2, s3); returns the advanced state."""
        s = (ctypes.c_int * 3)(*[int(v) for v in seeds])
        self._check(self.lib.imhd_generate_force(self._c, box, int(k_max), float(edot), float(dt_turb), s))
        return tuple(s)

    def get_force(self, box=0):
        sh, _ = self.shapes(box)
        delv = np.zeros(sh + (3,), order="F")
        self._check(self.lib.imhd_get_force(self._c, box, _p(delv)))
        return delv

    _SPECTRA = {"density": 0, "velocity": 1, "bfield": 2, "vpar": 3, "vperp": 4}

    def spectrum(self, field="velocity", box=0):
        """compute_pk of the reference's analysis program (analysis_main.f90:543-633) on the
        device state: field 'density', 'velocity', 'bfield', 'vpar' or 'vperp'; returns
        (k, Pk), n shells each (cubic boxes)."""
        n = self.boxes[box][0]
        k = np.zeros(n)
        pk = np.zeros(n)
        self._check(self.lib.imhd_spectrum(self._c, box, self._SPECTRA[field], _p(k), _p(pk)))
        return k, pk

    def mode_spectra(self, box=0):
        """mode_decomposition of the analysis program (analysis_main.f90:305-540): returns
        (k, P), P of shape (8, n): v_Alfven, v_fast, v_slow, b_Alfven, b_fast, b_slow,
        rho_fast, rho_slow."""
        n = self.boxes[box][0]
        k = np.zeros(n)
        P = np.zeros(8 * n)
        self._check(self.lib.imhd_mode_spectra(self._c, box, _p(k), _p(P)))
        return k, P.reshape(8, n)

    def density_pdf(self, nbins=None, rho_max=4.0, box=0):
        """compute_density_pdf (analysis_main.f90:637-698): (bin centres, PDF); the
        reference uses nbins = n on [0, 4)."""
        nbins = int(nbins or min(self.boxes[box]))
        r = np.zeros(nbins)
        pdf = np.zeros(nbins)
        self._check(self.lib.imhd_density_pdf(self._c, box, nbins, float(rho_max), _p(r), _p(pdf)))
        return r, pdf

    def mhd_update(self, dt):
        """driver.f90:378-389: the configured integrator's solve on every box."""
        self.vanleer_solve_3d(dt)

    def boundary_conditions(self):
        """boundary_conditions_mod.f90:20-234 (one rank)."""
        self._check(self.lib.imhd_fill_ghosts(self._c))

    def get_max_speed_vanleer(self):
        """vanleer_mod.f90:3266-3334; one value per box."""
        out = np.zeros(len(self.boxes))
        self._check(self.lib.imhd_max_speed(self._c, _p(out)))
        return out if len(self.boxes) > 1 else float(out[0])

    def copy_velocity(self, predictor_stage):
        """passive_mod.f90:204-244."""
        self._check(self.lib.imhd_copy_velocity(self._c, int(bool(predictor_stage))))

    def passive_solve_3d(self, dt):
        """passive_mod.f90:118-201."""
        self._check(self.lib.imhd_passive_step(self._c, float(dt)))

    def diagnostics(self):
        """compute_KE + compute_divB (driver.f90:492-553), plus max|div B| and mass."""
        out = (L.ImhdDiag * len(self.boxes))()
        self._check(self.lib.imhd_diagnostics(self._c, out))
        res = [{n: getattr(d, n) for n, _ in L.ImhdDiag._fields_} for d in out]
        return res if len(self.boxes) > 1 else res[0]

    def floor_count(self):
        n = ctypes.c_uint()
        self._check(self.lib.imhd_floor_count(self._c, ctypes.byref(n)))
        return n.value

    def device_bytes(self):
        """Device memory held by this context's field and work arrays (bytes)."""
        n = ctypes.c_ulonglong()
        self._check(self.lib.imhd_device_bytes(self._c, ctypes.byref(n)))
        return n.value

    def set_ou_forcing(self, kmin=1.0, kmax=2.0, t_corr=0.5, sigma=1.0, seed=3003, solenoidal=True):
        """Ornstein-Uhlenbeck forcing of the native loop (see imhd.h); sigma <= 0 switches
        it off."""
        o = L.ImhdOU(kmin, kmax, t_corr, sigma, seed, int(bool(solenoidal)), 0)
        self._check(self.lib.imhd_set_ou_forcing(self._c, ctypes.byref(o)))

    def set_impulsive_forcing(self, k_max=2, edot=1e-3, dt_turb=0.05, seeds=(1234, 5678, 9123)):
        """The reference driver's impulsive forcing inside run() (driver.f90:57-70, 102-111,
        336-344): draws the first field from the taus88 state seeds and kicks every dt_turb
        with that step's dt halved and the next one's COUR halved (see imhd.h); k_max <= 0
        switches it off."""
        sd = (ctypes.c_int * 3)(*[int(np.int32(v)) for v in seeds])
        self._check(self.lib.imhd_set_impulsive_forcing(self._c, int(k_max), float(edot), float(dt_turb), sd))

    def impulsive_forcing_state(self):
        """(taus88 state s1, s2, s3, next drive time)"""
        sd = (ctypes.c_int * 3)()
        tt = ctypes.c_double()
        self._check(self.lib.imhd_get_impulsive_forcing(self._c, sd, ctypes.byref(tt)))
        return tuple(sd), tt.value

    def ou_modes(self, box=0):
        """(integer wave vectors (nmode, 3), complex amplitudes (nmode, 3))"""
        n = ctypes.c_int()
        self._check(self.lib.imhd_get_ou_modes(self._c, box, ctypes.byref(n), None, None))
        k = np.zeros((n.value, 3), dtype=np.int32)
        a = np.zeros((n.value, 6))
        if n.value:
            self._check(self.lib.imhd_get_ou_modes(self._c, box, ctypes.byref(n), _p(k), _p(a)))
        return k, a[:, :3] + 1j * a[:, 3:]

    # -- native loop ----------------------------------------------------------------
    def set_time(self, t=None, dt=None):
        nb = len(self.boxes)
        tv = None if t is None else np.broadcast_to(np.asarray(t, dtype=np.float64), (nb,)).copy()
        dv = None if dt is None else np.broadcast_to(np.asarray(dt, dtype=np.float64), (nb,)).copy()
        self._check(self.lib.imhd_set_time(self._c, _p(tv) if tv is not None else None,
                                           _p(dv) if dv is not None else None))

    def get_time(self):
        t = np.zeros(len(self.boxes))
        dt = np.zeros(len(self.boxes))
        self._check(self.lib.imhd_get_time(self._c, _p(t), _p(dt)))
        return t, dt

    def run(self, nsteps, t_end=1e30, graph=True):
        """Device-resident time loop (driver.f90:35-186) for nsteps steps."""
        self._check(self.lib.imhd_run(self._c, int(nsteps), float(t_end), int(bool(graph))))

    def autotune(self, steps=5):
        """Pick the fastest launch shape on the current state (state and time restored);
        returns the report string."""
        buf = ctypes.create_string_buffer(2048)
        self._check(self.lib.imhd_autotune(self._c, int(steps), buf, len(buf)))
        return buf.value.decode()

    def sync(self):
        self._check(self.lib.imhd_sync(self._c))

    def set_profiling(self, on=True):
        self._check(self.lib.imhd_set_profiling(self._c, int(bool(on))))

    def profile(self):
        """Per-kernel HIP-event time of the last eager call:
        {name: (ms_total, launches, work_items_total)}."""
        rep = self.lib.imhd_profile_report(self._c).decode()
        out = {}
        for item in filter(None, rep.split(";")):
            name, v = item.split("=")
            ms, n, items = v.split("/")
            out[name] = (float(ms), int(n), int(items))
        return out


class TvdSolver(VanLeerSolver):
    """K98 Roe-TVD (riemann='roe') or HLLE (riemann='hlle') + CTU + CT isothermal MHD
    (Iso_TVD_Solver_mod, isotvd_mod.f90) on MI355X; same state layout, boundary, passive
    and native-loop interface as VanLeerSolver."""

    integrator = L.IMHD_TVD

    def __init__(self, dims, nb=6, npass=0, riemann="roe", bc="periodic", precision=64,
                 compensated=None, cs=1.0, cour=0.35):
        if riemann not in ("roe", "hlle"):
            raise ValueError("the TVD integrator takes riemann='roe' (K98 Roe-TVD) or 'hlle'")
        super().__init__(dims, nb=nb, npass=npass, riemann=riemann, recon="plm", bc=bc,
                         precision=precision, compensated=compensated, cs=cs, cour=cour)

    def tvd_solve_3d(self, dt):
        """isotvd_mod.f90:226-364 on every box with the given dt."""
        self._check(self.lib.imhd_step(self._c, float(dt), 0))

    def vanleer_solve_3d(self, dt, drive_next=False):
        raise AttributeError("TvdSolver integrates with tvd_solve_3d")

    def mhd_update(self, dt):
        self.tvd_solve_3d(dt)

    def get_max_speed_tvd(self):
        """isotvd_mod.f90:2607-2673 (interface-mean fast speeds over [0, n]); one value per
        box.  Like the reference driver, call it before boundary_conditions."""
        out = np.zeros(len(self.boxes))
        self._check(self.lib.imhd_max_speed(self._c, _p(out)))
        return out if len(self.boxes) > 1 else float(out[0])

    def get_max_speed_vanleer(self):
        raise AttributeError("TvdSolver's CFL speed is get_max_speed_tvd")


def driver_step(solver, dt):
    """One pass of the reference driver loop body with a host dt (driver.f90:55-160):
    copy_velocity(.FALSE.), mhd_update, copy_velocity(.TRUE.), passive_solve_3d,
    boundary_conditions."""
    if solver.npass > 0:
        solver.copy_velocity(False)
    solver.mhd_update(dt)
    if solver.npass > 0:
        solver.copy_velocity(True)
        solver.passive_solve_3d(dt)
    solver.boundary_conditions()
