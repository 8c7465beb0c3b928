"""ctypes binding of libimhd.so (the HIP library).  No fallback: importing the
solver without the built library raises."""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libimhd.so")

IMHD_TVD, IMHD_VL = 1, 2
IMHD_ROE, IMHD_HLLD, IMHD_HLLE = 1, 2, 3
IMHD_PLM_PRIM, IMHD_PPM_PRIM, IMHD_PPM_CHAR = 0, 1, 2
IMHD_BC_OPEN, IMHD_BC_PERIODIC = 1, 2

ERRORS = {-1: "IMHD_EINVAL", -2: "IMHD_EHIP", -3: "IMHD_ESTATE", -4: "IMHD_ENOSPEED", -5: "IMHD_EIO"}

# Every symbol include/imhd.h declares.
EXPORTS = [
    "imhd_create", "imhd_create_ensemble", "imhd_destroy", "imhd_set_state", "imhd_get_state", "imhd_get_field",
    "imhd_init_turbulence",
    "imhd_autotune", "imhd_step", "imhd_set_force", "imhd_set_ou_forcing", "imhd_get_ou_modes",
    "imhd_set_impulsive_forcing", "imhd_get_impulsive_forcing", "imhd_fill_ghosts", "imhd_max_speed", "imhd_copy_velocity", "imhd_passive_step",
    "imhd_diagnostics", "imhd_diagnostics_static", "imhd_floor_count", "imhd_device_bytes", "imhd_run", "imhd_set_time", "imhd_get_time",
    "imhd_set_profiling", "imhd_profile_report", "imhd_sync", "imhd_last_error", "imhd_version",
    "imhd_write_restart", "imhd_read_restart", "imhd_write_cube", "imhd_generate_force", "imhd_get_force",
    "imhd_spectrum", "imhd_density_pdf", "imhd_mode_spectra",
]


class ImhdConfig(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("nb", ctypes.c_int), ("npass", ctypes.c_int), ("integrator", ctypes.c_int),
                ("riemann", ctypes.c_int), ("recon", ctypes.c_int), ("bc", ctypes.c_int),
                ("precision", ctypes.c_int), ("compensated", ctypes.c_int),
                ("flux_protection", ctypes.c_int), ("cs", ctypes.c_double),
                ("mc_beta", ctypes.c_double), ("cour", ctypes.c_double)]


class ImhdDiag(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in
                ("ke", "vrms_sq", "momx", "momy", "momz", "divb_sum", "divb_max", "mass")]


class ImhdOU(ctypes.Structure):
    _fields_ = [("kmin", ctypes.c_double), ("kmax", ctypes.c_double), ("t_corr", ctypes.c_double),
                ("sigma", ctypes.c_double), ("seed", ctypes.c_ulonglong), ("solenoidal", ctypes.c_int),
                ("_reserved", ctypes.c_int)]


class ImhdTurbIC(ctypes.Structure):
    _fields_ = [("rho0", ctypes.c_double), ("sigma_ln", ctypes.c_double), ("b0", ctypes.c_double * 3),
                ("kmin", ctypes.c_double), ("kmax", ctypes.c_double), ("slope", ctypes.c_double),
                ("mach", ctypes.c_double), ("seed", ctypes.c_ulonglong), ("solenoidal", ctypes.c_int),
                ("passive", ctypes.c_int)]


class ImhdError(RuntimeError):
    pass


_lib = None


def lib():
    """Load libimhd.so; raise if it has not been built (there is no CPU fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImhdError(f"{LIB_PATH} is missing: build it with `python -m isothermal_mhd_amd.build` "
                        "(hipcc, gfx950); the solver has no CPU fallback")
    L = ctypes.CDLL(LIB_PATH)
    vp, ip, dp = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    L.imhd_create.argtypes = [ctypes.POINTER(ImhdConfig), ctypes.POINTER(vp)]
    L.imhd_create_ensemble.argtypes = [ctypes.POINTER(ImhdConfig), ip, ctypes.POINTER(vp)]
    L.imhd_destroy.argtypes = [vp]
    L.imhd_set_state.argtypes = [vp, ip, vp, vp, vp]
    L.imhd_get_state.argtypes = [vp, ip, vp, vp, vp]
    L.imhd_get_field.argtypes = [vp, ip, ip, vp]
    L.imhd_init_turbulence.argtypes = [vp, ip, ctypes.POINTER(ImhdTurbIC)]
    L.imhd_step.argtypes = [vp, dp, ip]
    L.imhd_set_force.argtypes = [vp, ip, vp, dp, dp]
    L.imhd_autotune.argtypes = [vp, ip, ctypes.c_char_p, ip]
    L.imhd_set_ou_forcing.argtypes = [vp, ctypes.POINTER(ImhdOU)]
    L.imhd_set_impulsive_forcing.argtypes = [vp, ip, dp, dp, ctypes.POINTER(ctypes.c_int)]
    L.imhd_get_impulsive_forcing.argtypes = [vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)]
    L.imhd_get_ou_modes.argtypes = [vp, ip, ctypes.POINTER(ctypes.c_int), vp, vp]
    L.imhd_fill_ghosts.argtypes = [vp]
    L.imhd_max_speed.argtypes = [vp, vp]
    L.imhd_copy_velocity.argtypes = [vp, ip]
    L.imhd_passive_step.argtypes = [vp, dp]
    L.imhd_diagnostics.argtypes = [vp, ctypes.POINTER(ImhdDiag)]
    L.imhd_diagnostics_static.argtypes = [vp, ctypes.POINTER(ImhdDiag)]
    L.imhd_floor_count.argtypes = [vp, ctypes.POINTER(ctypes.c_uint)]
    L.imhd_device_bytes.argtypes = [vp, ctypes.POINTER(ctypes.c_ulonglong)]
    L.imhd_run.argtypes = [vp, ip, dp, ip]
    L.imhd_set_time.argtypes = [vp, vp, vp]
    L.imhd_get_time.argtypes = [vp, vp, vp]
    L.imhd_set_profiling.argtypes = [vp, ip]
    L.imhd_profile_report.argtypes = [vp]
    L.imhd_profile_report.restype = ctypes.c_char_p
    L.imhd_sync.argtypes = [vp]
    L.imhd_write_restart.argtypes = [vp, ip, ctypes.c_char_p, dp, ip, vp]
    L.imhd_read_restart.argtypes = [vp, ip, ctypes.c_char_p, vp, vp, vp]
    L.imhd_write_cube.argtypes = [vp, ip, ctypes.c_char_p, ip, ip]
    L.imhd_generate_force.argtypes = [vp, ip, ip, dp, dp, vp]
    L.imhd_get_force.argtypes = [vp, ip, vp]
    L.imhd_spectrum.argtypes = [vp, ip, ip, vp, vp]
    L.imhd_density_pdf.argtypes = [vp, ip, ip, dp, vp, vp]
    L.imhd_mode_spectra.argtypes = [vp, ip, vp, vp]
    L.imhd_last_error.argtypes = [vp]
    L.imhd_last_error.restype = ctypes.c_char_p
    L.imhd_version.restype = ctypes.c_char_p
    for f in EXPORTS:
        if f not in ("imhd_profile_report", "imhd_last_error", "imhd_version"):
            getattr(L, f).restype = ctypes.c_int
    _lib = L
    return L
