"""Seeded inputs of the golden vectors (tests/golden/reference_vectors.npz)."""
import numpy as np

STEP_DTS = (1e-15, 2e-3, 1.5e-3)
# name -> (reference build, dims, npass, seed)
STEP_CASES = {
    "hlld_p12x10x8": ("vl_hlld_p12x10x8", (12, 10, 8), 2, 31),
    "hlld_o12x10x8": ("vl_hlld_o12x10x8", (12, 10, 8), 1, 32),
    "roe_p16": ("vl_roe_p16", (16, 16, 16), 1, 33),
    "hlle_p16": ("vl_hlle_p16", (16, 16, 16), 1, 34),
    "tvd_roe_p16": ("tvd_roe_p16", (16, 16, 16), 1, 35),
    "tvd_hlle_p12x10x8": ("tvd_hlle_p12x10x8", (12, 10, 8), 1, 36),
    "tvd_roe_o12x10x8": ("tvd_roe_o12x10x8", (12, 10, 8), 1, 37),
}
# TVD-module plane fluxes: (kind, predictor, dt); kind 2 K98 Roe-TVD, 1 HLLE; built from
# the 16^3 TVD build (dx = 1/16)
TVD_FLUX_CASES = ((2, 1, 2e-3), (2, 0, 2e-3), (2, 0, 1e-15), (1, 0, 2e-3))
TVD_FLUX_DX = 1.0 / 16


def step_case_params(name):
    """(mhd, riemann, bc) in the reference's own numbering for a STEP_CASES entry."""
    build = STEP_CASES[name][0]
    bc = 1 if "_o" in build else 2
    if build.startswith("tvd_"):
        return 1, {"roe": 2, "hlle": 1}[build.split("_")[1]], bc
    return 2, {"hlld": 2, "hlle": 3, "roe": 1}[build.split("_")[1]], bc


def tvd_flux_inputs(ni=28, nj=3, seed=300):
    """cell states with their own (independent) face fields in slot 4"""
    A, _ = riemann_inputs(2, ni, nj)
    rng = np.random.default_rng(seed)
    A = np.array(A, order="F")
    A[:, :, 4] = rng.normal(0, 1.0, (ni, nj))
    A[3, :, 4] = 0.0
    return A


def riemann_inputs(kind, ni=27, nj=3):
    rng = np.random.default_rng(100 + kind)
    n = ni * nj
    qL = np.zeros((n, 7))
    qR = np.zeros((n, 7))
    for q in (qL, qR):
        q[:, 0] = 0.2 + 2.0 * rng.random(n)
        q[:, 1:4] = q[:, [0]] * rng.normal(0, 1.5, (n, 3))
        q[:, 5:7] = rng.normal(0, 1.0, (n, 2))
    qL[:, 4] = rng.normal(0, 1.0, n)
    qL[:n // 8, 4] = 0.0                      # degenerate normal field
    qL[n // 8:n // 4, 5:7] = 0.0              # no transverse field
    qR[n // 8:n // 4, 5:7] = 0.0
    qR[:, 4] = qL[:, 4]
    A = np.asfortranarray(qL.reshape(nj, ni, 7).transpose(1, 0, 2))
    B = np.asfortranarray(qR.reshape(nj, ni, 7).transpose(1, 0, 2))
    return A, B


def recon_inputs(kind, ni=28, nj=3):
    rng = np.random.default_rng(200 + kind)
    x = np.linspace(0, 2 * np.pi, ni)
    w = np.zeros((ni, nj, 6))
    w[..., 0] = 1.0 + 0.3 * np.sin(x)[:, None] + 0.05 * rng.random((ni, nj))
    for m in range(1, 4):
        w[..., m] = np.cos((m + 1) * x)[:, None] + 0.2 * rng.standard_normal((ni, nj))
    w[..., 4:6] = rng.standard_normal((ni, nj, 2))
    w[3] = w[2]
    wb = 0.5 + 0.5 * rng.standard_normal((ni, nj))
    return np.asfortranarray(w), np.asfortranarray(wb)


# driven VL steps (reference build with drive_turbulence = .TRUE., Edot = 1e-3,
# dt_turb = 0.05): (reference build, dims, npass, state seed, force seed, dts, drive flags)
DRIVE_CASE = ("vl_hlld_p16_drive", (16, 16, 16), 1, 71, 72, (1e-15, 2e-3, 1.5e-3), (True, False, True))
DRIVE_EDOT, DRIVE_DT_TURB = 1e-3, 0.05


def drive_force(shape, seed):
    """a seeded force field on the padded grid, the shape of the driver's delv_re"""
    rng = np.random.default_rng(seed)
    return np.asfortranarray(rng.standard_normal(tuple(shape) + (3,)))
