"""The fused VL stage kernel (one launch per stage: directional fluxes, corner EMFs and the
update) against the three-kernel path it replaces (k_flux3 -> k_corner -> k_update_z), and
its tile shapes against each other.  FP64 must agree bit for bit over native-loop runs
(ghost fills, passive scalars, the CFL dt sequence and the density-floor count included);
FP32 within rounding.  The path and shape are chosen per context from IMHD_VSTAGE and
IMHD_VSHAPE, which imhd_create reads."""
import sys

import numpy as np
import pytest

from refutil import random_state

pytestmark = pytest.mark.gpu

# dims, riemann, boundary, reconstruction, passives: multi-tile boxes in x and y, several
# z-segments, non-cubic and thin boxes, both boundaries, all three solvers, PPM (fused
# predictor with the three-kernel corrector) and PLM (both stages fused)
CASES = [((40, 36, 40), "hlld", "periodic", "plm", 2),
         ((13, 9, 11), "roe", "open", "plm", 1),
         ((34, 17, 9), "hlle", "open", "plm", 0),
         ((16, 16, 16), "hlld", "periodic", "ppm_char", 1),
         ((12, 20, 7), "roe", "periodic", "ppm_prim", 0)]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def run(monkeypatch, env, dims, riemann, bc, recon, npass, precision=64, steps=6, seed=41):
    from isothermal_mhd_amd import VanLeerSolver
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    try:
        S = VanLeerSolver(dims, npass=npass, riemann=riemann, recon=recon, bc=bc, precision=precision)
    finally:
        for k in env:
            monkeypatch.delenv(k)
    q, bf, qp = random_state(*dims, npass=npass, seed=seed)
    S.set_state(q, bf, qp if npass else None)
    S.boundary_conditions()
    S.run(steps, graph=False)
    out = S.get_state(), S.get_time(), S.floor_count()
    S.close()
    return out


@pytest.mark.parametrize("case", CASES, ids=[f"{c[1]}_{c[2]}_{c[3]}_{'x'.join(map(str, c[0]))}" for c in CASES])
def test_fused_stage_equals_three_kernel_path_fp64(monkeypatch, case):
    dims, riemann, bc, recon, npass = case
    say(f"[gpu] fused vs three-kernel {case}")
    ref = run(monkeypatch, {"IMHD_VSTAGE": "0"}, *case)
    for shape in ("0", "1"):
        got = run(monkeypatch, {"IMHD_VSTAGE": "1", "IMHD_VSHAPE": shape}, *case)
        for x, y in zip(ref[0], got[0]):
            if x is not None:
                assert np.array_equal(x, y), f"state differs (tile shape {shape})"
        assert np.array_equal(ref[1][0], got[1][0]) and np.array_equal(ref[1][1], got[1][1]), "t / dt differ"
        assert ref[2] == got[2]


def test_fused_stage_segments_fp64(monkeypatch):
    """The z-split of the fused stages -- uniform segments (IMHD_VSEG) or a whole number of
    rounds of resident workgroups with s or s+1 balanced segments per tile (IMHD_VROUNDS) --
    changes only the work split (segment boundary planes are solved twice); results are
    identical."""
    case = CASES[0]
    say("[gpu] fused stage z-segments")
    ref = run(monkeypatch, {"IMHD_VSEG": "16"}, *case)
    for env in ({"IMHD_VSEG": "1"}, {"IMHD_VSEG": "5"}, {"IMHD_VSEG": "64"},
                {"IMHD_VROUNDS": "1"}, {"IMHD_VROUNDS": "3"}, {"IMHD_VROUNDS": "1000"}):
        got = run(monkeypatch, env, *case)
        for x, y in zip(ref[0], got[0]):
            if x is not None:
                assert np.array_equal(x, y), f"state differs ({env})"


def test_fused_stage_fp32_matches_three_kernel_path(monkeypatch):
    """FP32: the fused kernel's arithmetic per value is the three-kernel path's; FMA
    contraction may differ between the two compilations, so agreement is to rounding."""
    dims, riemann, bc, recon, npass = CASES[0]
    say("[gpu] fused vs three-kernel fp32")
    ref = run(monkeypatch, {"IMHD_VSTAGE": "0"}, dims, riemann, bc, recon, npass, precision=32)
    got = run(monkeypatch, {"IMHD_VSTAGE": "1"}, dims, riemann, bc, recon, npass, precision=32)
    for x, y in zip(ref[0][:2], got[0][:2]):
        assert np.max(np.abs(x - y)) <= 1e-5 * max(1.0, np.max(np.abs(x)))
    assert abs(ref[1][1][0] - got[1][1][0]) <= 1e-5 * ref[1][1][0]


def test_fused_ppm_corrector_fp32_matches_three_kernel_path(monkeypatch):
    """FP32 fused PPM correctors (characteristic and primitive, two waves per SIMD) against
    the three-kernel path (IMHD_VPPM=0): to rounding, on a multi-tile periodic box and an
    open one."""
    for case in ((( 40, 36, 24), "hlld", "periodic", "ppm_char", 1), ((22, 17, 13), "roe", "open", "ppm_prim", 0)):
        say(f"[gpu] fp32 fused PPM corrector vs three-kernel {case}")
        ref = run(monkeypatch, {"IMHD_VPPM": "0"}, *case, precision=32)
        got = run(monkeypatch, {}, *case, precision=32)
        for x, y in zip(ref[0][:2], got[0][:2]):
            assert np.max(np.abs(x - y)) <= 1e-5 * max(1.0, np.max(np.abs(x)))
        assert abs(ref[1][1][0] - got[1][1][0]) <= 1e-5 * ref[1][1][0]


def test_fused_stage_fp32_div_b_every_z_split(monkeypatch):
    """FP32 bulk with FP64 face B and corner EMFs: constrained transport keeps div B at FP64
    rounding only if every corner EMF that two z-segments (or two tiles) both form has the
    same bits in both.  A plane shared by two segments must come from the same compiled
    code in each (with FMA contraction, two inlined copies may round differently), so
    max|div B| stays at ~1e-14 for every z-split.  The case is the 128^3 member of the
    turbulence IC family of the 256^3 benchmark: a version that formed each segment's first
    plane in a separate prologue reached 2.1e-8 here within 15 steps (64^3 stayed clean)."""
    from isothermal_mhd_amd import VanLeerSolver, ics
    n = 128
    q, bf, qp = ics.config_c2(n=n)
    for env in ({}, {"IMHD_VSEG": "5"}, {"IMHD_VSEG": "16"}, {"IMHD_VROUNDS": "3"},
                {"IMHD_VSHAPE": "0", "IMHD_VSEG": "7"}, {"IMHD_VSHAPE": "2"}, {"IMHD_VSHAPE": "2", "IMHD_VSEG": "5"}):
        say(f"[gpu] fp32 fused stage div B {n}^3 {env}")
        for k, v in env.items():
            monkeypatch.setenv(k, v)
        try:
            S = VanLeerSolver((n, n, n), npass=1, riemann="hlld", recon="plm", bc="periodic", precision=32,
                              cour=0.35)
        finally:
            for k in env:
                monkeypatch.delenv(k)
        S.set_state(q, bf, qp)
        S.boundary_conditions()
        S.set_time(t=0.0, dt=1.0e-15)
        d0 = S.diagnostics()["divb_max"]
        S.run(15, graph=False)
        d1 = S.diagnostics()["divb_max"]
        S.close()
        assert d0 <= 1e-12, d0
        assert d1 <= 1e-12, f"{env}: max|div B| {d1}"


def test_fused_stage_fp32_32x32_tiles(monkeypatch):
    """The 32 x 32 tile shape (FP32 only; its LDS fills a CU) forms every face, corner and
    cell with the same per-value code as the 32 x 16 shape, over a different tiling:
    agreement to rounding (contraction may differ between the instantiations), on
    multi-tile, thin and open boxes with every solver; in an FP64 context the shape is not
    available and the default stays."""
    for case in CASES[:3]:
        say(f"[gpu] fp32 32x32 tiles vs 32x16 {case}")
        ref = run(monkeypatch, {"IMHD_VSHAPE": "1"}, *case, precision=32)
        got = run(monkeypatch, {"IMHD_VSHAPE": "2"}, *case, precision=32)
        for x, y in zip(ref[0][:2], got[0][:2]):
            assert np.max(np.abs(x - y)) <= 1e-5 * max(1.0, np.max(np.abs(x)))
        assert abs(ref[1][1][0] - got[1][1][0]) <= 1e-5 * ref[1][1][0]
        assert ref[2] == got[2]
    ref = run(monkeypatch, {}, *CASES[0])
    got = run(monkeypatch, {"IMHD_VSHAPE": "2"}, *CASES[0])
    for x, y in zip(ref[0], got[0]):
        if x is not None:
            assert np.array_equal(x, y)


def test_fused_fp32_lds_stencil_kernels_match_per_face_loads(monkeypatch):
    """The FP32 predictor and PLM corrector that take their stencils from LDS and lane
    shifts (the default) against the per-face-load kernels (IMHD_VSH=0), to rounding, with
    the corrector on 32 x 16 and on 16 x 16 tiles (IMHD_VSHAPE_C=0: three workgroups per
    CU), on a multi-tile periodic box and thin open ones with every solver; the
    floor-reset count and the dt sequence agree too."""
    for case in CASES[:3]:
        say(f"[gpu] fp32 LDS-stencil kernels vs per-face loads {case}")
        ref = run(monkeypatch, {"IMHD_VSH": "0"}, *case, precision=32)
        for env in ({}, {"IMHD_VSHAPE_C": "0"}):
            got = run(monkeypatch, env, *case, precision=32)
            for x, y in zip(ref[0][:2], got[0][:2]):
                assert np.max(np.abs(x - y)) <= 1e-5 * max(1.0, np.max(np.abs(x))), env
            assert abs(ref[1][1][0] - got[1][1][0]) <= 1e-5 * ref[1][1][0]
            assert ref[2] == got[2]
