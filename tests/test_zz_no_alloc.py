"""imhd_diagnostics_static, the entry point that allocates nothing, called through the C
ABI: its output equals imhd_diagnostics's bit for bit (same kernels, same fixed reduction
order) on two calls in a row -- a three-box ensemble, then a smaller single box, so the
second call reuses the fixed device scratch the first one filled -- and it allocates no
device memory (imhd_device_bytes unchanged, where imhd_diagnostics's first call adds its
partial rows).  A context above the 64-box limit is refused with a clear error."""
import ctypes
import sys

import numpy as np
import pytest

from refutil import random_state

pytestmark = pytest.mark.gpu

FIELDS = ("ke", "vrms_sq", "momx", "momy", "momz", "divb_sum", "divb_max", "mass")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def c_abi_diag(S, static):
    from isothermal_mhd_amd import _lib as L
    out = (L.ImhdDiag * len(S.boxes))()
    fn = S.lib.imhd_diagnostics_static if static else S.lib.imhd_diagnostics
    rc = fn(S._c, out)
    return rc, [{n: getattr(d, n) for n in FIELDS} for d in out]


def device_bytes(S):
    n = ctypes.c_ulonglong()
    assert S.lib.imhd_device_bytes(S._c, ctypes.byref(n)) == 0
    return n.value


@pytest.mark.parametrize("precision", [64, 32])
def test_static_diagnostics_equal_allocating_ones(precision):
    from isothermal_mhd_amd import VanLeerSolver
    cases = [[(20, 14, 9), (33, 17, 21), (12, 12, 40)], [(10, 8, 6)]]   # the second call is smaller
    for n, boxes in enumerate(cases):
        say(f"[gpu] imhd_diagnostics_static fp{precision}, call {n + 1}: {boxes}")
        S = VanLeerSolver(boxes if len(boxes) > 1 else boxes[0], npass=0, riemann="hlld", bc="periodic",
                          precision=precision)
        for b, d in enumerate(boxes):
            q, bf, _ = random_state(*d, npass=0, seed=90 + 10 * n + b)
            S.set_state(q, bf, None, box=b)
        S.boundary_conditions()
        S.run(2, graph=False)
        S.sync()
        b0 = device_bytes(S)
        rc, st = c_abi_diag(S, True)
        assert rc == 0
        assert device_bytes(S) == b0             # nothing allocated
        rc, al = c_abi_diag(S, False)
        assert rc == 0
        assert device_bytes(S) > b0              # the allocating entry point's partial rows
        for b in range(len(boxes)):
            for k in FIELDS:
                assert st[b][k] == al[b][k], (n, b, k, st[b][k], al[b][k])
        rc, st2 = c_abi_diag(S, True)            # again, after the allocating call
        assert rc == 0 and st2 == st
        assert np.isfinite([st[0][k] for k in FIELDS]).all() and st[0]["mass"] > 0
        S.close()


def test_static_diagnostics_box_limit():
    from isothermal_mhd_amd import VanLeerSolver
    boxes = [(8, 8, 8)] * 65
    say("[gpu] imhd_diagnostics_static above the box limit")
    S = VanLeerSolver(boxes, npass=0, riemann="hlle", bc="periodic", precision=64)
    for b, d in enumerate(boxes):
        q, bf, _ = random_state(*d, npass=0, seed=b)
        S.set_state(q, bf, None, box=b)
    rc, _ = c_abi_diag(S, True)
    assert rc == -1
    assert b"64 boxes" in S.lib.imhd_last_error(S._c)
    rc, _ = c_abi_diag(S, False)
    assert rc == 0
    S.close()
