"""The round-end smoke entry point (__graft_entry__.smoke): one tiny drop-in VL2 + CT step
with a passive scalar on the GPU, bit for bit against the oracle."""
import os
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_smoke_entry_point():
    sys.path.insert(0, ROOT)
    import __graft_entry__ as g
    g.smoke()
