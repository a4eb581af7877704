"""aero_hip_iterate_noalloc (the allocation-free variant of aero_hip_iterate, its history sums in a static
device array) against aero_hip_iterate, through the C ABI: two calls in a row, the second on a smaller
mesh with fewer iterations, give the same state and density-residual history bit for bit; a history
longer than the static array is rejected before any GPU work."""
import ctypes as C
import sys

import numpy as np
import pytest

import cases
import helpers
from aero_hip import _lib

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def call(hm, name, cfg, n, rk, cfl, hist):
    err = _lib.errbuf()
    rc = getattr(_lib.lib(), name)(hm.h, C.byref(cfg), n, rk, cfl,
                                   hist.ctypes.data_as(C.POINTER(C.c_double)) if hist is not None else None,
                                   err, len(err))
    return rc, err.value.decode()


@pytest.fixture(scope="module")
def hip():
    import aero_hip
    assert aero_hip.hip_available()
    return aero_hip


def test_noalloc_iterate_matches_iterate(hip):
    from aero_hip import freestream
    runs = [("64x14", dict(cells="mixed", quad_layers=4, jitter=0.1, seed=3, rcm=True), 40, 3, 0.9),
            ("40x8", {}, 9, 1, 0.5)]  # the second call: a smaller mesh, fewer iterations, RK1
    for name, kw, n, rk, cfl in runs:
        m = helpers.mesh(name, **kw)
        ff = freestream(0.5, 2.0)[1]
        U = helpers.perturbed_state(m.cell_center, ff, 0.05, 17)
        cfg = helpers.res_cfg(cases.residual_cfg_values((1, 1, 1, 0, 0, 0, 0), ff))
        say(f"[noalloc] {name} n={n} rk={rk}")
        with hip.HipMesh(m) as a, hip.HipMesh(m) as b:
            a.set_state(U)
            b.set_state(U)
            ha, hb = np.zeros(n), np.full(n, np.nan)
            assert call(a, "aero_hip_iterate", cfg, n, rk, cfl, ha)[0] == 0
            rc, msg = call(b, "aero_hip_iterate_noalloc", cfg, n, rk, cfl, hb)
            assert rc == 0, msg
            assert np.array_equal(ha, hb), "density-residual history differs"
            assert np.array_equal(a.state(), b.state()), "state differs"
            assert np.all(np.isfinite(hb))


def test_noalloc_rejects_long_history(hip):
    from aero_hip import freestream
    m = helpers.mesh("40x8")
    ff = freestream(0.5, 2.0)[1]
    cfg = helpers.res_cfg(cases.residual_cfg_values((1, 1, 1, 0, 0, 0, 0), ff))
    with hip.HipMesh(m) as hm:
        n = 65536 + 1
        rc, msg = call(hm, "aero_hip_iterate_noalloc", cfg, n, 3, 0.5, np.zeros(n))
        assert rc == _lib.AERO_EINVAL and "exceeds" in msg, (rc, msg)
