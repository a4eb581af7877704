"""The compiled reference's residual on a mesh built from arrays (ref_mesh_new: the entry through which
bench.py times the reference's CPU path on the benchmark meshes) against the same residual on the
reference's own generated mesh, bitwise.  Runs where oracle/_ref/libaero_ref.so is built; skipped
elsewhere."""
import ctypes as C

import numpy as np
import pytest

import cases
import helpers
import oracle_lib as ol


def _generated(ref, name):
    m = cases.MESHES[name]
    cfg = ol.RefMeshCfg(m["num_circumferential"], m["num_radial"], m["farfield_radius"], m["radial_stretch"], 1.0,
                        b"0012")
    n = [C.c_int() for _ in range(7)]
    ref.ref_mesh_sizes(C.byref(cfg), *[C.byref(x) for x in n])
    N, F, P, ws, wc, _, _ = [x.value for x in n]
    a = dict(owner=np.zeros(F, np.int32), neighbor=np.zeros(F, np.int32), bc=np.zeros(F, np.int32),
             face_vertices=np.zeros(2 * F, np.int32), normal=np.zeros(3 * F, np.float32),
             face_center=np.zeros(3 * F, np.float32), area=np.zeros(F, np.float32), volume=np.zeros(N, np.float32),
             cell_center=np.zeros(3 * N, np.float32), points=np.zeros(3 * P, np.float32))
    ref.ref_mesh_fill(C.byref(cfg), *[ol.ptr(a[k]) for k in ("owner", "neighbor", "bc", "face_vertices", "normal",
                                                               "face_center", "area", "volume", "cell_center",
                                                               "points")])
    return cfg, N, F, ws, wc, a


@pytest.mark.parametrize("name", ["40x8", "72x18"])
def test_reference_residual_on_array_mesh(name):
    ref = ol.ref()
    if ref is None:
        pytest.skip("the compiled reference (oracle/_ref) is not built here")
    mcfg, N, F, ws, wc, a = _generated(ref, name)
    from aero_hip import freestream
    ff = np.asarray(freestream(0.5, 1.25)[1], np.float32)
    U = helpers.perturbed_state(a["cell_center"].reshape(N, 3).astype(np.float64), ff.astype(np.float64), 0.05,
                                11).astype(np.float32)
    with ol.RefMesh(ref, N, F, a["owner"], a["neighbor"], a["normal"], a["face_center"], a["area"], a["volume"],
                    a["cell_center"], ws, wc) as rm:
        for combo in [(1, 1, 1, 0, 0, 0, 0), (0, 0, 1, 1, 0, 0, 0), (1, 0, 0, 0, 1, 1, 1)]:
            cfg = ol.ref_residual_cfg(cases.residual_cfg_values(combo, [float(x) for x in ff]))
            R1, l1, d1 = rm.residual(cfg, U)
            R2 = np.zeros((N, 5), np.float32)
            l2 = np.zeros(N, np.float32)
            d2 = np.zeros(3, np.int32)
            ref.ref_residual(C.byref(mcfg), C.byref(cfg), ol.ptr(U), ol.ptr(R2), ol.ptr(l2), ol.ptr(d2))
            assert np.array_equal(R1, R2) and np.array_equal(l1, l2) and np.array_equal(d1, d2), combo
            assert np.all(np.isfinite(R1))
