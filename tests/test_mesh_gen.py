"""The native O-grid generator: FP32 triangles reproduce make_airfoil_ogrid_mesh exactly; the FP64 quad,
mixed, jittered and RCM-renumbered meshes are closed, positively oriented and deterministic."""
import hashlib

import numpy as np
import pytest

import cases
import helpers


def digest(m):
    h = hashlib.sha256()
    for a in (m.owner, m.neighbor, m.bc, m.face_vertices.ravel(), m.normal.ravel(), m.face_center.ravel(), m.area,
              m.volume, m.cell_center.ravel(), m.points.ravel()):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def test_fp32_triangles_match_reference_arrays(golden):
    g = golden("mesh")
    m = helpers.mesh("40x8", precision=4)
    for k, a in (("owner", m.owner), ("neighbor", m.neighbor), ("bc", m.bc), ("face_vertices", m.face_vertices),
                 ("normal", m.normal), ("face_center", m.face_center), ("area", m.area), ("volume", m.volume),
                 ("cell_center", m.cell_center), ("points", m.points)):
        assert np.array_equal(a.ravel(), g[f"40x8_{k}"]), k


@pytest.mark.parametrize("name", list(cases.MESHES))
def test_fp32_triangles_match_reference_digests(golden, name):
    g = golden("mesh")
    m = helpers.mesh(name, precision=4)
    sizes = g[f"sizes_{name}"]
    assert [m.num_cells, m.num_faces, m.wall_start, m.wall_count, m.farfield_start, m.farfield_count] == list(sizes)
    assert digest(m) == str(g[f"digest_{name}"])


def closure(m):
    """Per cell sum of n A over its faces (outward from the cell): zero for a closed polygon."""
    s = np.zeros((m.num_cells, 2))
    nA = m.normal[:, :2] * m.area[:, None]
    np.add.at(s, m.owner, nA)
    inner = m.neighbor >= 0
    np.add.at(s, m.neighbor[inner], -nA[inner])
    return s


@pytest.mark.parametrize("kw", [dict(cells="tri"), dict(cells="quad"),
                                dict(cells="mixed", quad_layers=5, jitter=0.1, seed=3, rcm=True)])
def test_fp64_meshes_are_closed_and_oriented(kw):
    m = helpers.mesh("64x14", **kw)
    assert np.all(m.volume > 0)
    assert np.abs(np.linalg.norm(m.normal, axis=1) - 1.0).max() < 1e-14
    perim = np.zeros(m.num_cells)
    np.add.at(perim, m.owner, m.area)
    np.add.at(perim, m.neighbor[m.neighbor >= 0], m.area[m.neighbor >= 0])
    assert np.abs(closure(m)).max() <= 1e-13 * perim.max()
    # normals point out of the owner: from owner centre to face centre
    d = m.face_center[:, :2] - m.cell_center[m.owner, :2]
    assert np.all((d * m.normal[:, :2]).sum(1) > 0)
    # face ordering: interior, then wall, then farfield (airfoil_mesh.cpp:287-369)
    assert np.all(m.neighbor[: m.wall_start] >= 0) and np.all(m.neighbor[m.wall_start:] == -1)
    assert np.all(m.bc[m.wall_start: m.wall_start + m.wall_count] == 2)
    assert np.all(m.bc[m.farfield_start: m.farfield_start + m.farfield_count] == 1)
    assert m.wall_count == m.farfield_count == 64


def test_cell_counts():
    nt, nr = 64, 14
    assert helpers.mesh("64x14", cells="tri").num_cells == 2 * nt * nr
    q = helpers.mesh("64x14", cells="quad")
    assert q.num_cells == nt * nr and q.num_faces == nt * nr * 2 + nt
    mx = helpers.mesh("64x14", cells="mixed", quad_layers=4)
    assert mx.num_cells == nt * 4 + 2 * nt * (nr - 4)


def test_jitter_and_rcm_are_deterministic_permutations():
    a = helpers.mesh("64x14", cells="mixed", quad_layers=4, jitter=0.1, seed=7, rcm=True)
    b = helpers.mesh("64x14", cells="mixed", quad_layers=4, jitter=0.1, seed=7, rcm=True)
    c = helpers.mesh("64x14", cells="mixed", quad_layers=4, jitter=0.1, seed=7, rcm=False)
    assert digest(a) == digest(b)
    # RCM reorders cells only: the multiset of cell volumes and the total area are unchanged
    assert np.array_equal(np.sort(a.volume), np.sort(c.volume))
    # and it reduces the cell-graph bandwidth
    bw = lambda m: np.abs(m.owner[m.neighbor >= 0] - m.neighbor[m.neighbor >= 0]).max()
    assert bw(a) < bw(c)
    # jitter moves interior nodes only, by at most 10% of local spacing
    d = helpers.mesh("64x14", cells="mixed", quad_layers=4, jitter=0.0)
    moved = np.linalg.norm(c.points - d.points, axis=1)
    ring = lambda j: slice(j * 64, (j + 1) * 64)
    assert moved[ring(0)].max() == 0 and moved[ring(14)].max() == 0
    assert moved[ring(5)].max() > 0
