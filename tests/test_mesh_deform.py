"""Hicks-Henne + IDW shape deformation (no reference counterpart): pinned by an independent numpy
restatement of the bump and IDW definitions, zero-amplitude identity, exact wall displacement,
fixed farfield, and central differences of the geometry for the analytic tangents."""
import numpy as np
import pytest

import helpers


def _base():
    return helpers.mesh("64x14", cells="mixed", quad_layers=4, jitter=0.1, seed=3, rcm=True)


def _numpy_deform(m, amp, n=10, h_min=0.05, h_max=0.95, k=3.0):
    """Restatement of the definitions: bumps sin^3(pi x^(ln .5/ln h)), upper / lower surface by the
    outward normal, IDW over wall + farfield nodes with weights |r|^-k."""
    pts = m.points[:, :2].astype(np.float64)
    fv = m.face_vertices
    wall_f = np.arange(m.wall_start, m.wall_start + m.wall_count)
    far_f = np.arange(m.farfield_start, m.farfield_start + m.farfield_count)
    wall_nodes = np.unique(fv[wall_f].ravel())
    far_nodes = np.unique(fv[far_f].ravel())
    ny_sum = np.zeros(len(pts))
    for f in wall_f:
        ny_sum[fv[f]] += m.normal[f, 1]
    xw = pts[wall_nodes, 0]
    x = np.clip((xw - xw.min()) / (xw.max() - xw.min()), 0.0, 1.0)
    upper = -ny_sum[wall_nodes] >= 0.0
    hs = np.linspace(h_min, h_max, n)
    dy_wall = np.zeros(len(wall_nodes))
    for i, h in enumerate(hs):
        b = np.sin(np.pi * x ** (np.log(0.5) / np.log(h))) ** 3
        dy_wall += np.where(upper, amp[i], amp[n + i]) * b
    src = np.concatenate([wall_nodes, far_nodes])
    src_dy = np.concatenate([dy_wall, np.zeros(len(far_nodes))])
    out = pts.copy()
    is_src = np.zeros(len(pts), bool)
    is_src[src] = True
    out[wall_nodes, 1] += dy_wall
    q = np.where(~is_src)[0]
    d2 = ((pts[q, None, :] - pts[None, src, :]) ** 2).sum(axis=2)
    w = d2 ** (-0.5 * k)
    out[q, 1] += (w * src_dy[None, :]).sum(axis=1) / w.sum(axis=1)
    return out


def test_zero_amplitude_is_identity():
    from aero_hip import deform_mesh
    m = _base()
    d, _ = deform_mesh(m, np.zeros(20), tangents=False)
    for k in ("normal", "face_center", "area", "cell_center", "points"):
        assert np.array_equal(getattr(d, k), getattr(m, k)), k
    assert np.abs(d.volume - m.volume).max() <= 1e-15 * np.abs(m.volume).max()


def test_matches_numpy_restatement():
    from aero_hip import deform_mesh
    m = _base()
    rng = np.random.default_rng(4242)
    amp = rng.uniform(-0.002, 0.002, 20)
    d, _ = deform_mesh(m, amp, tangents=False)
    ref = _numpy_deform(m, amp)
    assert np.abs(d.points[:, :2] - ref).max() <= 1e-13
    far = np.unique(m.face_vertices[m.farfield_start:m.farfield_start + m.farfield_count].ravel())
    assert np.array_equal(d.points[far], m.points[far])
    assert np.array_equal(d.points[:, 0], m.points[:, 0])  # Hicks-Henne moves nodes in y only
    assert np.all(d.volume > 0)


def test_geometry_tangents_match_central_difference():
    from aero_hip import deform_mesh
    m = _base()
    rng = np.random.default_rng(1)
    amp = rng.uniform(-0.002, 0.002, 20)
    _, tan = deform_mesh(m, amp)
    h = 1e-6
    for p in (0, 4, 9, 10, 15, 19):
        e = np.zeros(20)
        e[p] = h
        dp, _ = deform_mesh(m, amp + e, tangents=False)
        dm, _ = deform_mesh(m, amp - e, tangents=False)
        for k, key in (("normal", "normal"), ("face_center", "face_center"), ("area", "area"),
                       ("cell_center", "cell_center"), ("volume", "volume"), ("points", "points")):
            fd = (getattr(dp, k) - getattr(dm, k)) / (2 * h)
            scale = max(np.abs(fd).max(), 1e-30)
            assert np.abs(tan[key][p] - fd).max() <= 1e-6 * scale, (p, k)


def test_bumps_are_local_and_surface_specific():
    from aero_hip import deform_mesh
    m = _base()
    _, tan = deform_mesh(m, np.zeros(20))
    wall = np.unique(m.face_vertices[m.wall_start:m.wall_start + m.wall_count].ravel())
    y = m.points[wall, 1]
    up, lo = wall[y > 1e-9], wall[y < -1e-9]
    for i in range(10):
        assert np.all(tan["points"][i][lo, 1] == 0.0) and np.abs(tan["points"][i][up, 1]).max() > 0.3
        assert np.all(tan["points"][10 + i][up, 1] == 0.0) and np.abs(tan["points"][10 + i][lo, 1]).max() > 0.3
        assert tan["points"][i][:, 1].max() <= 1.0 + 1e-12


def test_inverting_deformation_is_rejected():
    from aero_hip import deform_mesh, _lib
    m = _base()
    amp = np.zeros(20)
    amp[5] = -0.5  # pushes the upper surface half a chord down through the lower one
    with pytest.raises(_lib.HipError):
        deform_mesh(m, amp, tangents=False)
