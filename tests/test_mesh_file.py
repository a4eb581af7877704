"""Binary mesh files (aero_mesh_write / aero_mesh_read): exact round trips at both precisions, and
rejection of truncated, foreign and corrupted files (the reader's checks keep out-of-range indices away
from the device kernels)."""
import numpy as np
import pytest

from aero_hip import generate_mesh, load_mesh, save_mesh

FIELDS = ("owner", "neighbor", "bc", "face_vertices", "normal", "face_center", "area", "volume", "cell_center",
          "points", "cell_offsets", "cell_connectivity")


@pytest.mark.parametrize("precision,cells", [(8, "mixed"), (4, "tri"), (8, "quad")])
def test_round_trip(tmp_path, precision, cells):
    m = generate_mesh(num_circumferential=64, num_radial=14, farfield_radius=12.0, radial_stretch=1.35, cells=cells,
                      quad_layers=4, jitter=0.1 if cells == "mixed" else 0.0, rcm=cells == "mixed",
                      precision=precision)
    path = tmp_path / "m.amsh"
    save_mesh(m, path)
    r = load_mesh(path)
    assert (r.num_cells, r.num_faces) == (m.num_cells, m.num_faces)
    assert (r.wall_start, r.wall_count, r.farfield_start, r.farfield_count) == \
        (m.wall_start, m.wall_count, m.farfield_start, m.farfield_count)
    for k in FIELDS:
        a, b = np.asarray(getattr(m, k)), np.asarray(getattr(r, k))
        assert a.dtype == b.dtype and np.array_equal(a, b), k


def test_rejects_bad_files(tmp_path):
    m = generate_mesh(num_circumferential=40, num_radial=8, farfield_radius=10.0, radial_stretch=1.25)
    good = tmp_path / "good.amsh"
    save_mesh(m, good)
    raw = good.read_bytes()
    cases = {"truncated": raw[:-8], "trailing": raw + b"\0", "foreign": b"NOTAMESH" + raw[8:], "empty": b""}
    # an owner index past the last cell: the first face_owner entry sits right after the 8 + 36 byte head
    bad_owner = bytearray(raw)
    bad_owner[44:48] = np.int32(m.num_cells).tobytes()
    cases["owner out of range"] = bytes(bad_owner)
    for name, data in cases.items():
        p = tmp_path / f"{name.replace(' ', '_')}.amsh"
        p.write_bytes(data)
        with pytest.raises(RuntimeError):
            load_mesh(p)
    with pytest.raises(RuntimeError):
        load_mesh(tmp_path / "missing.amsh")


def test_cli_writes_the_case_mesh(tmp_path):
    """`python -m aero_hip mesh CASE FILE` writes the O-grid the runner builds for the case; the runner's
    case_and_mesh reads it back identically through mesh_path."""
    from aero_hip.__main__ import main
    from aero_hip.case_file import load_case
    from aero_hip.outputs import case_and_mesh
    case = tmp_path / "native_case.cfg"
    case.write_text("case_type=euler_airfoil_2d\nnum_circumferential=72\nnum_radial=18\nfarfield_radius=12\n"
                    "radial_stretch=1.35\nmach=0.5\naoa_deg=2\n")
    path = tmp_path / "case.amsh"
    assert main(["mesh", str(case), str(path)]) == 0
    kv = load_case(case)
    c1, m1 = case_and_mesh(kv)
    c2, m2 = case_and_mesh(kv, path)
    assert c1 == c2
    for k in FIELDS:
        assert np.array_equal(np.asarray(getattr(m1, k)), np.asarray(getattr(m2, k))), k
