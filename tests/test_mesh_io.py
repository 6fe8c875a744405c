"""Mesh import and result export (host, no GPU): the readers against what
the compiled reference's GmshLoader / VTKLoader / STLLoader read from the
same files (tests/golden/mesh/, made by tests/golden/gen_golden.py mesh),
boundary groups from tagged triangles, and the writers' formats."""
import json
import os

import numpy as np
import pytest

import scdat_amd as S
from scdat_amd import mesh_io as M

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "mesh")
REF = np.load(os.path.join(GOLD, "mesh_ref.npz"))


@pytest.mark.parametrize("name", ["box3_v2.msh", "box3_v41.msh", "box3.vtk", "tet.stl"])
def test_reader_matches_reference_loader(name):
    m = M.read_mesh(os.path.join(GOLD, name))
    key = name.replace(".", "_")
    xyz, et, en, tg = (REF[key + s] for s in ("_xyz", "_type", "_nodes", "_tag"))
    np.testing.assert_array_equal(m["xyz"], xyz)
    tet = et == 4
    tri = et == 3
    np.testing.assert_array_equal(m["tets"], en[tet])
    np.testing.assert_array_equal(m["tet_tag"], tg[tet])
    np.testing.assert_array_equal(m["tris"], en[tri][:, :3])
    np.testing.assert_array_equal(m["tri_tag"], tg[tri])
    assert len(m["tets"]) + len(m["tris"]) == len(et)


def test_gmsh_tags_give_the_box_boundary_groups():
    box = S.box_mesh(3, conductor=((1 / 3,) * 3, (2 / 3,) * 3))
    for name in ("box3_v2.msh", "box3_v41.msh"):
        m = M.read_mesh(os.path.join(GOLD, name), tag_groups={200 + g: g for g in range(12)}, default_group=15)
        np.testing.assert_array_equal(m["xyz"], box["xyz"])
        np.testing.assert_array_equal(m["tets"], box["tets"])
        np.testing.assert_array_equal(m["group"], box["group"])
        assert m["faces_matched"] == int((box["group"] >= 0).sum())
        assert np.all(m["tet_tag"] == 100)
    # an unmapped tag takes the default group
    m = M.read_mesh(os.path.join(GOLD, "box3_v2.msh"), tag_groups={200: 0}, default_group=9)
    assert set(np.unique(m["group"])) == {-1, 0, 9}


def test_vtk_writer_round_trip(tmp_path):
    box = S.box_mesh(4)
    n = len(box["xyz"])
    phi = np.sin(box["xyz"][:, 0]) + 0.5
    E = np.stack([box["xyz"][:, 1], -box["xyz"][:, 2], np.ones(n)], axis=1)
    p = tmp_path / "fields.vtk"
    M.write_vtk(p, box["xyz"], box["tets"], {"potential_v": phi, "electric_field_v_per_m": E})
    m = M.read_mesh(p)
    np.testing.assert_array_equal(m["xyz"], box["xyz"])
    np.testing.assert_array_equal(m["tets"], box["tets"])
    text = p.read_text()
    assert "SCALARS potential_v double 1" in text and "VECTORS electric_field_v_per_m double" in text
    tail = text.split("LOOKUP_TABLE default\n")[1].split("\n")[:n]
    np.testing.assert_array_equal(np.array(tail, float), phi)


def test_columnar_export_follows_the_output_contract(tmp_path):
    z = np.linspace(0, 1, 5)
    series = {"potential_v": z ** 2, "electron_density_m3": 1e9 * (1 + z), "ion_density_m3": 2e9 - z}
    csv = tmp_path / "profile.csv"
    M.export_columns(csv, "z_m", z, series)
    lines = csv.read_text().splitlines()
    assert lines[0] == "z_m,electron_density_m3,ion_density_m3,potential_v"  # series in name order
    row = [float(v) for v in lines[2].split(",")]
    np.testing.assert_allclose(row, [z[1], series["electron_density_m3"][1], series["ion_density_m3"][1],
                                     series["potential_v"][1]], rtol=1e-8)
    assert lines[1].split(",")[0] == "%.8e" % z[0]
    vtk = (tmp_path / "profile.vtk").read_text()
    assert "DATASET POLYDATA" in vtk and "POINTS 5 float" in vtk and "SCALARS potential_v float 1" in vtk
    h5 = (tmp_path / "profile.h5").read_text()
    assert h5.startswith("# SCDAT HDF5-like archive") and "[dataset:z_m]" in h5
    for art, fmt in (("profile.csv", "csv"), ("profile.vtk", "vtk"), ("profile.h5", "h5")):
        meta = json.loads((tmp_path / (art + ".metadata.json")).read_text())
        assert meta["artifact_format"] == fmt and meta["output_contract_schema"] == "scdat.output.contract.v1"
        md = meta["metadata"]
        assert md["output_row_count"] == "5" and md["output_series_count"] == "3"
        assert md["output_series_names"] == "electron_density_m3;ion_density_m3;potential_v"
        assert md["module"] == "PIC-MCC" and md["output_axis_name"] == "z_m"
    with pytest.raises(S.ScdatError):
        M.export_columns(tmp_path / "bad.csv", "z_m", z, {"x": np.array([0, 1, np.nan, 3, 4.0])})


def test_reader_errors_are_reported(tmp_path):
    with pytest.raises(S.ScdatError):
        M.read_mesh(tmp_path / "missing.msh")
    bad = tmp_path / "bad.msh"
    bad.write_text("$Nodes\n1\n1 0 0 0\n$EndNodes\n$Elements\n1\n1 4 0 1 2 3 4\n$EndElements\n")
    with pytest.raises(S.ScdatError):
        M.read_mesh(bad)
    with pytest.raises(S.ScdatError):
        M.read_mesh(tmp_path / "x.obj")
