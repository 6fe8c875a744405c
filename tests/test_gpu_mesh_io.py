"""Imported meshes drive the device cycle, and the layer profile of the PIC
harness (main_PICcore.cpp:230-310) from the device state."""
import os

import numpy as np
import pytest

import scdat_amd as S
from scdat_amd import mesh_io as M
from helpers import QE, plasma_setup

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "mesh")


def _ctx(xyz, tets, kind, group, st, precision=1):
    c = S.Context(precision=precision, dt=st["dt"], pcg_tol=1e-12, sort_interval=2)
    c.set_mesh(xyz, tets, kind, group=group)
    for g in range(6, 12):
        c.set_boundary(g, S.BC_CONDUCTOR)
    n = len(st["e"][0])
    c.add_species(-QE, S.ELECTRON_MASS, st["w"], 2 * n)
    c.add_species(QE, S.PROTON_MASS, st["w"], 2 * n)
    c.set_particles(0, *st["e"])
    c.set_particles(1, *st["i"])
    c.initialize()
    return c


def test_imported_gmsh_mesh_runs_like_the_box_mesh():
    cond = ((1 / 3,) * 3, (2 / 3,) * 3)
    box = S.box_mesh(3, conductor=cond)
    st = plasma_setup(n=3, npart=2000, seed=4)
    spec = S.box_spec(3, conductor=cond)
    for sp in ("e", "i"):
        st[sp] = (st[sp][0], st[sp][1], S.loading.box_locate(spec, st[sp][0]))
        keep = st[sp][2] >= 0
        st[sp] = tuple(a[keep] for a in st[sp])
    m = M.read_mesh(os.path.join(GOLD, "box3_v41.msh"), tag_groups={200 + g: g for g in range(12)})
    c1 = _ctx(box["xyz"], box["tets"], box["kind"], box["group"], st)
    c2 = _ctx(m["xyz"], m["tets"], box["kind"], m["group"], st)
    c1.step(3)
    c2.step(3)
    for a, b in zip(c1.fields(), c2.fields()):
        np.testing.assert_array_equal(a, b)
    assert c1.stats()["floating_potential"] == c2.stats()["floating_potential"]


def test_layer_profile_matches_its_definition(tmp_path):
    st = plasma_setup(n=6, npart=5000, seed=9, Te_eV=2.0)
    m = st["mesh"]
    c = _ctx(m["xyz"], m["tets"], m["kind"], m["group"], st, precision=0)
    c.step(2)
    prof = c.layer_profile(axis=2, length=1.0, n_cells=6, area=1.0)
    rho, phi, E = c.fields()
    z = np.round(m["xyz"][:, 2] * 1e9) * 1e-9
    layers = np.unique(z)
    np.testing.assert_allclose(prof["position"], layers)
    np.testing.assert_allclose(prof["potential_v"], [phi[z == L].mean() for L in layers], rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(prof["electric_field_v_per_m"], [E[z == L, 2].mean() for L in layers],
                               rtol=1e-12, atol=1e-12)
    dz, vol = 1.0 / 6, 1.0 / 6
    ne, ni, acc = np.zeros(len(layers)), np.zeros(len(layers)), np.zeros(len(layers))
    for s, dens in ((0, ne), (1, ni)):
        x, v, cell, _ = c.get_particles(s)
        a = cell >= 0
        b = np.clip(np.rint(np.clip(x[a, 2], 0, 1) / dz), 0, len(layers) - 1).astype(int)
        np.add.at(dens, b, st["w"] / vol)
        if s == 0:
            ke = 0.5 * S.ELECTRON_MASS * (v[a] ** 2).sum(1) / QE
            np.add.at(acc, b, st["w"] * ke)
    np.testing.assert_allclose(prof["electron_density_m3"], ne, rtol=1e-12)
    np.testing.assert_allclose(prof["ion_density_m3"], ni, rtol=1e-12)
    te = np.where(ne > 0, acc / np.maximum(1.0, ne * vol) / 1.5, 0.0)
    np.testing.assert_allclose(prof["electron_temperature_ev"], te, rtol=1e-9)
    assert 0.5 < np.median(te[te > 0]) < 4.0  # a 2 eV plasma
    series = {k: v for k, v in prof.items() if k != "position"}
    M.export_columns(tmp_path / "layers.csv", "z_m", prof["position"], series)
    M.write_vtk(tmp_path / "fields.vtk", m["xyz"], m["tets"], {"potential_v": phi, "electric_field_v_per_m": E})
    assert (tmp_path / "layers.csv").exists() and (tmp_path / "fields.vtk").exists()
