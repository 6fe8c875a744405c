"""Host-side logic of the C ABI (no device needed): mesher, locate, RNG,
loader, and the library's exported symbols against include/scdat_amd.h."""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle as O
import scdat_amd as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
COND = ((0.4, 0.4, 0.4), (0.6, 0.6, 0.6))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "scdat_amd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(scdat_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    names = header_functions()
    assert len(names) >= 30
    out = subprocess.run(["nm", "-D", "--defined-only", S.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (scdat_[a-z0-9_]+)", out))
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    # and the Python binding declares a signature for each of them
    assert set(names) == set(S.SIGNATURES), set(names) ^ set(S.SIGNATURES)


def test_create_without_device_fails_cleanly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    with pytest.raises(S.ScdatError) as e:
        S.Context()
    assert e.value.code == S.ERR_DEVICE


def test_philox_known_answers():
    # Random123 kat_vectors for philox4x32_10
    assert S.philox4x32([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert S.philox4x32([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert S.philox4x32([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]) == \
        [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


@pytest.mark.parametrize("n,cond", [(4, None), (8, COND), ((5, 7, 9), COND), (16, COND)])
def test_box_mesh_is_conforming(n, cond):
    m = S.box_mesh(n, conductor=cond)
    xyz, tets = m["xyz"], m["tets"]
    o = O.Oracle(xyz, tets, m["kind"], group=m["group"])
    _, _, vol = o.geometry()
    expect = 1.0 - (0.2 ** 3 if cond else 0.0)
    assert abs(vol.sum() - expect) < 1e-12
    assert vol.min() > 0
    # positive orientation of every tet
    p = xyz[tets]
    det = np.einsum("ij,ij->i", p[:, 1] - p[:, 0], np.cross(p[:, 2] - p[:, 0], p[:, 3] - p[:, 0]))
    assert (det > 0).all()
    nb = o.neighbors()
    bnd = nb < 0
    # boundary faces carry a group, interior faces none
    assert (m["group"][bnd] >= 0).all() and (m["group"][~bnd] == -1).all()
    # the boundary faces lie on the planes their group names
    for t, k in zip(*np.nonzero(bnd)):
        g = m["group"][t, k]
        face = xyz[np.delete(tets[t], k)]
        ax = (g % 6) // 2
        vals = face[:, ax]
        assert np.ptp(vals) < 1e-12
        if g < 6:
            assert abs(vals[0] - (0.0 if g % 2 == 0 else 1.0)) < 1e-12
        else:
            assert abs(vals[0] - (0.4 if g % 2 == 0 else 0.6)) < 1e-12
    # node kinds: walls Dirichlet, conductor surface floating
    on_wall = np.any((xyz < 1e-12) | (xyz > 1 - 1e-12), axis=1)
    assert np.array_equal(m["kind"] == 1, on_wall)
    if cond:
        inside = np.all((xyz > 0.4 - 1e-12) & (xyz < 0.6 + 1e-12), axis=1)
        assert np.array_equal(m["kind"] == 2, inside)


def test_box_mesh_total_face_area_of_conductor():
    m = S.box_mesh(8, conductor=COND)
    xyz, tets = m["xyz"], m["tets"]
    area = 0.0
    for t, k in zip(*np.nonzero(m["group"] >= 6)):
        f = xyz[np.delete(tets[t], k)]
        area += 0.5 * np.linalg.norm(np.cross(f[1] - f[0], f[2] - f[0]))
    assert abs(area - 6 * 0.2 ** 2) < 1e-12


def test_box_locate_agrees_with_barycentric():
    m = S.box_mesh(7, conductor=COND)
    spec = S.box_spec(7, conductor=COND)
    x, v, cell = S.maxwellian_box(spec, 5000, 3, 0, 1.0)
    assert (cell >= 0).all()
    o = O.Oracle(m["xyz"], m["tets"], m["kind"], group=m["group"])
    lam = np.array([o.barycentric(int(c), p) for c, p in zip(cell, x)])
    assert lam.min() > -1e-12
    # points inside the conductor or outside the box are not located
    bad = np.array([[0.5, 0.5, 0.5], [1.5, 0.2, 0.2], [-0.1, 0.5, 0.5]])
    assert (S.loading.box_locate(spec, bad) == -1).all()


def test_maxwellian_loader_moments_and_determinism():
    spec = S.box_spec(8, conductor=COND)
    x, v, c = S.maxwellian_box(spec, 200000, 42, 0, 2.0)
    assert abs(v.mean()) < 0.02
    assert abs(v.std() - 2.0) < 0.01
    inside = np.all((x > 0.4) & (x < 0.6), axis=1)
    assert not inside.any()
    # uniform in the free volume: fraction in the lower half-box
    assert abs((x[:, 2] < 0.5).mean() - 0.5) < 0.01
    x2, v2, _ = S.maxwellian_box(spec, 1000, 42, 0, 2.0)
    assert np.array_equal(x2, x[:1000]) and np.array_equal(v2, v[:1000])
    x3, _, _ = S.maxwellian_box(spec, 1000, 42, 1, 2.0)
    assert not np.array_equal(x3, x[:1000])


def test_collision_process_defaults_are_the_reference_defaults():
    # MonteCarloCollisionHandler::initializeDefaultCollisionParameters
    # (CollisionAlgorithm.cpp:590-646); the same sets drive mcc_sigma.npz
    g = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "mcc_sigma.npz"))
    prm = g["prm"]
    cases = {S.MCC_ELASTIC: prm[0], S.MCC_INELASTIC: prm[1], S.MCC_CHARGE_EXCHANGE: prm[2],
             S.MCC_IONIZATION: prm[4], S.MCC_EXCITATION: prm[6], S.MCC_RECOMBINATION: prm[8]}
    for t, p in cases.items():
        d = S.mcc_process(t, 0)
        assert (d.threshold_ev, d.sigma_max_m2) == (p[0], p[1]), t
        if t in (S.MCC_ELASTIC, S.MCC_INELASTIC, S.MCC_CHARGE_EXCHANGE, S.MCC_RECOMBINATION):
            assert (d.reference_energy_ev, d.power_index) == (p[2], p[3]), t
    assert S.mcc_process(S.MCC_EXCITATION, 0).alpha == 1.2 and S.mcc_process(S.MCC_EXCITATION, 0).beta == 0.35
    assert S.mcc_process(S.MCC_IONIZATION, 0).energy_loss_ev == 15.76
    assert abs(S.mcc_process(S.MCC_COULOMB, 0).coulomb_log - g["coulomb_log"][0]) < 1e-12


def test_cpp_facade_builds_and_fails_loudly_without_a_device():
    # the facade links against the library; with no GPU the context cannot be
    # created and the program stops with the library's error, no CPU fallback
    here = os.path.dirname(os.path.abspath(__file__))
    subprocess.run(["make", "-s", "-C", os.path.join(here, "cpp")], check=True)
    if os.environ.get("HIP_VISIBLE_DEVICES") is None and os.path.exists("/dev/kfd"):
        pytest.skip("a GPU may be visible here")
    r = subprocess.run([os.path.join(here, "cpp", "picycle_facade_test")], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode != 0 and "scdat_create failed" in r.stderr
