"""ctypes binding of oracle/_ref/libscdat_ref.so: the reference's own
routines compiled from its sources plus ref_harness*.cpp.

TEST INFRASTRUCTURE ONLY: used in this development container to generate the
golden vectors under tests/golden/ that pin the oracle.  The GPU path never
loads it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "_ref", "libscdat_ref.so")
_D = C.POINTER(C.c_double)
_I32 = C.POINTER(C.c_int32)
_I64 = C.POINTER(C.c_int64)
_I = C.POINTER(C.c_int32)
_lib = None


def available():
    return os.path.exists(LIB) or os.path.isdir("/root/reference")


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            subprocess.run(["make", "-s", "-C", os.path.join(_HERE, "ref"), "-j8"], check=True)
        L = C.CDLL(LIB)
        L.ref_barycentric.argtypes = [C.c_int, _D, C.c_int, _I, C.c_int, _I, _D, _D, _I]
        L.ref_deposit.argtypes = [C.c_int, _D, C.c_int, _I, C.c_int, _D, _I, C.c_double, C.c_double, _D]
        L.ref_cross.argtypes = [C.c_int, _D, C.c_int, _I, _D, C.c_int, _D, _D, C.c_int, C.c_int, C.c_int,
                                _D, _I, C.c_double, C.c_double, C.c_double, C.c_double, _D, _I, _I, _D]
        L.ref_multiboundary.argtypes = [C.c_int, _D, C.c_int, _I, _I, _D, _D, C.c_double, C.c_double,
                                        C.c_int, C.c_double, _D]
        L.ref_poisson.argtypes = [C.c_int, _D, C.c_int, _I, _I, _D, _D, _D, _D]
        _lib = L
    return _lib


def _p(a, t):
    return None if a is None else a.ctypes.data_as(C.POINTER(t))


def _mesh(xyz, tets):
    return np.ascontiguousarray(xyz, np.float64), np.ascontiguousarray(tets, np.int32)


def barycentric(xyz, tets, cells, pts):
    xyz, tets = _mesh(xyz, tets)
    cells = np.ascontiguousarray(cells, np.int32)
    pts = np.ascontiguousarray(pts, np.float64)
    lam = np.zeros((len(cells), 4))
    ins = np.zeros(len(cells), np.int32)
    lib().ref_barycentric(len(xyz), _p(xyz, C.c_double), len(tets), _p(tets, C.c_int32), len(cells),
                          _p(cells, C.c_int32), _p(pts, C.c_double), _p(lam, C.c_double), _p(ins, C.c_int32))
    return lam, ins


def deposit(xyz, tets, pos, cells, charge, weight):
    xyz, tets = _mesh(xyz, tets)
    pos = np.ascontiguousarray(pos, np.float64)
    cells = np.ascontiguousarray(cells, np.int32)
    rho = np.zeros(len(xyz))
    lib().ref_deposit(len(xyz), _p(xyz, C.c_double), len(tets), _p(tets, C.c_int32), len(cells),
                      _p(pos, C.c_double), _p(cells, C.c_int32), charge, weight, _p(rho, C.c_double))
    return rho


def cross(xyz, tets, state, mass, charge, weight, dt, nodal_E=None, uniform_E=None, uniform_B=None,
          kernel=0, segments=10, hint=None):
    xyz, tets = _mesh(xyz, tets)
    state = np.ascontiguousarray(state, np.float64).reshape(-1, 6)
    mode = 1 if uniform_E is not None else 0
    E = None if nodal_E is None else np.ascontiguousarray(nodal_E, np.float64)
    uE = np.ascontiguousarray(uniform_E if uniform_E is not None else [0, 0, 0], np.float64)
    uB = np.ascontiguousarray(uniform_B if uniform_B is not None else [0, 0, 0], np.float64)
    h = None if hint is None else np.ascontiguousarray(hint, np.int32)
    out = np.zeros_like(state)
    el = np.zeros(len(state), np.int32)
    lost = np.zeros(len(state), np.int32)
    rho = np.zeros(len(xyz))
    lib().ref_cross(len(xyz), _p(xyz, C.c_double), len(tets), _p(tets, C.c_int32), _p(E, C.c_double), mode,
                    _p(uE, C.c_double), _p(uB, C.c_double), kernel, segments, len(state),
                    _p(state, C.c_double), _p(h, C.c_int32), mass, charge, weight, dt,
                    _p(out, C.c_double), _p(el, C.c_int32), _p(lost, C.c_int32), _p(rho, C.c_double))
    return out, el, lost, rho


def multiboundary(xyz, tets, kind, dval, density, q_float, permittivity, solver=3, tol=1e-12):
    xyz, tets = _mesh(xyz, tets)
    kind = np.ascontiguousarray(kind, np.int32)
    dv = None if dval is None else np.ascontiguousarray(dval, np.float64)
    de = None if density is None else np.ascontiguousarray(density, np.float64)
    phi = np.zeros(len(xyz))
    rc = lib().ref_multiboundary(len(xyz), _p(xyz, C.c_double), len(tets), _p(tets, C.c_int32),
                                 _p(kind, C.c_int32), _p(dv, C.c_double), _p(de, C.c_double), q_float,
                                 permittivity, solver, tol, _p(phi, C.c_double))
    return phi, rc


def poisson(xyz, tets, kind, dval, density):
    xyz, tets = _mesh(xyz, tets)
    kind = np.ascontiguousarray(kind, np.int32)
    dv = None if dval is None else np.ascontiguousarray(dval, np.float64)
    de = None if density is None else np.ascontiguousarray(density, np.float64)
    phi = np.zeros(len(xyz))
    E = np.zeros((len(xyz), 3))
    rc = lib().ref_poisson(len(xyz), _p(xyz, C.c_double), len(tets), _p(tets, C.c_int32), _p(kind, C.c_int32),
                           _p(dv, C.c_double), _p(de, C.c_double), _p(phi, C.c_double), _p(E, C.c_double))
    return phi, E, rc


SURF_LIB = os.path.join(_HERE, "_ref", "libscdat_ref_surface.so")
_slib = None


def surface_lib():
    """Cross sections and surface yields of the reference (a second library,
    linked with --gc-sections; see oracle/ref/Makefile)."""
    global _slib
    if _slib is None:
        if not os.path.exists(SURF_LIB):
            subprocess.run(["make", "-s", "-C", os.path.join(_HERE, "ref"), "-j8"], check=True)
        L = C.CDLL(SURF_LIB)
        L.ref_sigma.argtypes = [C.c_int, _D, C.c_int64, _D, _D]
        L.ref_coulomb_log.argtypes = [C.c_double, C.c_double]
        L.ref_coulomb_log.restype = C.c_double
        L.ref_surface.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double, C.c_int64, _D, _D, _D,
                                  C.c_double, C.c_double, _D, _D]
        L.ref_circuit_advance.argtypes = [C.c_int, _D, C.c_int, _D, C.c_double, _D, _D, _D, _D, C.c_int, _D,
                                          C.c_double, _D, _D, _D]
        _slib = L
    return _slib


def sigma(kind, prm, e_ev):
    """kind: 0 elastic, 1 ionisation, 2 excitation, 3 recombination, 4 Coulomb;
    prm: threshold, sigma_max, reference_energy, power_index, alpha, beta,
    electron_density, electron_temperature_K."""
    e = np.ascontiguousarray(e_ev, np.float64)
    p = np.ascontiguousarray(prm, np.float64)
    out = np.zeros_like(e)
    surface_lib().ref_sigma(kind, _p(p, C.c_double), len(e), _p(e, C.c_double), _p(out, C.c_double))
    return out


def coulomb_log(temperature_k, density_m3):
    return surface_lib().ref_coulomb_log(temperature_k, density_m3)


def surface(see_max, see_peak_ev, photo_yield, photo_temp_ev, e_ev, angle_rad, surf_v, ref_v, source_j):
    e = np.ascontiguousarray(e_ev, np.float64)
    a = np.ascontiguousarray(angle_rad, np.float64)
    v = np.ascontiguousarray(surf_v, np.float64)
    y, j = np.zeros_like(e), np.zeros_like(e)
    surface_lib().ref_surface(see_max, see_peak_ev, photo_yield, photo_temp_ev, len(e), _p(e, C.c_double),
                              _p(a, C.c_double), _p(v, C.c_double), ref_v, source_j, _p(y, C.c_double),
                              _p(j, C.c_double))
    return y, j


def circuit_advance(nodes, branches, dt, current, didv, rhs=None, diag=None, entries=(), max_delta=np.inf):
    """One SurfaceCircuitCoupling::advanceImplicit.  nodes: (v, C, G, fixed, fixed_v)
    rows; branches: (from, to, G, bias) rows; entries: (row, col, coef) rows."""
    nd = np.ascontiguousarray(nodes, np.float64).reshape(-1, 5)
    br = np.ascontiguousarray(branches, np.float64).reshape(-1, 4)
    en = np.ascontiguousarray(entries, np.float64).reshape(-1, 3)
    n = len(nd)
    z = np.zeros(n)
    cur = np.ascontiguousarray(current, np.float64)
    dv = np.ascontiguousarray(didv, np.float64)
    rh = np.ascontiguousarray(z if rhs is None else rhs, np.float64)
    dg = np.ascontiguousarray(z if diag is None else diag, np.float64)
    v, i, b = np.zeros(n), np.zeros(n), np.zeros(max(1, len(br)))
    rc = surface_lib().ref_circuit_advance(n, _p(nd, C.c_double), len(br), _p(br, C.c_double), dt,
                                           _p(cur, C.c_double), _p(dv, C.c_double), _p(rh, C.c_double),
                                           _p(dg, C.c_double), len(en), _p(en, C.c_double), max_delta,
                                           _p(v, C.c_double), _p(i, C.c_double), _p(b, C.c_double))
    return rc, v, i, b[:len(br)]


def mesh_read(path, fmt, cap_nodes=100000, cap_el=400000):
    """The reference's GmshLoader (fmt 1) / VTKLoader (2) / STLLoader (3):
    (xyz, element type, element nodes (4, -1 padded), material id)."""
    L = lib()
    L.ref_mesh_read.argtypes = [C.c_char_p, C.c_int, C.c_int64, C.c_int64, _I64, _I64, _D, _I32, _I32, _I32]
    nn, ne = C.c_int64(), C.c_int64()
    xyz = np.zeros((cap_nodes, 3))
    et = np.zeros(cap_el, np.int32)
    en = np.zeros((cap_el, 4), np.int32)
    tg = np.zeros(cap_el, np.int32)
    rc = L.ref_mesh_read(str(path).encode(), int(fmt), cap_nodes, cap_el, C.byref(nn), C.byref(ne),
                         _p(xyz, C.c_double), _p(et, C.c_int32), _p(en, C.c_int32), _p(tg, C.c_int32))
    if rc != 0:
        raise RuntimeError("reference mesh loader failed (%d)" % rc)
    return xyz[:nn.value], et[:ne.value], en[:ne.value], tg[:ne.value]
