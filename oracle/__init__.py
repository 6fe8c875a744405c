"""ctypes binding of the FP64 CPU oracle (oracle/oracle.cpp).

TEST INFRASTRUCTURE ONLY: the checker for the GPU path and the timed CPU
baseline.  Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
leg import this package; the product (scdat_amd) never does.
"""
import ctypes as C
import math
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
REF_LIB_PATH = os.path.join(_HERE, "_ref", "libscdat_ref.so")

_D = C.POINTER(C.c_double)
_F = C.POINTER(C.c_float)
_I32 = C.POINTER(C.c_int32)
_I64 = C.POINTER(C.c_int64)
_P = C.c_void_p

_SIG = {
    "orc_create": (_P, [C.c_int, _D, C.c_int, _I32, _I32, _D, _I32]),
    "orc_destroy": (None, [_P]),
    "orc_num_threads": (C.c_int, []),
    "orc_set_threads": (None, [C.c_int]),
    "orc_set_poisson": (None, [_P, C.c_int, C.c_double, C.c_int, C.c_double, C.c_int]),
    "orc_set_push": (None, [_P, C.c_double, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int]),
    "orc_set_group": (None, [_P, C.c_int, C.c_int, _D]),
    "orc_set_partner": (None, [_P, _I32]),
    "orc_get_neighbors": (None, [_P, _I32]),
    "orc_get_geometry": (None, [_P, _D, _D, _D]),
    "orc_kfull_nnz": (C.c_int64, [_P]),
    "orc_get_kfull": (None, [_P, _I64, _I32, _D]),
    "orc_add_species": (C.c_int, [_P, C.c_double, C.c_double, C.c_double]),
    "orc_set_particles": (None, [_P, C.c_int, C.c_int64, _D, _D, _I32]),
    "orc_get_particles": (C.c_int64, [_P, C.c_int, _D, _D, _I32]),
    "orc_set_efield": (None, [_P, _D]),
    "orc_set_bfield": (None, [_P, _D]),
    "orc_set_rho": (None, [_P, _D]),
    "orc_set_phi": (None, [_P, _D]),
    "orc_set_conductor_charge": (None, [_P, C.c_double]),
    "orc_conductor_charge": (C.c_double, [_P]),
    "orc_floating_potential": (C.c_double, [_P]),
    "orc_get_fields": (None, [_P, _D, _D, _D]),
    "orc_solve": (C.c_int, [_P, _D, _I32]),
    "orc_efield": (None, [_P]),
    "orc_push": (None, [_P]),
    "orc_deposit": (None, [_P]),
    "orc_step": (C.c_int, [_P, C.c_int]),
    "orc_get_ledger": (None, [_P, _D, C.c_int]),
    "orc_lost": (C.c_int64, [_P]),
    "orc_kinetic_energy": (C.c_double, [_P]),
    "orc_boris": (None, [_D, _D, _D, C.c_double, C.c_double]),
    "orc_adaptive_dt": (C.c_double, [_D, _D, _D, C.c_double, C.c_double, C.c_double]),
    "orc_barycentric": (None, [_P, C.c_int, _D, _D]),
    "orc_deposit_fixed_f32": (None, [C.c_int, _F, _I32, C.c_int64, _F, _I32, C.c_int64, _I64]),
    "orc_set_seed": (None, [_P, C.c_uint32, C.c_uint32]),
    "orc_set_step": (None, [_P, C.c_int64]),
    "orc_mcc_configure": (None, [_P, C.c_int, _D]),
    "orc_set_emission": (None, [_P, C.c_int, _D]),
    "orc_post_push": (None, [_P]),
    "orc_charging_forward": (C.c_double, [_P, _D, C.c_int]),
    "orc_mcc_counts": (None, [_P, _I64]),
    "orc_get_ids": (None, [_P, C.c_int, _I32]),
    "orc_sigma": (None, [_D, C.c_int64, _D, _D]),
    "orc_coulomb_log": (C.c_double, [C.c_double, C.c_double]),
    "orc_see_yield": (None, [C.c_int, C.c_double, C.c_double, C.c_int64, _D, _D, _D]),
    "orc_photo_current": (None, [C.c_double, C.c_double, C.c_double, C.c_int64, _D, _D]),
    "orc_philox": (None, [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
}

# Collision process types (CollisionProcessFactory, CollisionAlgorithm.cpp:738-812)
MCC_ELASTIC, MCC_INELASTIC, MCC_IONIZATION, MCC_EXCITATION = 0, 1, 2, 3
MCC_CHARGE_EXCHANGE, MCC_RECOMBINATION, MCC_COULOMB = 4, 5, 6
_PROC_KEYS = ("type", "species", "e_species", "i_species", "density", "threshold", "sigma_max",
              "energy_loss", "reference_energy", "power_index", "alpha", "beta", "coulomb_log")
_PROC_DEFAULT = dict(type=0, species=0, e_species=-1, i_species=-1, density=0.0, threshold=0.0,
                     sigma_max=0.0, energy_loss=0.0, reference_energy=1.0, power_index=0.5,
                     alpha=1.2, beta=0.35, coulomb_log=10.0)


def pack_process(p):
    d = dict(_PROC_DEFAULT)
    d.update(p)
    a = np.zeros(16)
    for k, key in enumerate(_PROC_KEYS):
        a[k] = d[key]
    return a


def sigma(proc, e_ev):
    e = np.ascontiguousarray(e_ev, np.float64)
    out = np.zeros_like(e)
    prm = pack_process(proc)
    lib().orc_sigma(_p(prm, C.c_double), len(e), _p(e, C.c_double), _p(out, C.c_double))
    return out


def coulomb_log(temperature_k, density_m3):
    return lib().orc_coulomb_log(temperature_k, density_m3)


def see_yield(model, dmax, emax, e_ev, cos_inc):
    e = np.ascontiguousarray(e_ev, np.float64)
    c = np.ascontiguousarray(np.broadcast_to(cos_inc, e.shape), np.float64)
    out = np.zeros_like(e)
    lib().orc_see_yield(model, dmax, emax, len(e), _p(e, C.c_double), _p(c, C.c_double), _p(out, C.c_double))
    return out


def photo_current(j0y, t_ph, v_ref, V):
    v = np.ascontiguousarray(V, np.float64)
    out = np.zeros_like(v)
    lib().orc_photo_current(j0y, t_ph, v_ref, len(v), _p(v, C.c_double), _p(out, C.c_double))
    return out


def field_emission(params, hits, dt, pending=(0.0, 0.0), next_id=0):
    """Ion-driven field and secondary emission of one cathode, restating
    FieldEmissionBoundaryCondition::processParticle
    (Tools/Boundary/src/FieldEmissionBoundaryCondition.cpp:88-178) in
    impact order.

    params: dict with the fields of FieldEmissionBoundaryCondition::Parameters
    (emission_current_density_a_per_m2, emission_area_m2,
    emitted_electron_speed_m_per_s, max_emitted_particles_per_event,
    enable_secondary_electron_emission, secondary_electron_yield_per_ion,
    secondary_electron_speed_m_per_s, secondary_electron_weight,
    max_secondary_particles_per_event).  hits: (point, normal, ion_weight)
    per ion impact.  Returns (electrons, pending) with electrons a list of
    (id, position, velocity, kind) where kind is "field" or "secondary".
    """
    e = 1.602176634e-19
    pf, ps = pending
    out = []
    for point, normal, w_ion in hits:
        n = np.asarray(normal, np.float64)
        mag = math.sqrt(float(n @ n))
        u = n / mag if mag > 1e-12 else np.array([0.0, 0.0, 1.0])
        made = 0

        def emit(count, speed, kind):
            nonlocal made, next_id
            for _ in range(count):
                off = 1.0e-10 * float(made + 1)
                out.append((next_id, np.asarray(point, np.float64) + u * off, u * speed, kind))
                next_id += 1
                made += 1

        j = params.get("emission_current_density_a_per_m2", 0.0)
        if dt > 0.0 and j > 0.0:
            expect = j * params.get("emission_area_m2", 1e-8) * dt / e
            if math.isfinite(expect) and expect > 0.0:
                pf += expect
                cnt = min(int(math.floor(pf)), params.get("max_emitted_particles_per_event", 64))
                pf = max(0.0, pf - cnt)
                emit(cnt, params.get("emitted_electron_speed_m_per_s", 2e6), "field")
        y = params.get("secondary_electron_yield_per_ion", 0.0)
        if params.get("enable_secondary_electron_emission", False) and y > 0.0:
            expect = y * max(0.0, w_ion) / max(1e-12, params.get("secondary_electron_weight", 1.0))
            if math.isfinite(expect) and expect > 0.0:
                ps += expect
                cnt = min(int(math.floor(ps)), params.get("max_secondary_particles_per_event", 64))
                ps = max(0.0, ps - cnt)
                emit(cnt, params.get("secondary_electron_speed_m_per_s", 8e5), "secondary")
    return out, (pf, ps)


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().orc_philox(c, k, o)
    return list(o)

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        for k, (r, a) in _SIG.items():
            f = getattr(L, k)
            f.restype = r
            f.argtypes = a
        _lib = L
    return _lib


def _p(a, t):
    return None if a is None else a.ctypes.data_as(C.POINTER(t))


def boris(v, E, B, qm, dt):
    v = np.array(v, np.float64)
    E = np.ascontiguousarray(E, np.float64)
    B = np.ascontiguousarray(B, np.float64)
    lib().orc_boris(_p(v, C.c_double), _p(E, C.c_double), _p(B, C.c_double), qm, dt)
    return v


def adaptive_dt(v0, E, B, qm, dt_max, floor_ratio=1e-3):
    a = [np.ascontiguousarray(x, np.float64) for x in (v0, E, B)]
    return lib().orc_adaptive_dt(*[_p(x, C.c_double) for x in a], qm, dt_max, floor_ratio)


def deposit_fixed_f32(maps, tets, x32, cell, cs, n_nodes, rho=None):
    """Integer deposition checker (see orc_deposit_fixed_f32)."""
    maps = np.ascontiguousarray(maps, np.float32)
    tets = np.ascontiguousarray(tets, np.int32)
    x32 = np.ascontiguousarray(x32, np.float32).reshape(-1, 3)
    cell = np.ascontiguousarray(cell, np.int32)
    out = np.zeros(n_nodes, np.int64) if rho is None else rho
    lib().orc_deposit_fixed_f32(len(maps), _p(maps, C.c_float), _p(tets, C.c_int32), len(x32),
                                _p(x32, C.c_float), _p(cell, C.c_int32), int(cs), _p(out, C.c_int64))
    return out


class Oracle:
    """FP64 restatement of the PIC-MCC cycle on one mesh."""

    def __init__(self, xyz, tets, kind, dirichlet=None, group=None, threads=None):
        self.xyz = np.ascontiguousarray(xyz, np.float64)
        self.tets = np.ascontiguousarray(tets, np.int32)
        self.kind = np.ascontiguousarray(kind, np.int32)
        self.nn, self.nt = len(self.kind), len(self.tets)
        d = None if dirichlet is None else np.ascontiguousarray(dirichlet, np.float64)
        g = None if group is None else np.ascontiguousarray(group, np.int32)
        if threads:
            lib().orc_set_threads(int(threads))
        self.h = lib().orc_create(self.nn, _p(self.xyz, C.c_double), self.nt,
                                  _p(self.tets, C.c_int32), _p(self.kind, C.c_int32),
                                  _p(d, C.c_double), _p(g, C.c_int32))
        self.n_species = 0

    def close(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_poisson(self, op=0, eps=8.8541878128e-12, load_sign=1, tol=1e-10, max_it=20000):
        lib().orc_set_poisson(self.h, op, eps, load_sign, tol, max_it)

    def set_push(self, dt, adaptive=1, floor_ratio=1e-3, max_sub=64, deposit_mode=0, nseg=10):
        lib().orc_set_push(self.h, dt, adaptive, floor_ratio, max_sub, deposit_mode, nseg)

    def set_group(self, g, bc_type, shift=None):
        sh = None if shift is None else np.ascontiguousarray(shift, np.float64)
        lib().orc_set_group(self.h, g, bc_type, _p(sh, C.c_double))

    def set_partner(self, partner):
        p = np.ascontiguousarray(partner, np.int32)
        lib().orc_set_partner(self.h, _p(p, C.c_int32))

    def neighbors(self):
        nb = np.zeros((self.nt, 4), np.int32)
        lib().orc_get_neighbors(self.h, _p(nb, C.c_int32))
        return nb

    def geometry(self):
        A = np.zeros((self.nt, 9))
        X0 = np.zeros((self.nt, 3))
        vol = np.zeros(self.nt)
        lib().orc_get_geometry(self.h, _p(A, C.c_double), _p(X0, C.c_double), _p(vol, C.c_double))
        return A, X0, vol

    def kfull(self):
        import scipy.sparse as sp
        nnz = lib().orc_kfull_nnz(self.h)
        rp = np.zeros(self.nn + 1, np.int64)
        ci = np.zeros(nnz, np.int32)
        va = np.zeros(nnz)
        lib().orc_get_kfull(self.h, _p(rp, C.c_int64), _p(ci, C.c_int32), _p(va, C.c_double))
        return sp.csr_matrix((va, ci, rp), shape=(self.nn, self.nn))

    def add_species(self, q, m, w):
        self.n_species += 1
        return lib().orc_add_species(self.h, q, m, w)

    def set_particles(self, s, x, v, cell):
        x = np.ascontiguousarray(x, np.float64).reshape(-1, 3)
        v = np.ascontiguousarray(v, np.float64).reshape(-1, 3)
        c = np.ascontiguousarray(cell, np.int32)
        lib().orc_set_particles(self.h, s, len(x), _p(x, C.c_double), _p(v, C.c_double), _p(c, C.c_int32))

    def get_particles(self, s):
        n = lib().orc_get_particles(self.h, s, None, None, None)
        x, v, c = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros(n, np.int32)
        lib().orc_get_particles(self.h, s, _p(x, C.c_double), _p(v, C.c_double), _p(c, C.c_int32))
        return x, v, c

    def set_efield(self, E):
        E = np.ascontiguousarray(E, np.float64).reshape(-1, 3)
        lib().orc_set_efield(self.h, _p(E, C.c_double))

    def set_bfield(self, B):
        B = np.ascontiguousarray(B, np.float64).reshape(-1, 3)
        lib().orc_set_bfield(self.h, _p(B, C.c_double))

    def set_rho(self, rho):
        rho = np.ascontiguousarray(rho, np.float64)
        lib().orc_set_rho(self.h, _p(rho, C.c_double))

    def set_phi(self, phi):
        phi = np.ascontiguousarray(phi, np.float64)
        lib().orc_set_phi(self.h, _p(phi, C.c_double))

    def set_conductor_charge(self, q):
        lib().orc_set_conductor_charge(self.h, q)

    def conductor_charge(self):
        return lib().orc_conductor_charge(self.h)

    def floating_potential(self):
        return lib().orc_floating_potential(self.h)

    def fields(self):
        rho, phi, E = np.zeros(self.nn), np.zeros(self.nn), np.zeros((self.nn, 3))
        lib().orc_get_fields(self.h, _p(rho, C.c_double), _p(phi, C.c_double), _p(E, C.c_double))
        return rho, phi, E

    def solve(self):
        rr = C.c_double()
        cv = C.c_int32()
        it = lib().orc_solve(self.h, C.byref(rr), C.byref(cv))
        return it, rr.value, bool(cv.value)

    def efield(self):
        lib().orc_efield(self.h)

    def push(self):
        lib().orc_push(self.h)

    def deposit(self):
        lib().orc_deposit(self.h)

    def step(self, n=1):
        return lib().orc_step(self.h, n)

    def ledger(self, n_groups=12):
        out = np.zeros((n_groups, 7))
        lib().orc_get_ledger(self.h, _p(out, C.c_double), n_groups)
        return out

    def lost(self):
        return lib().orc_lost(self.h)

    def kinetic_energy(self):
        return lib().orc_kinetic_energy(self.h)

    def set_seed(self, seed):
        lib().orc_set_seed(self.h, seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF)

    def set_step(self, step):
        lib().orc_set_step(self.h, step)

    def mcc_configure(self, procs):
        a = np.concatenate([pack_process(p) for p in procs]) if procs else np.zeros(16)
        lib().orc_mcc_configure(self.h, len(procs), _p(a, C.c_double))

    def set_emission(self, group, see_model=0, delta_max=0.0, energy_max=400.0, see_temperature=2.0,
                     photo_j=0.0, photo_temperature=2.0, photo_reference_v=0.0, electron_species=0):
        a = np.array([see_model, delta_max, energy_max, see_temperature, photo_j, photo_temperature,
                      photo_reference_v, electron_species], np.float64)
        lib().orc_set_emission(self.h, group, _p(a, C.c_double))

    def charging_forward(self, params, steps):
        """Floating potential after `steps` steps of the differentiable
        charging model; params has the fields of scdat_charging_params."""
        q = np.array([params.see_delta_max, params.see_energy_max_ev] + list(params.face_yield)
                     + [params.photo_current_density, params.photo_temperature_ev, params.photo_reference_v,
                        float(params.sunlit_faces), params.surface_conductance_s], dtype=np.float64)
        return lib().orc_charging_forward(self.h, _p(q, C.c_double), int(steps))

    def post_push(self):
        lib().orc_post_push(self.h)

    def mcc_counts(self):
        out = np.zeros(22, np.int64)
        lib().orc_mcc_counts(self.h, _p(out, C.c_int64))
        return dict(collisions=out[:16].copy(), created=int(out[16]), removed=int(out[17]),
                    see_emitted=int(out[18]), photo_emitted=int(out[19]), candidates=int(out[20]),
                    majorant_exceeded=int(out[21]))

    def get_ids(self, s):
        n = lib().orc_get_particles(self.h, s, None, None, None)
        ids = np.zeros(n, np.int32)
        lib().orc_get_ids(self.h, s, _p(ids, C.c_int32))
        return ids

    def barycentric(self, t, p):
        p = np.ascontiguousarray(p, np.float64)
        lam = np.zeros(4)
        lib().orc_barycentric(self.h, t, _p(p, C.c_double), _p(lam, C.c_double))
        return lam
