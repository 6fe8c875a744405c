"""MI355X-native PIC-MCC cycle of SCDAT: Python binding over the C ABI.

The compute path is the gfx950 library ``libscdat_amd.so`` built in-tree
(``python -m scdat_amd.build``).  This module is a thin ctypes layer over
``include/scdat_amd.h``; it never computes physics itself.  Importing it
without the built library raises immediately: there is no CPU fallback.

The class :class:`PICCycle` mirrors the reference orchestrator
``SCDAT::PICcore::PICCycle`` (Tools/PICcore/include/PICCycle.h:33-185):
``setMesh`` / ``setBoundaryCondition`` / ``initialize`` / ``executeTimeStep`` /
``getChargeDensity`` / ``getPotential`` / ``getElectricField`` /
``getStatistics`` / ``getSurfaceCurrentLedger``.
"""
import atexit
import ctypes as C
import os
import sys
import weakref

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libscdat_amd.so")

# physical constants, Tools/Basic/include/Constants.h:46-53
ELEMENTARY_CHARGE = 1.602176634e-19
ELECTRON_MASS = 9.1093837015e-31
PROTON_MASS = 1.67262192369e-27
ATOMIC_MASS_UNIT = 1.66053906660e-27
VACUUM_PERMITTIVITY = 8.8541878128e-12
BOLTZMANN = 1.380649e-23

OK, ERR_ARG, ERR_STATE, ERR_DEVICE, ERR_NOMEM, ERR_SOLVER = 0, -1, -2, -3, -4, -5
BC_ABSORB, BC_PERIODIC, BC_REFLECT, BC_CONDUCTOR, BC_ESCAPE, BC_HOST = 0, 1, 2, 3, 4, 5
KERNEL_PUSH, KERNEL_PCG = 0, 1
NODE_FREE, NODE_DIRICHLET, NODE_CONDUCTOR = 0, 1, 2
MAX_GROUPS = 16


class ScdatError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"scdat error {code}: {msg}")
        self.code = code


class SolverFailure(ScdatError):
    """A Poisson solve missed its tolerance (SCDAT_ERR_SOLVER).  The call
    still completed: the failed steps kept the previous potential and field,
    as PICCycle::solvePoissonEquation does (PICCycle.cpp:731-767)."""


class BoxSpec(C.Structure):
    _fields_ = [("lo", C.c_double * 3), ("hi", C.c_double * 3), ("n", C.c_int32 * 3),
                ("conductor", C.c_int32), ("c_lo", C.c_double * 3), ("c_hi", C.c_double * 3)]


class Config(C.Structure):
    _fields_ = [("precision", C.c_int32), ("dt", C.c_double), ("adaptive_substep", C.c_int32),
                ("floor_ratio", C.c_double), ("max_substeps", C.c_int32),
                ("poisson_operator", C.c_int32), ("permittivity", C.c_double),
                ("load_sign", C.c_int32), ("pcg_tol", C.c_double), ("pcg_max_iter", C.c_int32),
                ("inner_tol", C.c_double), ("sort_interval", C.c_int32),
                ("bfield", C.c_double * 3)]


class Stats(C.Structure):
    _fields_ = [("completed_steps", C.c_int64), ("simulation_time", C.c_double),
                ("total_charge", C.c_double), ("total_energy", C.c_double),
                ("active_particles", C.c_int64), ("surface_deposited_charge", C.c_double),
                ("conductor_charge", C.c_double), ("floating_potential", C.c_double),
                ("pcg_iterations", C.c_int32), ("pcg_outer", C.c_int32),
                ("pcg_rel_residual", C.c_double), ("pcg_converged", C.c_int32),
                ("lost_particles", C.c_int64), ("particle_pushes", C.c_int64),
                ("pcg_initial_rel_residual", C.c_double), ("pcg_iterations_total", C.c_int64),
                ("pcg_failed_solves", C.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class LedgerEntry(C.Structure):
    _fields_ = [("absorbed_electron_charge_c", C.c_double), ("absorbed_ion_charge_c", C.c_double),
                ("emitted_electron_charge_c", C.c_double), ("incident_energy_j", C.c_double),
                ("absorbed_electron_particles", C.c_int64), ("absorbed_ion_particles", C.c_int64),
                ("emitted_electron_particles", C.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Impact(C.Structure):
    _fields_ = [("species", C.c_int32), ("id", C.c_int32), ("group", C.c_int32), ("cell", C.c_int32),
                ("x", C.c_double * 3), ("v", C.c_double * 3)]


MCC_ELASTIC, MCC_INELASTIC, MCC_IONIZATION, MCC_EXCITATION = 0, 1, 2, 3
MCC_CHARGE_EXCHANGE, MCC_RECOMBINATION, MCC_COULOMB = 4, 5, 6
MAX_MCC_PROCESSES = 16
SEE_OFF, SEE_REFERENCE, SEE_STERNGLASS = 0, 1, 2


class MccProcess(C.Structure):
    _fields_ = [("type", C.c_int32), ("species", C.c_int32), ("product_electron_species", C.c_int32),
                ("product_ion_species", C.c_int32), ("target_density", C.c_double),
                ("threshold_ev", C.c_double), ("sigma_max_m2", C.c_double), ("energy_loss_ev", C.c_double),
                ("reference_energy_ev", C.c_double), ("power_index", C.c_double), ("alpha", C.c_double),
                ("beta", C.c_double), ("coulomb_log", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class SurfaceEmission(C.Structure):
    _fields_ = [("see_model", C.c_int32), ("electron_species", C.c_int32), ("see_delta_max", C.c_double),
                ("see_energy_max_ev", C.c_double), ("see_temperature_ev", C.c_double),
                ("photo_current_density", C.c_double), ("photo_temperature_ev", C.c_double),
                ("photo_reference_v", C.c_double)]


EVENTS_FREE, EVENTS_RECORD, EVENTS_REPLAY = 0, 1, 2
SEGMENT_ALIGN = 512
PRECOND_JACOBI, PRECOND_TWO_LEVEL = 0, 1
ORDER_MORTON, ORDER_GIVEN, ORDER_TILED = 0, 1, 2


class SolverInfo(C.Structure):
    _fields_ = [("rows", C.c_int64), ("nonzeros", C.c_int64), ("ell_width", C.c_int32),
                ("preconditioner", C.c_int32), ("coarse_unknowns", C.c_int32), ("row_patterns", C.c_int32),
                ("pcg_block", C.c_int32), ("coarse_in_kernel", C.c_int32),
                ("tile_staged", C.c_int32), ("box_patterns", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}
SHAPE_PUSH_CHUNKS = 0
SHAPE_ROW_PATTERNS = 1
SHAPE_PCG_BLOCK = 2
SHAPE_TILE_STAGING = 3
SHAPE_CELL_ORDER = 4
SHAPE_STEP_GRAPH = 5
SHAPE_LATTICE_PUSH = 6
SHAPE_DEPOSIT_BUCKETS = 7


class ScenarioStats(C.Structure):
    _fields_ = [("completed_steps", C.c_int64), ("simulation_time", C.c_double), ("time_step", C.c_double),
                ("weight_scale", C.c_double), ("floating_potential", C.c_double),
                ("conductor_charge", C.c_double), ("pcg_iterations", C.c_int32), ("pcg_converged", C.c_int32),
                ("pcg_rel_residual", C.c_double), ("first_node", C.c_int64), ("n_nodes", C.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


GRAD_DELTA_MAX, GRAD_PHOTO_J, GRAD_CONDUCTANCE, GRAD_FACE_YIELD, N_GRAD = 0, 1, 2, 3, 9


class FieldEmission(C.Structure):
    _fields_ = [("electron_species", C.c_int32), ("emission_current_density_a_per_m2", C.c_double),
                ("emission_area_m2", C.c_double), ("emitted_electron_speed_m_per_s", C.c_double),
                ("emitted_electron_weight", C.c_double), ("max_emitted_particles_per_event", C.c_int32),
                ("enable_secondary_electron_emission", C.c_int32), ("secondary_electron_yield_per_ion", C.c_double),
                ("secondary_electron_speed_m_per_s", C.c_double), ("secondary_electron_weight", C.c_double),
                ("max_secondary_particles_per_event", C.c_int32)]


def field_emission(**kw):
    p = FieldEmission()
    lib().scdat_field_emission_default(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


class ChargingParams(C.Structure):
    _fields_ = [("see_delta_max", C.c_double), ("see_energy_max_ev", C.c_double), ("face_yield", C.c_double * 6),
                ("photo_current_density", C.c_double), ("photo_temperature_ev", C.c_double),
                ("photo_reference_v", C.c_double), ("sunlit_faces", C.c_int32),
                ("surface_conductance_s", C.c_double)]

    def vector(self):
        """(delta_max, J0, G, face_yield[6]) in gradient order."""
        return np.array([self.see_delta_max, self.photo_current_density, self.surface_conductance_s]
                        + list(self.face_yield))

    def set_vector(self, theta):
        self.see_delta_max, self.photo_current_density, self.surface_conductance_s = (float(t) for t in theta[:3])
        for k in range(6):
            self.face_yield[k] = float(theta[3 + k])
        return self


def charging_params(**kw):
    p = ChargingParams()
    lib().scdat_charging_params_default(C.byref(p))
    for k, v in kw.items():
        if k == "face_yield":
            for i in range(6):
                p.face_yield[i] = v[i]
        else:
            setattr(p, k, v)
    return p


class EventCounts(C.Structure):
    _fields_ = [("collisions", C.c_int64 * MAX_MCC_PROCESSES), ("created", C.c_int64), ("removed", C.c_int64),
                ("see_emitted", C.c_int64), ("photo_emitted", C.c_int64), ("overflow", C.c_int64),
                ("impacts_dropped", C.c_int64), ("mcc_candidates", C.c_int64),
                ("mcc_majorant_exceeded", C.c_int64)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_ if k != "collisions"}
        d["collisions"] = list(self.collisions)
        return d


_P = C.c_void_p
_D = C.POINTER(C.c_double)
_F = C.POINTER(C.c_float)
_I32 = C.POINTER(C.c_int32)
_I64 = C.POINTER(C.c_int64)

# exported symbol -> (restype, argtypes); the test-suite checks that this
# table, the header and the library agree.
SIGNATURES = {
    "scdat_config_default": (None, [C.POINTER(Config)]),
    "scdat_box_mesh_size": (C.c_int, [C.POINTER(BoxSpec), _I64, _I64]),
    "scdat_box_mesh_build": (C.c_int, [C.POINTER(BoxSpec), _D, _I32, _I32, _I32]),
    "scdat_box_locate": (C.c_int, [C.POINTER(BoxSpec), C.c_int64, _D, _I32]),
    "scdat_philox4x32": (None, [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "scdat_load_maxwellian": (C.c_int, [C.POINTER(BoxSpec), C.c_int64, C.c_uint64, C.c_uint32,
                                        C.c_double, _D, _D, _D, _I32]),
    "scdat_create": (C.c_int, [C.POINTER(Config), C.POINTER(_P)]),
    "scdat_destroy": (None, [_P]),
    "scdat_last_error": (C.c_char_p, [_P]),
    "scdat_device_info": (C.c_int, [C.c_char_p, C.c_int32, _I32]),
    "scdat_set_mesh": (C.c_int, [_P, C.c_int64, _D, C.c_int64, _I32, _I32, _D, _I32]),
    "scdat_set_boundary": (C.c_int, [_P, C.c_int32, C.c_int32, C.c_int32, _D]),
    "scdat_get_escaped": (C.c_int, [_P, C.c_int32, _I64]),
    "scdat_get_host_impacts": (C.c_int, [_P, C.c_int64, C.POINTER(Impact), _I64]),
    "scdat_add_species": (C.c_int, [_P, C.c_double, C.c_double, C.c_double, C.c_int64, _I32]),
    "scdat_set_particles": (C.c_int, [_P, C.c_int32, C.c_int64, _D, _D, _I32]),
    "scdat_get_particle_count": (C.c_int, [_P, C.c_int32, _I64, _I64]),
    "scdat_get_particles": (C.c_int, [_P, C.c_int32, _D, _D, _I32, _I32]),
    "scdat_set_conductor_charge": (C.c_int, [_P, C.c_double]),
    "scdat_initialize": (C.c_int, [_P]),
    "scdat_step": (C.c_int, [_P, C.c_int32]),
    "scdat_push": (C.c_int, [_P]),
    "scdat_deposit": (C.c_int, [_P]),
    "scdat_solve": (C.c_int, [_P]),
    "scdat_compute_efield": (C.c_int, [_P]),
    "scdat_sort": (C.c_int, [_P]),
    "scdat_set_time_step": (C.c_int, [_P, C.c_double]),
    "scdat_set_pcg_controls": (C.c_int, [_P, C.c_double, C.c_int32]),
    "scdat_set_efield": (C.c_int, [_P, _D]),
    "scdat_set_potential": (C.c_int, [_P, _D]),
    "scdat_set_rho": (C.c_int, [_P, _D]),
    "scdat_get_fields": (C.c_int, [_P, _D, _D, _D]),
    "scdat_get_rho_fixed": (C.c_int, [_P, _I64, _D]),
    "scdat_get_tet_maps": (C.c_int, [_P, _F]),
    "scdat_get_species_charge_fixed": (C.c_int, [_P, C.c_int32, _I64]),
    "scdat_get_stats": (C.c_int, [_P, C.POINTER(Stats)]),
    "scdat_get_stats_noalloc": (C.c_int, [_P, C.POINTER(Stats)]),
    "scdat_get_ledger": (C.c_int, [_P, C.POINTER(LedgerEntry), C.c_int32]),
    "scdat_get_stage_times": (C.c_int, [_P, _D]),
    "scdat_get_push_kernel_time": (C.c_int, [_P, _D, _I64]),
    "scdat_set_kernel_timing": (C.c_int, [_P, C.c_int32]),
    "scdat_get_kernel_time": (C.c_int, [_P, C.c_int32, _D, _I64]),
    "scdat_sync": (C.c_int, [_P]),
    "scdat_mcc_process_default": (C.c_int, [C.c_int32, C.POINTER(MccProcess)]),
    "scdat_mcc_configure": (C.c_int, [_P, C.POINTER(MccProcess), C.c_int32]),
    "scdat_mcc_set_density": (C.c_int, [_P, C.c_int32, C.c_double]),
    "scdat_set_seed": (C.c_int, [_P, C.c_uint64]),
    "scdat_set_surface_emission": (C.c_int, [_P, C.c_int32, C.POINTER(SurfaceEmission)]),
    "scdat_get_event_counts": (C.c_int, [_P, C.POINTER(EventCounts)]),
    "scdat_sampler_config_default": (None, [C.c_void_p]),
    "scdat_sample_surface_current": (C.c_int, [C.c_void_p, C.c_double, C.c_void_p]),
    "scdat_sampler_plate": (None, [C.c_void_p, C.POINTER(BoxSpec)]),
    "scdat_sampler_counts": (None, [C.c_void_p, _I64, _I64, _D]),
    "scdat_sampler_direct": (None, [C.c_int64, _D]),
    "scdat_circuit_size": (C.c_int, [_P, _I32, _I32]),
    "scdat_set_conductor_circuit": (C.c_int, [_P, C.c_int32, _I32, _P, _I32, _I32]),
    "scdat_set_scenarios": (C.c_int, [_P, C.c_int32, _I64, _D, _D, C.POINTER(C.c_uint64)]),
    "scdat_charging_params_default": (None, [C.POINTER(ChargingParams)]),
    "scdat_set_launch_shape": (C.c_int, [_P, C.c_int32, C.c_int32]),
    "scdat_get_launch_shape": (C.c_int, [_P, C.c_int32, _I32]),
    "scdat_autotune": (C.c_int, [_P]),
    "scdat_set_preconditioner": (C.c_int, [_P, C.c_int32]),
    "scdat_set_node_order": (C.c_int, [_P, C.c_int32]),
    "scdat_io_last_error": (C.c_char_p, []),
    "scdat_field_emission_default": (None, [C.POINTER(FieldEmission)]),
    "scdat_set_field_emission": (C.c_int, [_P, C.c_int32, C.POINTER(FieldEmission)]),
    "scdat_set_emission_current_density": (C.c_int, [_P, C.c_int32, C.c_double]),
    "scdat_append_particles": (C.c_int, [_P, C.c_int32, C.c_int64, _D, _D, _I32]),
    "scdat_mesh_read": (C.c_int, [C.c_char_p, C.c_int32, C.POINTER(C.c_void_p)]),
    "scdat_mesh_free": (None, [_P]),
    "scdat_mesh_sizes": (C.c_int, [_P, _I64, _I64, _I64]),
    "scdat_mesh_arrays": (C.c_int, [_P, _D, _I32, _I32, _I32, _I32]),
    "scdat_mesh_face_groups": (C.c_int, [_P, C.c_int32, _I32, _I32, C.c_int32, _I32, _I64]),
    "scdat_write_vtk": (C.c_int, [C.c_char_p, C.c_int64, _D, C.c_int64, _I32, C.c_int32,
                                  C.POINTER(C.c_char_p), _I32, C.POINTER(_D)]),
    "scdat_export_columns": (C.c_int, [C.c_char_p, C.c_char_p, C.c_char_p, C.c_int64, _D, C.c_int32,
                                       C.POINTER(C.c_char_p), C.POINTER(_D)]),
    "scdat_layer_count": (C.c_int, [_P, C.c_int32, _I64]),
    "scdat_layer_profile": (C.c_int, [_P, C.c_int32, C.c_double, C.c_int32, C.c_double, _D, _D, _D, _D, _D, _D]),
    "scdat_get_solver_info": (C.c_int, [_P, C.POINTER(SolverInfo)]),
    "scdat_charging_forward": (C.c_int, [_P, C.POINTER(ChargingParams), C.c_int32, _D]),
    "scdat_charging_forward_events": (C.c_int, [_P, C.POINTER(ChargingParams), C.c_int32, C.c_int32, _D]),
    "scdat_charging_adjoint": (C.c_int, [_P, C.POINTER(ChargingParams), C.c_int32, C.c_int32, _D, _D]),
    "scdat_set_scenario_emission": (C.c_int, [_P, C.c_int32, C.c_int32, C.POINTER(SurfaceEmission)]),
    "scdat_set_scenario_mcc_density": (C.c_int, [_P, C.c_int32, C.c_int32, C.c_double]),
    "scdat_get_scenario_stats": (C.c_int, [_P, C.c_int32, C.POINTER(ScenarioStats)]),
    "scdat_get_scenario_ledger": (C.c_int, [_P, C.c_int32, C.POINTER(LedgerEntry), C.c_int32]),
    "scdat_set_patch_didv": (C.c_int, [_P, _D]),
    "scdat_get_patch_potentials": (C.c_int, [_P, _D]),
    "scdat_get_patch_basis": (C.c_int, [_P, C.c_int32, _D]),
    "scdat_circuit_create": (C.c_int, [C.POINTER(_P)]),
    "scdat_circuit_destroy": (None, [_P]),
    "scdat_circuit_last_error": (C.c_char_p, [_P]),
    "scdat_circuit_add_node": (C.c_int, [_P, C.c_double, C.c_double, C.c_double, C.c_int32, C.c_double, _I32]),
    "scdat_circuit_add_branch": (C.c_int, [_P, C.c_int32, C.c_int32, C.c_double, C.c_double, _I32]),
    "scdat_circuit_set_node": (C.c_int, [_P, C.c_int32, C.c_double, C.c_double, C.c_double, C.c_int32, C.c_double]),
    "scdat_circuit_set_branch_conductance": (C.c_int, [_P, C.c_int32, C.c_double]),
    "scdat_circuit_get_potentials": (C.c_int, [_P, _D]),
    "scdat_circuit_advance": (C.c_int, [_P, C.c_double, _D, _D, _D, _D, C.c_void_p, C.c_int32, C.c_double,
                                        _D, _D, _D, _D]),
}


def mcc_process(type, species, **kw):
    """A collision process with the reference defaults of its type, overridden by kw."""
    p = MccProcess()
    if lib().scdat_mcc_process_default(type, C.byref(p)) != OK:
        raise ValueError(f"unknown collision type {type}")
    p.species = species
    for k, v in kw.items():
        setattr(p, k, v)
    return p

_lib = None

# Live handles, released at interpreter exit in creation-reverse order while
# the HIP runtime is still up: Python's atexit hooks run before the C
# runtime's exit handlers (the HIP runtime's and a profiler's among them),
# whereas a finalizer left to interpreter shutdown may run after them.
_live = weakref.WeakValueDictionary()
_live_seq = [0]


def _track(obj):
    _live_seq[0] += 1
    _live[_live_seq[0]] = obj


@atexit.register
def _close_all():
    for k in sorted(list(_live.keys()), reverse=True):
        o = _live.get(k)
        if o is not None:
            try:
                o.close()
            except Exception:
                pass


def lib():
    """Load the gfx950 library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m scdat_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _ptr(a, t):
    return a.ctypes.data_as(C.POINTER(t)) if a is not None else None


def device_info():
    name = C.create_string_buffer(256)
    ncu = C.c_int32(0)
    rc = lib().scdat_device_info(name, 256, C.byref(ncu))
    if rc != OK:
        raise ScdatError(rc, "no HIP device")
    return name.value.decode(), ncu.value


# ---------------------------------------------------------------------------
# mesh generation (host only)
# ---------------------------------------------------------------------------
def box_spec(n, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), conductor=None):
    s = BoxSpec()
    nn = (n, n, n) if np.isscalar(n) else tuple(n)
    for r in range(3):
        s.lo[r], s.hi[r], s.n[r] = lo[r], hi[r], nn[r]
    if conductor is not None:
        s.conductor = 1
        for r in range(3):
            s.c_lo[r], s.c_hi[r] = conductor[0][r], conductor[1][r]
    return s


def box_mesh(n, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), conductor=None):
    """Structured 6-tet box mesh (optionally with a conductor box removed).

    Returns dict(xyz[nn,3], tets[nt,4], kind[nn], group[nt,4]).
    """
    s = box_spec(n, lo, hi, conductor)
    nn, nt = C.c_int64(), C.c_int64()
    rc = lib().scdat_box_mesh_size(C.byref(s), C.byref(nn), C.byref(nt))
    if rc != OK:
        raise ScdatError(rc, "invalid box spec")
    xyz = np.zeros((nn.value, 3))
    tets = np.zeros((nt.value, 4), np.int32)
    kind = np.zeros(nn.value, np.int32)
    grp = np.zeros((nt.value, 4), np.int32)
    rc = lib().scdat_box_mesh_build(C.byref(s), _ptr(xyz, C.c_double), _ptr(tets, C.c_int32),
                                    _ptr(kind, C.c_int32), _ptr(grp, C.c_int32))
    if rc != OK:
        raise ScdatError(rc, "box mesh build failed")
    return {"xyz": xyz, "tets": tets, "kind": kind, "group": grp}


def default_config(**kw):
    c = Config()
    lib().scdat_config_default(C.byref(c))
    for k, v in kw.items():
        if k == "bfield":
            for r in range(3):
                c.bfield[r] = v[r]
        else:
            setattr(c, k, v)
    return c


# ---------------------------------------------------------------------------
# context
# ---------------------------------------------------------------------------
class Context:
    """One device simulation (owns all device memory; one HIP stream)."""

    def __init__(self, config=None, **kw):
        self.cfg = config if config is not None else default_config(**kw)
        h = _P()
        rc = lib().scdat_create(C.byref(self.cfg), C.byref(h))
        if rc != OK:
            raise ScdatError(rc, "scdat_create failed (no device or bad config)")
        self.h = h
        self.n_nodes = 0
        self.n_tets = 0
        self.species = []
        _track(self)

    def close(self):
        if getattr(self, "h", None):
            lib().scdat_destroy(self.h)
            self.h = None

    def __del__(self):
        # never during interpreter shutdown: the atexit hook has closed every
        # live context by then, and the runtime may already be gone
        if sys is None or sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _chk(self, rc):
        if rc != OK:
            msg = lib().scdat_last_error(self.h)
            cls = SolverFailure if rc == ERR_SOLVER else ScdatError
            raise cls(rc, msg.decode() if msg else "")
        return rc

    # -- setup --------------------------------------------------------------
    def set_mesh(self, xyz, tets, kind, dirichlet=None, group=None):
        xyz = np.ascontiguousarray(xyz, np.float64)
        tets = np.ascontiguousarray(tets, np.int32)
        kind = np.ascontiguousarray(kind, np.int32)
        d = None if dirichlet is None else np.ascontiguousarray(dirichlet, np.float64)
        g = None if group is None else np.ascontiguousarray(group, np.int32)
        self._chk(lib().scdat_set_mesh(self.h, len(kind), _ptr(xyz, C.c_double), len(tets),
                                       _ptr(tets, C.c_int32), _ptr(kind, C.c_int32),
                                       _ptr(d, C.c_double), _ptr(g, C.c_int32)))
        self.n_nodes, self.n_tets = len(kind), len(tets)

    def set_boundary(self, group, bc_type, partner=-1, shift=None):
        sh = None if shift is None else np.ascontiguousarray(shift, np.float64)
        self._chk(lib().scdat_set_boundary(self.h, group, bc_type, partner, _ptr(sh, C.c_double)))

    def escaped(self, species):
        n = C.c_int64()
        self._chk(lib().scdat_get_escaped(self.h, species, C.byref(n)))
        return n.value

    def host_impacts(self):
        """Impacts on BC_HOST groups in the last step: a list of dicts with
        species, id, group, cell, x (crossing point) and v (velocity)."""
        n = C.c_int64()
        self._chk(lib().scdat_get_host_impacts(self.h, 0, None, C.byref(n)))
        if n.value == 0:
            return []
        arr = (Impact * n.value)()
        self._chk(lib().scdat_get_host_impacts(self.h, n.value, arr, C.byref(n)))
        return [dict(species=e.species, id=e.id, group=e.group, cell=e.cell, x=np.array(e.x[:]),
                     v=np.array(e.v[:])) for e in arr[:n.value]]

    def add_species(self, q, m, w, capacity):
        sid = C.c_int32()
        self._chk(lib().scdat_add_species(self.h, q, m, w, int(capacity), C.byref(sid)))
        self.species.append((q, m, w))
        return sid.value

    def set_particles(self, species, x, v, cell=None):
        x = np.ascontiguousarray(x, np.float64).reshape(-1, 3)
        v = np.ascontiguousarray(v, np.float64).reshape(-1, 3)
        c = None if cell is None else np.ascontiguousarray(cell, np.int32)
        self._chk(lib().scdat_set_particles(self.h, species, len(x), _ptr(x, C.c_double),
                                            _ptr(v, C.c_double), _ptr(c, C.c_int32)))

    def particle_count(self, species):
        a, s = C.c_int64(), C.c_int64()
        self._chk(lib().scdat_get_particle_count(self.h, species, C.byref(a), C.byref(s)))
        return a.value, s.value

    def get_particles(self, species):
        """Returns (x[n,3], v[n,3], cell[n], id[n]) in device order."""
        _, n = self.particle_count(species)
        x = np.zeros((n, 3))
        v = np.zeros((n, 3))
        c = np.zeros(n, np.int32)
        i = np.zeros(n, np.int32)
        self._chk(lib().scdat_get_particles(self.h, species, _ptr(x, C.c_double), _ptr(v, C.c_double),
                                            _ptr(c, C.c_int32), _ptr(i, C.c_int32)))
        return x, v, c, i

    def get_particles_by_id(self, species, n_ids):
        """Particles scattered to upload order; removed ones get cell -1, NaN state."""
        x, v, c, i = self.get_particles(species)
        X = np.full((n_ids, 3), np.nan)
        V = np.full((n_ids, 3), np.nan)
        Cc = np.full(n_ids, -1, np.int32)
        m = (i >= 0) & (i < n_ids)
        X[i[m]] = x[m]
        V[i[m]] = v[m]
        Cc[i[m]] = c[m]
        return X, V, Cc

    def set_conductor_charge(self, q):
        self._chk(lib().scdat_set_conductor_charge(self.h, q))

    # -- cycle --------------------------------------------------------------
    def initialize(self):
        self._chk(lib().scdat_initialize(self.h))

    def step(self, n=1):
        self._chk(lib().scdat_step(self.h, n))

    def push(self):
        self._chk(lib().scdat_push(self.h))

    def deposit(self):
        self._chk(lib().scdat_deposit(self.h))

    def solve(self):
        self._chk(lib().scdat_solve(self.h))

    def compute_efield(self):
        self._chk(lib().scdat_compute_efield(self.h))

    def sort(self):
        self._chk(lib().scdat_sort(self.h))

    def set_efield(self, E):
        E = np.ascontiguousarray(E, np.float64).reshape(-1, 3)
        self._chk(lib().scdat_set_efield(self.h, _ptr(E, C.c_double)))

    def set_potential(self, phi):
        phi = np.ascontiguousarray(phi, np.float64)
        self._chk(lib().scdat_set_potential(self.h, _ptr(phi, C.c_double)))

    def set_rho(self, rho):
        rho = np.ascontiguousarray(rho, np.float64)
        self._chk(lib().scdat_set_rho(self.h, _ptr(rho, C.c_double)))

    def fields(self):
        n = self.n_nodes
        rho, phi, E = np.zeros(n), np.zeros(n), np.zeros((n, 3))
        self._chk(lib().scdat_get_fields(self.h, _ptr(rho, C.c_double), _ptr(phi, C.c_double),
                                         _ptr(E, C.c_double)))
        return rho, phi, E

    def rho_fixed(self):
        r = np.zeros(self.n_nodes, np.int64)
        u = C.c_double()
        self._chk(lib().scdat_get_rho_fixed(self.h, _ptr(r, C.c_int64), C.byref(u)))
        return r, u.value

    def tet_maps(self):
        m = np.zeros((self.n_tets, 12), np.float32)
        self._chk(lib().scdat_get_tet_maps(self.h, _ptr(m, C.c_float)))
        return m

    def species_charge_fixed(self, s):
        v = C.c_int64()
        self._chk(lib().scdat_get_species_charge_fixed(self.h, s, C.byref(v)))
        return v.value

    def set_seed(self, seed):
        self._chk(lib().scdat_set_seed(self.h, seed))

    def mcc_configure(self, procs):
        arr = (MccProcess * max(1, len(procs)))(*procs)
        self._chk(lib().scdat_mcc_configure(self.h, arr, len(procs)))

    def mcc_set_density(self, process, density):
        self._chk(lib().scdat_mcc_set_density(self.h, process, density))

    def set_surface_emission(self, group, see_model=SEE_OFF, electron_species=0, delta_max=0.0,
                             energy_max=400.0, see_temperature=2.0, photo_j=0.0, photo_temperature=2.0,
                             photo_reference_v=0.0):
        e = SurfaceEmission(see_model, electron_species, delta_max, energy_max, see_temperature, photo_j,
                            photo_temperature, photo_reference_v)
        self._chk(lib().scdat_set_surface_emission(self.h, group, C.byref(e)))

    def set_scenarios(self, node_start, weight_scale=None, time_step=None, seed=None):
        """Split the mesh into independent scenarios (batched charging cases)."""
        ns = np.ascontiguousarray(node_start, np.int64)
        ws = None if weight_scale is None else np.ascontiguousarray(weight_scale, np.float64)
        dt = None if time_step is None else np.ascontiguousarray(time_step, np.float64)
        sd = None if seed is None else np.ascontiguousarray(seed, np.uint64)
        self._chk(lib().scdat_set_scenarios(self.h, len(ns) - 1, _ptr(ns, C.c_int64), _ptr(ws, C.c_double),
                                            _ptr(dt, C.c_double), _ptr(sd, C.c_uint64)))
        self.n_scenarios = len(ns) - 1

    def set_scenario_emission(self, scenario, group, see_model=SEE_OFF, electron_species=0, delta_max=0.0,
                              energy_max=400.0, see_temperature=2.0, photo_j=0.0, photo_temperature=2.0,
                              photo_reference_v=0.0):
        e = SurfaceEmission(see_model, electron_species, delta_max, energy_max, see_temperature, photo_j,
                            photo_temperature, photo_reference_v)
        self._chk(lib().scdat_set_scenario_emission(self.h, scenario, group, C.byref(e)))

    def set_scenario_mcc_density(self, scenario, process, density):
        """Target density of collision process `process` in one scenario."""
        self._chk(lib().scdat_set_scenario_mcc_density(self.h, scenario, process, density))

    def scenario_stats(self, scenario):
        st = ScenarioStats()
        self._chk(lib().scdat_get_scenario_stats(self.h, scenario, C.byref(st)))
        return st.as_dict()

    def scenario_ledger(self, scenario, n_groups=12):
        arr = (LedgerEntry * n_groups)()
        self._chk(lib().scdat_get_scenario_ledger(self.h, scenario, arr, n_groups))
        return [e.as_dict() for e in arr]

    def set_field_emission(self, group, params):
        self._chk(lib().scdat_set_field_emission(self.h, group, C.byref(params)))

    def set_emission_current_density(self, group, j):
        self._chk(lib().scdat_set_emission_current_density(self.h, group, float(j)))

    def append_particles(self, species, x, v, cell=None):
        x = np.ascontiguousarray(x, np.float64).reshape(-1, 3)
        v = np.ascontiguousarray(v, np.float64).reshape(-1, 3)
        cl = None if cell is None else np.ascontiguousarray(cell, np.int32)
        self._chk(lib().scdat_append_particles(self.h, species, len(x), _ptr(x, C.c_double), _ptr(v, C.c_double),
                                               _ptr(cl, C.c_int32)))

    def layer_profile(self, axis=2, length=1.0, n_cells=1, area=1.0):
        """collectLayerProfile of the PIC harness: per node layer along `axis`
        the mean potential and field component, electron and ion densities and
        the electron temperature (eV)."""
        n = C.c_int64()
        self._chk(lib().scdat_layer_count(self.h, axis, C.byref(n)))
        out = {k: np.zeros(n.value) for k in ("position", "potential_v", "electric_field_v_per_m",
                                             "electron_density_m3", "ion_density_m3", "electron_temperature_ev")}
        self._chk(lib().scdat_layer_profile(self.h, axis, float(length), int(n_cells), float(area),
                                            *[_ptr(out[k], C.c_double) for k in out]))
        return out

    def set_node_order(self, order):
        self._chk(lib().scdat_set_node_order(self.h, order))

    def set_preconditioner(self, kind):
        self._chk(lib().scdat_set_preconditioner(self.h, kind))

    def solver_info(self):
        i = SolverInfo()
        self._chk(lib().scdat_get_solver_info(self.h, C.byref(i)))
        return i.as_dict()

    def autotune(self):
        """Time the push kernel's launch shapes on a scratch copy and keep the
        fastest; returns the chosen particles per thread."""
        self._chk(lib().scdat_autotune(self.h))
        return self.push_chunks

    @property
    def push_chunks(self):
        v = C.c_int32()
        self._chk(lib().scdat_get_launch_shape(self.h, SHAPE_PUSH_CHUNKS, C.byref(v)))
        return v.value

    @push_chunks.setter
    def push_chunks(self, n):
        self._chk(lib().scdat_set_launch_shape(self.h, SHAPE_PUSH_CHUNKS, int(n)))

    @property
    def row_patterns(self):
        """0: the FP32 inner iteration reads the ELL slots; 1 (default): row
        patterns where they pay (few distinct rows); 2: row patterns with
        4-byte ids."""
        v = C.c_int32()
        self._chk(lib().scdat_get_launch_shape(self.h, SHAPE_ROW_PATTERNS, C.byref(v)))
        return v.value

    @row_patterns.setter
    def row_patterns(self, mode):
        self._chk(lib().scdat_set_launch_shape(self.h, SHAPE_ROW_PATTERNS, int(mode)))

    @property
    def pcg_block(self):
        """Rows of a PCG row block: 0 (by size), 256 or 512."""
        v = C.c_int32()
        self._chk(lib().scdat_get_launch_shape(self.h, SHAPE_PCG_BLOCK, C.byref(v)))
        return v.value

    @pcg_block.setter
    def pcg_block(self, rows):
        self._chk(lib().scdat_set_launch_shape(self.h, SHAPE_PCG_BLOCK, int(rows)))

    @property
    def tile_staging(self):
        """1 (default): in the tiled order the mixed-precision iteration stages
        each tile's neighbour values in LDS where the mesh allows; 0: row
        patterns.  Same bits either way."""
        v = C.c_int32()
        self._chk(lib().scdat_get_launch_shape(self.h, SHAPE_TILE_STAGING, C.byref(v)))
        return v.value

    @tile_staging.setter
    def tile_staging(self, mode):
        self._chk(lib().scdat_set_launch_shape(self.h, SHAPE_TILE_STAGING, int(mode)))

    @property
    def cell_order(self):
        """Order of the cells along the sorted particle arrays: 0 (default)
        the tet index, 1 the centroids' Morton order.  Same fields, particles
        and ledger counts either way."""
        v = C.c_int32()
        self._chk(lib().scdat_get_launch_shape(self.h, SHAPE_CELL_ORDER, C.byref(v)))
        return v.value

    @cell_order.setter
    def cell_order(self, mode):
        self._chk(lib().scdat_set_launch_shape(self.h, SHAPE_CELL_ORDER, int(mode)))

    @property
    def step_graph(self):
        """1: step() replays the launches around the Poisson iteration from
        HIP graphs recorded in the call (same results as eager launches)."""
        v = C.c_int32()
        self._chk(lib().scdat_get_launch_shape(self.h, SHAPE_STEP_GRAPH, C.byref(v)))
        return v.value

    @step_graph.setter
    def step_graph(self, mode):
        self._chk(lib().scdat_set_launch_shape(self.h, SHAPE_STEP_GRAPH, int(mode)))

    @property
    def lattice_push(self):
        """1 when the FP32 push locates sub-steps clear of boundary faces on
        the grid of a Kuhn-lattice mesh instead of walking tet faces (same
        results; off by default, autotune() picks the faster)."""
        v = C.c_int32()
        self._chk(lib().scdat_get_launch_shape(self.h, SHAPE_LATTICE_PUSH, C.byref(v)))
        return v.value

    @lattice_push.setter
    def lattice_push(self, mode):
        self._chk(lib().scdat_set_launch_shape(self.h, SHAPE_LATTICE_PUSH, int(mode)))

    @property
    def deposit_buckets(self):
        """1 (default): push blocks hand their reduced charge to per-row-range
        buckets summed after the pushes; 0: one global atomic per node."""
        v = C.c_int32()
        self._chk(lib().scdat_get_launch_shape(self.h, SHAPE_DEPOSIT_BUCKETS, C.byref(v)))
        return v.value

    @deposit_buckets.setter
    def deposit_buckets(self, mode):
        self._chk(lib().scdat_set_launch_shape(self.h, SHAPE_DEPOSIT_BUCKETS, int(mode)))

    def charging_forward(self, params, steps, events=EVENTS_FREE):
        """Floating potential after `steps` steps of the differentiable charging
        model, from the initialized state (the context is left unchanged).
        events: EVENTS_RECORD keeps each particle's absorption step and group,
        EVENTS_REPLAY imposes the recorded ones (the event-frozen map)."""
        v = C.c_double()
        self._chk(lib().scdat_charging_forward_events(self.h, C.byref(params), steps, int(events), C.byref(v)))
        return v.value

    def charging_adjoint(self, params, steps, checkpoint_interval=20):
        """(V_final, dV/d(delta_max, J0, G, face_yield[0..5])) by the discrete adjoint."""
        v = C.c_double()
        g = np.zeros(N_GRAD)
        self._chk(lib().scdat_charging_adjoint(self.h, C.byref(params), steps, checkpoint_interval, C.byref(v),
                                               _ptr(g, C.c_double)))
        return v.value, g

    def event_counts(self):
        e = EventCounts()
        self._chk(lib().scdat_get_event_counts(self.h, C.byref(e)))
        return e.as_dict()

    def set_conductor_circuit(self, circuit, patch_circuit_node, n_patches=7, node_patch=None, group_patch=None):
        """Hold conductor patches at circuit potentials (default: six faces + body)."""
        pc = np.ascontiguousarray(patch_circuit_node, np.int32)
        npa = None if node_patch is None else np.ascontiguousarray(node_patch, np.int32)
        gp = None if group_patch is None else np.ascontiguousarray(group_patch, np.int32)
        self._circuit = circuit  # keep alive
        self._n_patch = n_patches
        self._chk(lib().scdat_set_conductor_circuit(self.h, n_patches, _ptr(npa, C.c_int32), circuit.h,
                                                     _ptr(pc, C.c_int32), _ptr(gp, C.c_int32)))

    def set_patch_didv(self, didv):
        d = np.ascontiguousarray(didv, np.float64)
        self._chk(lib().scdat_set_patch_didv(self.h, _ptr(d, C.c_double)))

    def patch_potentials(self):
        v = np.zeros(max(1, getattr(self, "_n_patch", 0)))
        self._chk(lib().scdat_get_patch_potentials(self.h, _ptr(v, C.c_double)))
        return v

    def patch_basis(self, k):
        phi = np.zeros(self.n_nodes)
        self._chk(lib().scdat_get_patch_basis(self.h, k, _ptr(phi, C.c_double)))
        return phi

    def set_time_step(self, dt):
        self._chk(lib().scdat_set_time_step(self.h, dt))

    def set_pcg_controls(self, tol, max_iter=0):
        """Relative tolerance and inner-iteration cap of later solves."""
        self._chk(lib().scdat_set_pcg_controls(self.h, float(tol), int(max_iter)))

    def stats(self):
        st = Stats()
        self._chk(lib().scdat_get_stats(self.h, C.byref(st)))
        return st.as_dict()

    def stats_noalloc(self):
        """stats() through scdat_get_stats_noalloc (no allocation in the library)."""
        st = Stats()
        self._chk(lib().scdat_get_stats_noalloc(self.h, C.byref(st)))
        return st.as_dict()

    def ledger(self, n_groups=12):
        arr = (LedgerEntry * n_groups)()
        self._chk(lib().scdat_get_ledger(self.h, arr, n_groups))
        return [e.as_dict() for e in arr]

    def stage_times(self):
        a = np.zeros(5)
        self._chk(lib().scdat_get_stage_times(self.h, _ptr(a, C.c_double)))
        return dict(zip(["push", "rhs", "pcg", "field", "sort"], a.tolist()))

    def push_kernel_time(self):
        t, n = C.c_double(), C.c_int64()
        self._chk(lib().scdat_get_push_kernel_time(self.h, C.byref(t), C.byref(n)))
        return t.value, n.value

    def set_kernel_timing(self, on=True, period=1):
        """Time every push launch and every `period`-th PCG iteration launch."""
        self._chk(lib().scdat_set_kernel_timing(self.h, int(period) if on else 0))

    def kernel_time(self, which):
        """(average ms, launches) for KERNEL_PUSH or KERNEL_PCG (one PCG iteration)."""
        t, n = C.c_double(), C.c_int64()
        self._chk(lib().scdat_get_kernel_time(self.h, which, C.byref(t), C.byref(n)))
        return t.value, n.value

    def sync(self):
        self._chk(lib().scdat_sync(self.h))


from .loading import maxwellian_box, philox4x32  # noqa: E402,F401
from .piccycle import PICCycle  # noqa: E402,F401


class CircuitEntry(C.Structure):
    _fields_ = [("row", C.c_int32), ("column", C.c_int32), ("coefficient", C.c_double)]


class Circuit:
    """SurfaceCircuitCoupling (Tools/Coupling/include/SurfaceCircuitCoupling.h:324-361)
    over the C ABI: host FP64 backward-Euler circuit of surface patches."""

    def __init__(self):
        h = C.c_void_p()
        if lib().scdat_circuit_create(C.byref(h)) != OK:
            raise ScdatError(ERR_ARG, "circuit_create failed")
        self.h = h
        self.n_nodes = 0
        self.n_branches = 0
        _track(self)

    def close(self):
        if self.h:
            lib().scdat_circuit_destroy(self.h)
            self.h = None

    def __del__(self):
        if sys is None or sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        if rc != OK:
            raise ScdatError(rc, lib().scdat_circuit_last_error(self.h).decode())

    def addNode(self, potential_v=0.0, capacitance_f=0.0, shunt_conductance_s=0.0, fixed=False,
                fixed_value_v=0.0):
        i = C.c_int32()
        self._chk(lib().scdat_circuit_add_node(self.h, potential_v, capacitance_f, shunt_conductance_s,
                                               int(bool(fixed)), fixed_value_v, C.byref(i)))
        self.n_nodes += 1
        return i.value

    def addBranch(self, from_node, to_node, conductance_s, bias_v=0.0):
        i = C.c_int32()
        self._chk(lib().scdat_circuit_add_branch(self.h, from_node, to_node, conductance_s, bias_v, C.byref(i)))
        self.n_branches += 1
        return i.value

    def setNode(self, node, potential_v=0.0, capacitance_f=0.0, shunt_conductance_s=0.0, fixed=False,
                fixed_value_v=0.0):
        self._chk(lib().scdat_circuit_set_node(self.h, node, potential_v, capacitance_f, shunt_conductance_s,
                                               int(bool(fixed)), fixed_value_v))

    def setBranchConductance(self, branch, conductance_s):
        self._chk(lib().scdat_circuit_set_branch_conductance(self.h, branch, conductance_s))

    def size(self):
        n, b = C.c_int32(), C.c_int32()
        self._chk(lib().scdat_circuit_size(self.h, C.byref(n), C.byref(b)))
        return n.value, b.value

    def potentials(self):
        v = np.zeros(self.n_nodes)
        self._chk(lib().scdat_circuit_get_potentials(self.h, _ptr(v, C.c_double)))
        return v

    def advanceImplicit(self, dt, plasma_current, plasma_didv, additional_rhs=None, additional_diagonal=None,
                        entries=(), max_delta_potential_v=float("inf")):
        """Returns (node potentials, plasma currents, branch currents, max |dV|)."""
        n = self.n_nodes

        def arr(a):
            return None if a is None else np.ascontiguousarray(a, np.float64)
        cur, dv, rh, dg = arr(plasma_current), arr(plasma_didv), arr(additional_rhs), arr(additional_diagonal)
        ent = (CircuitEntry * max(1, len(entries)))(*[CircuitEntry(int(r), int(c), float(x)) for r, c, x in entries])
        v, i, b = np.zeros(n), np.zeros(n), np.zeros(max(1, self.n_branches))
        md = C.c_double()
        p = lambda a: None if a is None else _ptr(a, C.c_double)  # noqa: E731
        self._chk(lib().scdat_circuit_advance(self.h, dt, p(cur), p(dv), p(rh), p(dg), C.cast(ent, C.c_void_p),
                                              len(entries), max_delta_potential_v, p(v), p(i), p(b), C.byref(md)))
        return v, i, b[:self.n_branches], md.value


class SamplerConfig(C.Structure):
    _fields_ = [("surface_area_m2", C.c_double), ("gap_distance_m", C.c_double),
                ("surface_aspect_ratio", C.c_double), ("surface_potential_v", C.c_double),
                ("plasma_reference_potential_v", C.c_double), ("electron_density_m3", C.c_double),
                ("electron_temperature_ev", C.c_double), ("ion_density_m3", C.c_double),
                ("ion_temperature_ev", C.c_double), ("ion_mass_amu", C.c_double),
                ("neutral_density_m3", C.c_double), ("enable_mcc", C.c_int32), ("nx", C.c_int32),
                ("ny", C.c_int32), ("z_layers", C.c_int32), ("particles_per_element", C.c_int32),
                ("window_steps", C.c_int32), ("precision", C.c_int32), ("seed", C.c_uint64),
                ("time_step", C.c_double)]


class CurrentSample(C.Structure):
    _fields_ = [("valid", C.c_int32), ("sampled_surface_potential_v", C.c_double),
                ("electron_collection_current_density_a_per_m2", C.c_double),
                ("ion_collection_current_density_a_per_m2", C.c_double),
                ("emitted_electron_current_density_a_per_m2", C.c_double),
                ("net_collection_current_density_a_per_m2", C.c_double),
                ("current_derivative_a_per_m2_per_v", C.c_double), ("total_collisions", C.c_int64),
                ("time_step", C.c_double), ("macro_electrons", C.c_int64), ("macro_ions", C.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def sampler_config(**kw):
    """PicMccSurfaceSamplerConfig with the reference defaults, overridden by kw."""
    c = SamplerConfig()
    lib().scdat_sampler_config_default(C.byref(c))
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def sample_surface_current(cfg, probe_delta_v=0.0):
    """PicMccSurfaceCurrentSampler::sample (probe_delta_v = 0) or
    sampleWithDerivative (three concurrent windows)."""
    out = CurrentSample()
    rc = lib().scdat_sample_surface_current(C.byref(cfg), probe_delta_v, C.byref(out))
    if rc != OK:
        raise ScdatError(rc, "sample_surface_current failed")
    return out.as_dict()


def sampler_inputs(cfg):
    """The window's mesh spec, macro-particle counts/weight and initial
    particles (for checkers)."""
    spec = BoxSpec()
    lib().scdat_sampler_plate(C.byref(cfg), C.byref(spec))
    ne, ni, w = C.c_int64(), C.c_int64(), C.c_double()
    lib().scdat_sampler_counts(C.byref(cfg), C.byref(ne), C.byref(ni), C.byref(w))
    return spec, ne.value, ni.value, w.value


def sampler_direct(v):
    v = np.ascontiguousarray(v, np.float64)
    lib().scdat_sampler_direct(len(v), _ptr(v, C.c_double))
    return v
