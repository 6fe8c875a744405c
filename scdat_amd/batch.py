"""Batched charging scenarios: many independent cases in one context.

A sweep of charging cases (mesh size, particle count, plasma temperature and
density, secondary-emission yield) is concatenated into one mesh and one
particle set; `Context.set_scenarios` tells the device which nodes, cells and
particles belong to which case.  Every step then advances all cases with one
launch per kernel, and each case gives the results it would give alone
(`standalone_context` builds that single-case context for comparison).

Each case is the charging geometry of the PIC benchmark: a 1 m box with the
0.2 m conductor cube [0.4, 0.6]^3 at its centre, floating, and a Maxwellian
electron-proton plasma loaded uniformly outside it
(Main/main_PICcore.cpp:151-229 builds the reference's equivalent case).
"""
import math

import numpy as np

from . import (BC_CONDUCTOR, ELECTRON_MASS, ELEMENTARY_CHARGE, ORDER_GIVEN, PRECOND_JACOBI, PROTON_MASS,
               SEE_STERNGLASS, SEGMENT_ALIGN, VACUUM_PERMITTIVITY, Context, box_mesh, box_spec)
from .loading import maxwellian_box, philox4x32

CONDUCTOR = ((0.4, 0.4, 0.4), (0.6, 0.6, 0.6))
CONDUCTOR_GROUPS = range(6, 12)
_SWEEP_TAG = 0x5357


def _u01(word):
    return (word + 0.5) / 4294967296.0


def sweep(n_scenarios, seed, edge=(8, 32), particles=(2.0e4, 2.0e6), te_ev=(1.0, 2.0e4),
          density=(1.0e6, 1.0e11), delta_max=(1.0, 3.0)):
    """Parameters of a sweep of charging cases, drawn from one Philox stream
    keyed by `seed`: mesh edge (hexes per axis), particles per species,
    temperature (Te = Ti) and density log-uniform, SEE delta_max uniform.
    Case k also gets its own loading seed."""
    key = [seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF]

    def logu(u, lo, hi):
        return math.exp(math.log(lo) + u * (math.log(hi) - math.log(lo)))

    out = []
    for k in range(n_scenarios):
        a = philox4x32([k, 0, 0, _SWEEP_TAG], key)
        b = philox4x32([k, 1, 0, _SWEEP_TAG], key)
        out.append(dict(
            edge=int(round(logu(_u01(a[0]), *edge))),
            n_particles=int(round(logu(_u01(a[1]), *particles))),
            te_ev=logu(_u01(a[2]), *te_ev),
            density=logu(_u01(a[3]), *density),
            delta_max=delta_max[0] + _u01(b[0]) * (delta_max[1] - delta_max[0]),
            seed=(b[1] << 32) | b[2],
        ))
    return out


def case_inputs(sc):
    """Mesh, particles, weight and time step of one case.

    dt = min(0.2 / omega_pe, 0.2 h / v_th,e): the plasma-period rule of the
    benchmark, capped so that a thermal electron crosses a fifth of a cell
    per step in the hot cases."""
    n = int(sc["edge"])
    m = box_mesh(n, conductor=CONDUCTOR)
    spec = box_spec(n, conductor=CONDUCTOR)
    qe = ELEMENTARY_CHARGE
    te = float(sc["te_ev"])
    vth_e = math.sqrt(te * qe / ELECTRON_MASS)
    vth_i = math.sqrt(sc.get("ti_ev", te) * qe / PROTON_MASS)
    npart = int(sc["n_particles"])
    xe, ve, ce = maxwellian_box(spec, npart, int(sc["seed"]), 0, vth_e)
    xi, vi, ci = maxwellian_box(spec, npart, int(sc["seed"]), 1, vth_i)
    vol = 1.0 - 0.2 ** 3
    w = float(sc["density"]) * vol / npart
    wpe = math.sqrt(float(sc["density"]) * qe ** 2 / (VACUUM_PERMITTIVITY * ELECTRON_MASS))
    dt = min(0.2 / wpe, 0.2 * (1.0 / n) / vth_e)
    return dict(mesh=m, e=(xe, ve, ce), i=(xi, vi, ci), w=w, dt=dt)


def all_case_inputs(scenarios, workers=16):
    """case_inputs of every scenario, built on a thread pool (the loader and
    the mesher run in the library, outside the interpreter lock)."""
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max_workers=max(1, int(workers))) as ex:
        return list(ex.map(case_inputs, scenarios))


SUNLIT_GROUP = 7  # the conductor's +x face
OXYGEN_ION = 2


def _emission(ctx, scenario, sc):
    """Sternglass secondary emission on the conductor and, when the case has
    `photo_j` (A/m^2), photoemission of 2 eV electrons from its +x face."""
    for g in CONDUCTOR_GROUPS:
        pj = float(sc.get("photo_j", 0.0)) if g == SUNLIT_GROUP else 0.0
        if scenario is None:
            ctx.set_surface_emission(g, SEE_STERNGLASS, 0, sc["delta_max"], 400.0, 2.0, photo_j=pj)
        else:
            ctx.set_scenario_emission(scenario, g, SEE_STERNGLASS, 0, sc["delta_max"], 400.0, 2.0, photo_j=pj)


def _has_mcc(scenarios):
    return any(float(sc.get("neutral_density", 0.0)) > 0.0 for sc in scenarios)


def _processes(density):
    from .geo import device_process, processes
    return [device_process(p) for p in processes(density)]


def standalone_context(sc, inputs=None, precision=0, tol=1e-8, sort_interval=20, see=True, capacity_factor=2,
                       autotune=False, like_batch=True, mcc=None):
    """The single-case context of case `sc`.  like_batch: the settings a
    batched case runs with (the caller's node order, Jacobi-PCG, 512-row
    PCG blocks), under which the batched case reproduces it bit for bit;
    otherwise the context's defaults.  mcc: the oxygen background of the
    GEO case at the case's `neutral_density` (with an O+ species); None: when
    that density is positive."""
    st = inputs or case_inputs(sc)
    m = st["mesh"]
    c = Context(precision=precision, dt=st["dt"], pcg_tol=tol, sort_interval=sort_interval)
    if like_batch:
        c.set_node_order(ORDER_GIVEN)
        c.set_preconditioner(PRECOND_JACOBI)
        c.pcg_block = 512  # the batched solve's row blocks
    c.set_mesh(m["xyz"], m["tets"], m["kind"], group=m["group"])
    for g in CONDUCTOR_GROUPS:
        c.set_boundary(g, BC_CONDUCTOR)
    c.set_seed(int(sc["seed"]))
    n = len(st["e"][0])
    c.add_species(-ELEMENTARY_CHARGE, ELECTRON_MASS, st["w"], capacity_factor * n)
    c.add_species(ELEMENTARY_CHARGE, PROTON_MASS, st["w"], capacity_factor * n)
    if mcc is None:
        mcc = _has_mcc([sc])
    if mcc:
        c.add_species(ELEMENTARY_CHARGE, 16.0 * PROTON_MASS, st["w"], n)
        c.mcc_configure(_processes(float(sc.get("neutral_density", 0.0))))
    if see:
        _emission(c, None, sc)
    c.set_particles(0, *st["e"])
    c.set_particles(1, *st["i"])
    c.initialize()
    if autotune:
        c.autotune()
    return c


def concat(cases):
    """Union of the cases' meshes and particles, each case's nodes starting
    at a multiple of SEGMENT_ALIGN (padding nodes belong to no tet)."""
    xyz, tets, kind, group = [], [], [], []
    node_start = [0]
    tet_off = []
    nn = nt = 0
    for st in cases:
        m = st["mesh"]
        n = len(m["xyz"])
        pad = (-n) % SEGMENT_ALIGN
        xyz += [m["xyz"], np.zeros((pad, 3))]
        kind += [m["kind"], np.ones(pad, np.int32)]
        tets.append(m["tets"] + nn)
        group.append(m["group"])
        tet_off.append(nt)
        nn += n + pad
        nt += len(m["tets"])
        node_start.append(nn)
    parts = []
    for sp in ("e", "i"):
        x = np.concatenate([st[sp][0] for st in cases])
        v = np.concatenate([st[sp][1] for st in cases])
        cell = np.concatenate([np.where(st[sp][2] >= 0, st[sp][2] + off, -1) for st, off in zip(cases, tet_off)])
        parts.append((x, v, cell.astype(np.int32)))
    return dict(xyz=np.concatenate(xyz), tets=np.concatenate(tets).astype(np.int32),
                kind=np.concatenate(kind).astype(np.int32), group=np.concatenate(group).astype(np.int32),
                node_start=np.array(node_start, np.int64), tet_offset=np.array(tet_off + [nt], np.int64),
                e=parts[0], i=parts[1])


def batch_context(scenarios, inputs=None, precision=0, tol=1e-8, sort_interval=20, see=True, capacity_factor=2,
                  autotune=False):
    """One context advancing every case of `scenarios` together; returns
    (context, union) where union holds the node and tet offsets per case."""
    cases = inputs or [case_inputs(sc) for sc in scenarios]
    u = concat(cases)
    c = Context(precision=precision, dt=cases[0]["dt"], pcg_tol=tol, sort_interval=sort_interval)
    c.set_mesh(u["xyz"], u["tets"], u["kind"], group=u["group"])
    for g in CONDUCTOR_GROUPS:
        c.set_boundary(g, BC_CONDUCTOR)
    c.set_scenarios(u["node_start"], weight_scale=[st["w"] for st in cases], time_step=[st["dt"] for st in cases],
                    seed=[int(sc["seed"]) for sc in scenarios])
    n = len(u["e"][0])
    c.add_species(-ELEMENTARY_CHARGE, ELECTRON_MASS, 1.0, capacity_factor * n)
    c.add_species(ELEMENTARY_CHARGE, PROTON_MASS, 1.0, capacity_factor * n)
    if _has_mcc(scenarios):
        # one process list; each case's background density is its own
        c.add_species(ELEMENTARY_CHARGE, 16.0 * PROTON_MASS, 1.0, n)
        c.mcc_configure(_processes(0.0))
        for k, sc in enumerate(scenarios):
            for p in range(3):
                c.set_scenario_mcc_density(k, p, float(sc.get("neutral_density", 0.0)))
    if see:
        for k, sc in enumerate(scenarios):
            _emission(c, k, sc)
    c.set_particles(0, *u["e"])
    c.set_particles(1, *u["i"])
    c.initialize()
    if autotune:
        c.autotune()
    return c, u


def split_particles(u, x, v, cell, ids):
    """Per-case (x, v, cell, id) of a particle download of the batch context,
    cells shifted back to the case's own numbering; absorbed particles
    (cell -1) are dropped."""
    out = []
    toff = u["tet_offset"]
    seg = np.searchsorted(toff, cell, side="right") - 1
    for k in range(len(toff) - 1):
        sel = (cell >= 0) & (seg == k)
        out.append((x[sel], v[sel], cell[sel] - toff[k], ids[sel]))
    return out
