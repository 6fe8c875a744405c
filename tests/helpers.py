"""Shared test set-ups: box meshes with a floating conductor, Maxwellian
plasma, and matched device/oracle contexts on the same seeded inputs."""
import numpy as np

import oracle
import scdat_amd as S

COND = ((0.4, 0.4, 0.4), (0.6, 0.6, 0.6))
EPS0 = S.VACUUM_PERMITTIVITY
QE = S.ELEMENTARY_CHARGE


def plasma_setup(n=8, npart=2000, seed=1, Te_eV=1.0, density=1e10, conductor=True):
    """Mesh + electron/ion Maxwellian states + dt = 0.2/omega_pe."""
    cond = COND if conductor else None
    m = S.box_mesh(n, conductor=cond)
    spec = S.box_spec(n, conductor=cond)
    vth_e = np.sqrt(Te_eV * QE / S.ELECTRON_MASS)
    vth_i = np.sqrt(Te_eV * QE / S.PROTON_MASS)
    xe, ve, ce = S.maxwellian_box(spec, npart, seed, 0, vth_e)
    xi, vi, ci = S.maxwellian_box(spec, npart, seed, 1, vth_i)
    vol = 1.0 - (0.2 ** 3 if conductor else 0.0)
    w = density * vol / npart
    wpe = np.sqrt(density * QE ** 2 / (EPS0 * S.ELECTRON_MASS))
    dt = 0.2 / wpe
    return dict(mesh=m, spec=spec, e=(xe, ve, ce), i=(xi, vi, ci), w=w, dt=dt)


def make_oracle(st, tol=1e-12, deposit_mode=0, adaptive=1, max_sub=64):
    m = st["mesh"]
    o = oracle.Oracle(m["xyz"], m["tets"], m["kind"], group=m["group"])
    for g in range(6, 12):
        o.set_group(g, S.BC_CONDUCTOR)
    o.set_poisson(0, EPS0, 1, tol, 50000)
    o.set_push(st["dt"], adaptive, 1e-3, max_sub, deposit_mode)
    se = o.add_species(-QE, S.ELECTRON_MASS, st["w"])
    si = o.add_species(QE, S.PROTON_MASS, st["w"])
    o.set_particles(se, *st["e"])
    o.set_particles(si, *st["i"])
    return o


def make_ctx(st, precision=0, tol=1e-8, sort_interval=20, adaptive=1, order=None):
    m = st["mesh"]
    c = S.Context(precision=precision, dt=st["dt"], pcg_tol=tol, sort_interval=sort_interval,
                  adaptive_substep=adaptive)
    if order is not None:
        c.set_node_order(order)
    c.set_mesh(m["xyz"], m["tets"], m["kind"], group=m["group"])
    for g in range(6, 12):
        c.set_boundary(g, S.BC_CONDUCTOR)
    n = len(st["e"][0])
    c.add_species(-QE, S.ELECTRON_MASS, st["w"], 2 * n)
    c.add_species(QE, S.PROTON_MASS, st["w"], 2 * n)
    c.set_particles(0, *st["e"])
    c.set_particles(1, *st["i"])
    return c


def smooth_field(xyz, amp=50.0):
    """A smooth non-linear nodal E field (V/m)."""
    x, y, z = xyz[:, 0], xyz[:, 1], xyz[:, 2]
    return amp * np.stack([np.sin(3 * y + 1) + 0.3, np.cos(2 * z) * x, np.sin(2 * x + z)], 1)


def rel_err(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def periodic_partners(m, pairs=((0, 1, 0), (2, 3, 1))):
    """Partner tet across periodic wall pairs (group_a, group_b, axis), by
    matching face centroids in the other two coordinates."""
    xyz, tets, grp = m["xyz"], m["tets"], m["group"]
    partner = -np.ones_like(tets)
    faces = {}
    for t, k in zip(*np.nonzero(grp >= 0)):
        faces.setdefault(int(grp[t, k]), []).append((t, xyz[np.delete(tets[t], k)].mean(0)))
    for ga, gb, ax in pairs:
        keep = [r for r in range(3) if r != ax]
        B = {tuple(np.round(c[keep], 9)): t for t, c in faces.get(gb, [])}
        for t, c in faces.get(ga, []):
            tb = B[tuple(np.round(c[keep], 9))]
            ka = [k for k in range(4) if grp[t, k] == ga][0]
            kb = [k for k in range(4) if grp[tb, k] == gb][0]
            partner[t, ka] = tb
            partner[tb, kb] = t
    return partner


def geo_oracle(st, tol=1e-10):
    """The oracle of a GEO case built by scdat_amd.geo.inputs: same mesh,
    species, processes, surface emission and seed as geo.context."""
    import scdat_amd.geo as G
    cfg = st["cfg"]
    m = st["mesh"]
    o = oracle.Oracle(m["xyz"], m["tets"], m["kind"], group=m["group"])
    for g in G.CONDUCTOR_GROUPS:
        o.set_group(g, S.BC_CONDUCTOR)
    o.set_poisson(0, EPS0, 1, tol, 50000)
    o.set_push(st["dt"], 1, 1e-3, 64, 0)
    for q, mass in G.species():
        o.add_species(q, mass, st["w"])
    o.set_particles(G.ELECTRON, *st["e"])
    o.set_particles(G.PROTON, *st["i"])
    o.set_particles(G.OXYGEN_ION, np.zeros((0, 3)), np.zeros((0, 3)), np.zeros(0, np.int32))
    o.mcc_configure(st["procs"])
    o.set_seed(int(cfg["seed"]))
    for g in G.CONDUCTOR_GROUPS:
        o.set_emission(g, see_model=S.SEE_STERNGLASS, delta_max=cfg["delta_max"], energy_max=cfg["energy_max_ev"],
                       see_temperature=cfg["see_temperature_ev"],
                       photo_j=cfg["photo_current_density"] if g == G.SUNLIT_GROUP else 0.0,
                       photo_temperature=cfg["photo_temperature_ev"], photo_reference_v=0.0,
                       electron_species=G.ELECTRON)
    return o
