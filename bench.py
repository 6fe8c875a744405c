#!/usr/bin/env python3
"""PIC-MCC cycle throughput on one MI355X.

Workload (the reference's scaled-up charging case): 1 m cube domain, 0.2 m
floating conducting cube at its centre, 64^3 hexes split into 6 tets each
(1,559,682 tets after the conductor cut-out, 272,897 nodes), phi = 0 on the
outer walls, 2^24 macro-particles (2^23 electrons + 2^23 H+), Maxwellian
Te = Ti = 1 eV, n = 1e10 m^-3, dt = 0.2 / omega_pe, collisionless.  FP32 push
/ gather / deposition, FP32 PCG with FP64 residual refinement to 1e-8.
Inputs are resident in HBM before the timed region.

One timed step = push + gather + walk + deposit of every particle, boundary
events, floating-conductor Poisson solve and nodal E (PICCycle order).  The
timed region is the configuration's whole run, 500 steps (--steps), over
which the outer walls absorb most electrons; value = particles pushed /
wall time of the timed steps.  The first 50 steps, with the plasma still
whole, are reported beside it (first_steps).

Prints one JSON line.  --gpus N with N > 1 exits: this build runs one GPU.

--workload sweep runs the batched charging sweep instead: --scenarios cases
(default 256) with mesh edge 8..32 hexes and 2e4..2e6 particles per species
(log-uniform), Te = Ti log-uniform in 1 eV..20 keV, n log-uniform in
1e6..1e11 m^-3, Sternglass SEE delta_max uniform in 1..3 on the conductor,
FP32 with FP64 refinement; all cases advance together in one context, and
the same cases are then run one after another in their own contexts, and
the batch continues with its steps replayed from HIP graphs.
value = particle pushes of all cases / wall time of the batched steps.

--workload adjoint runs the discrete adjoint of the floating potential on
the 32^3 box with 2^20 particles (Te = Ti = 1 eV, n = 1e10), 200 collisionless
steps, checkpoint every 20: the gradient with respect to SEE delta_max,
photoemission J0, surface conductance and the six face yields, checked
against central differences (h = 1e-4 relative) in FP64, with the FP32
adjoint's error against the FP64 one.  value = wall time of one gradient.

--workload geo runs the GEO charging case: 48^3 hexes, 2^23 particles
(electrons and protons, Te = Ti = 10 keV, n = 1e6 m^-3), an oxygen
background at 1e12 m^-3 (electron elastic and ionisation, proton elastic;
null-collision MCC), Sternglass secondary emission (delta_max 2 at 400 eV)
on the conductor and 20 uA/m^2 photoemission from its sunlit +x face, FP32
with FP64 refinement, dt = 0.2 h / v_th,e, the configuration's 2000 steps
timed after the warm-up (--geo-steps).  value = particle pushes / wall time.

--workload fp64 runs the reference-harness case in FP64: 16^3 hexes, 10^5
electrons + 10^5 protons (Te = Ti = 1 eV, n = 1e10), 100 collisionless steps,
Jacobi-PCG to 1e-10, beside the FP64 oracle on the same inputs (CPU
baseline and the potential's error after the steps).

--copy-calibration MB runs only a plain device copy of MB megabytes (the
reference kernel for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE counters).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s
TRAFFIC_FILE = os.path.join(HERE, "profiles", "pmc_traffic_64cube.json")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_threads():
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    return len(os.sched_getaffinity(0))


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def build_case(n_hex, n_part, seed):
    import scdat_amd as S
    cond = ((0.4, 0.4, 0.4), (0.6, 0.6, 0.6))
    t0 = time.time()
    mesh = S.box_mesh(n_hex, conductor=cond)
    spec = S.box_spec(n_hex, conductor=cond)
    half = n_part // 2
    QE, ME, MP, EPS0 = S.ELEMENTARY_CHARGE, S.ELECTRON_MASS, S.PROTON_MASS, S.VACUUM_PERMITTIVITY
    Te = 1.0
    dens = 1e10
    vth_e = np.sqrt(Te * QE / ME)
    vth_i = np.sqrt(Te * QE / MP)
    xe, ve, ce = S.maxwellian_box(spec, half, seed, 0, vth_e)
    xi, vi, ci = S.maxwellian_box(spec, half, seed, 1, vth_i)
    w = dens * (1.0 - 0.2 ** 3) / half
    wpe = np.sqrt(dens * QE ** 2 / (EPS0 * ME))
    dt = 0.2 / wpe
    log(f"case: {n_hex}^3 hexes, {len(mesh['tets'])} tets, {len(mesh['kind'])} nodes, "
        f"{n_part} particles, dt={dt:.4e}s, built in {time.time() - t0:.1f}s")
    return dict(mesh=mesh, spec=spec, e=(xe, ve, ce), i=(xi, vi, ci), w=w, dt=dt,
                qm=[(-QE, ME), (QE, MP)])


def structural_nnz(tets, n_nodes):
    """Entries of the node-adjacency pattern (self + neighbours sharing a tet)."""
    e = np.concatenate([tets[:, [a, b]] for a in range(4) for b in range(a + 1, 4)])
    e.sort(axis=1)
    key = e[:, 0].astype(np.int64) * n_nodes + e[:, 1]
    return int(2 * np.unique(key).size + n_nodes)


def box_context(case, sort_interval, row_patterns=True, precond="two-level", cell_order=0):
    """The box workload's mixed-precision context, initialized."""
    import scdat_amd as S
    m = case["mesh"]
    ctx = S.Context(precision=0, dt=case["dt"], pcg_tol=1e-8, sort_interval=sort_interval)
    if precond == "two-level":
        ctx.set_node_order(S.ORDER_TILED)
        ctx.set_preconditioner(S.PRECOND_TWO_LEVEL)
    ctx.set_mesh(m["xyz"], m["tets"], m["kind"], group=m["group"])
    for g in range(6, 12):
        ctx.set_boundary(g, S.BC_CONDUCTOR)
    for (q, mass), (x, v, c) in zip(case["qm"], (case["e"], case["i"])):
        sid = ctx.add_species(q, mass, case["w"], len(x))
        ctx.set_particles(sid, x, v, c)
    ctx.row_patterns = row_patterns
    ctx.cell_order = cell_order
    ctx.initialize()
    return ctx


def run_gpu(case, steps, warmup, sort_interval, kernel_timing=True, row_patterns=True, precond="two-level",
            cell_order=0, graph_steps=0):
    import scdat_amd as S
    t0 = time.time()
    ctx = box_context(case, sort_interval, row_patterns, precond, cell_order)
    chunks = ctx.autotune()
    log(f"gpu setup {time.time() - t0:.1f}s (push tile: {chunks} particles per thread); warmup {warmup} steps")
    for _ in range(warmup):
        ctx.step(1)
    ctx.sync()
    st0 = ctx.stats()
    # every push launch and every 64th PCG iteration carry dispatch-timestamp
    # events (timing every launch would slow the host's launch loop)
    ctx.set_kernel_timing(kernel_timing, period=64)
    # the first steps (the plasma still whole) as a window of their own, then
    # the rest of the run; the timed region is both
    first = min(steps, 50)
    t1 = time.perf_counter()
    ctx.step(first)
    ctx.sync()
    wall_first = time.perf_counter() - t1
    pushes_first = ctx.stats()["particle_pushes"] - st0["particle_pushes"]
    stages = ctx.stage_times()  # (per step call)
    t1b = time.perf_counter()
    if steps > first:
        ctx.step(steps - first)
        ctx.sync()
        stages = {k: v + ctx.stage_times()[k] for k, v in stages.items()}
    wall = wall_first + (time.perf_counter() - t1b)
    st1 = ctx.stats()
    kt = {"push": ctx.kernel_time(S.KERNEL_PUSH), "pcg_iter": ctx.kernel_time(S.KERNEL_PCG)}
    pushes = st1["particle_pushes"] - st0["particle_pushes"]
    log(f"gpu: {steps} steps in {wall:.3f}s, pushes {pushes}, stages(ms) {stages}, kernels {kt}, stats {st1}")
    solver = ctx.solver_info()
    graph = None
    if graph_steps > 0:
        # the same context continued with the launches around the PCG
        # iteration replayed from HIP graphs (no per-kernel events there)
        ctx.set_kernel_timing(False)
        ctx.step_graph = 1
        p0 = ctx.stats()["particle_pushes"]
        t2 = time.perf_counter()
        ctx.step(graph_steps)
        ctx.sync()
        gw = time.perf_counter() - t2
        gp = ctx.stats()["particle_pushes"] - p0
        gst = ctx.stage_times()
        ctx.step_graph = 0
        p1 = ctx.stats()["particle_pushes"]
        t3 = time.perf_counter()
        ctx.step(graph_steps)
        ctx.sync()
        ew = time.perf_counter() - t3
        ep = ctx.stats()["particle_pushes"] - p1
        graph = dict(value=gp / gw, ms_per_step=1e3 * gw / graph_steps, steps=graph_steps, stage_ms_total=gst,
                     eager_after=dict(value=ep / ew, ms_per_step=1e3 * ew / graph_steps))
        log(f"gpu, captured steps: {graph_steps} steps in {gw:.3f}s, pushes {gp}; eager after {ew:.3f}s, {ep}")
    return ctx, dict(wall=wall, pushes=pushes, stages=stages, kt=kt, stats=st1, stats0=st0, chunks=chunks,
                     solver=solver, graph=graph,
                     first=dict(steps=first, value=pushes_first / wall_first, ms_per_step=1e3 * wall_first / first))


def run_cpu_baseline(case, steps, threads):
    import oracle
    m = case["mesh"]
    t0 = time.time()
    o = oracle.Oracle(m["xyz"], m["tets"], m["kind"], group=m["group"], threads=threads)
    for g in range(6, 12):
        o.set_group(g, 3)
    o.set_poisson(0, 8.8541878128e-12, 1, 1e-8, 20000)
    o.set_push(case["dt"], 1, 1e-3, 64, 0)
    for (q, mass), (x, v, c) in zip(case["qm"], (case["e"], case["i"])):
        s = o.add_species(q, mass, case["w"])
        o.set_particles(s, x, v, c)
    o.deposit()
    o.solve()
    o.efield()
    o.step(1)  # warm start for the timed steps, as on the GPU
    setup = time.time() - t0
    alive = [sum(int((o.get_particles(s)[2] >= 0).sum()) for s in range(2))]
    wall = 0.0
    for _ in range(steps):
        t1 = time.perf_counter()
        o.step(1)
        wall += time.perf_counter() - t1
        alive.append(sum(int((o.get_particles(s)[2] >= 0).sum()) for s in range(2)))
    pushes = sum(alive[:-1])
    log(f"cpu baseline: setup {setup:.1f}s, {steps} steps {wall:.2f}s, threads {oracle.lib().orc_num_threads()}")
    _, phi, _ = o.fields()
    return dict(wall=wall, pushes=pushes, threads=oracle.lib().orc_num_threads(), steps=1 + steps, phi=phi,
                vfloat=o.floating_potential())


def run_sweep(args):
    import scdat_amd as S
    import scdat_amd.batch as B
    name, _ = S.device_info()
    t0 = time.time()
    sweep = B.sweep(args.scenarios, args.sweep_seed)
    cases = [B.case_inputs(sc) for sc in sweep]
    n_part = sum(2 * sc["n_particles"] for sc in sweep)
    log(f"sweep: {len(sweep)} cases, {n_part} particles, "
        f"{sum(len(c['mesh']['tets']) for c in cases)} tets, inputs in {time.time() - t0:.1f}s")

    def timed(ctx):
        for _ in range(args.warmup):
            ctx.step(1)
        ctx.sync()
        p0 = ctx.stats()["particle_pushes"]
        t1 = time.perf_counter()
        ctx.step(args.steps)
        ctx.sync()
        wall = time.perf_counter() - t1
        return wall, ctx.stats()["particle_pushes"] - p0

    t0 = time.time()
    bc, u = B.batch_context(sweep, cases, precision=0, tol=1e-8, sort_interval=args.sort_interval, autotune=True)
    log(f"batched setup {time.time() - t0:.1f}s")
    bwall, bpush = timed(bc)
    stages = bc.stage_times()
    conv = [bc.scenario_stats(k) for k in range(len(sweep))]
    log(f"batched: {args.steps} steps {bwall:.3f}s, pushes {bpush}, stages {stages}")
    # the same context continued with the steps replayed from HIP graphs,
    # then eagerly again (the particle count falls as the walls absorb, so
    # the captured window is compared with the eager windows either side)
    bc.step_graph = 1
    gwall, gpush = timed(bc)
    gstages = bc.stage_times()
    bc.step_graph = 0
    ewall, epush = timed(bc)
    log(f"batched, captured steps: {args.steps} steps {gwall:.3f}s, pushes {gpush}, stages {gstages}; "
        f"eager after: {ewall:.3f}s, pushes {epush}")
    bc.close()
    del bc
    swall, spush = 0.0, 0
    t0 = time.time()
    for k, (sc, st) in enumerate(zip(sweep, cases)):
        c = B.standalone_context(sc, st, precision=0, tol=1e-8, sort_interval=args.sort_interval, autotune=True,
                                 like_batch=False)
        w, p = timed(c)
        swall += w
        spush += p
        del c
        if k % 32 == 31:
            log(f"sequential: {k + 1} cases, {swall:.3f}s timed so far")
    log(f"sequential: {args.steps} steps per case {swall:.3f}s, pushes {spush}, total {time.time() - t0:.1f}s")
    value = bpush / bwall
    res = {
        "metric": "batched charging sweep: particle-steps/s over all cases; speed-up vs the cases run one by one",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * bwall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (push/SpMV) + f64 (refinement, dots, conductor)",
        "data": "synthetic Maxwellian plasmas (sweep seed %d)" % args.sweep_seed,
        "config": {"workload": "%d charging cases, box 1m with floating 0.2m conductor, edge 8..32 hexes, "
                               "2e4..2e6 particles per species, Te 1eV..20keV, n 1e6..1e11, Sternglass SEE "
                               "delta_max 1..3, PCG 1e-8; %d nodes, %d particles in all"
                               % (len(sweep), int(u["node_start"][-1]), n_part)},
        "device": name,
        "sequential": {"value": spush / swall, "ms_per_step": 1e3 * swall / args.steps,
                       "note": "the same cases, each in its own context, one after another (timed steps only)"},
        "speedup_vs_sequential": value / (spush / swall),
        "step_graph": {"value": gpush / gwall, "ms_per_step": 1e3 * gwall / args.steps, "stage_ms_total": gstages,
                       "eager_after": {"value": epush / ewall, "ms_per_step": 1e3 * ewall / args.steps},
                       "note": "the same context continued for as many steps with the launches around the "
                               "PCG iteration replayed from HIP graphs (SCDAT_SHAPE_STEP_GRAPH), then as many "
                               "eager steps again; value and ms_per_step above are the first eager window"},
        "stage_ms_total": stages,
        "cases_converged": int(sum(1 for q in conv if q["pcg_converged"])),
        "pcg_iterations_max": int(max(q["pcg_iterations"] for q in conv)),
    }
    print(json.dumps(res), flush=True)


def run_sampler(args):
    """Live-PIC surface-current sampler (PicMccSurfaceCurrentSampler::sample
    and sampleWithDerivative) at the reference's default window: wall time
    per call, one window and the three concurrent windows of dJ/dV."""
    import scdat_amd as S
    name, _ = S.device_info()
    cfg = S.sampler_config()
    _, ne, ni, _ = S.sampler_inputs(cfg)
    S.sample_surface_current(cfg, 0.0)  # warm-up: first contexts, code objects
    reps = max(1, args.steps // 10)
    t0 = time.perf_counter()
    for _ in range(reps):
        one = S.sample_surface_current(cfg, 0.0)
    t_one = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        three = S.sample_surface_current(cfg, 1.0)
    t_three = (time.perf_counter() - t0) / reps
    pushes = (ne + ni) * cfg.window_steps
    log(f"sampler: {ne}+{ni} macro-particles, {cfg.window_steps} steps; one window {t_one * 1e3:.1f} ms, "
        f"three {t_three * 1e3:.1f} ms")
    print(json.dumps({
        "metric": "live-PIC surface-current sample with dJ/dV (three concurrent windows): wall time per call",
        "value": t_three, "unit": "s", "n_gpus": 1, "steps": cfg.window_steps, "warmup": 1,
        "ms_per_step": 1e3 * t_three / cfg.window_steps, "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 (particles) + f64 (field refinement, currents)",
        "data": "synthetic (reference default sampler configuration, seeded loading)",
        "config": {"workload": "PicMccSurfaceSamplerConfig defaults: %d macro-electrons + %d macro-ions, %d steps, "
                               "argon MCC %s" % (ne, ni, cfg.window_steps, "on" if cfg.enable_mcc else "off")},
        "device": name,
        "one_window_s": t_one, "three_windows_s": t_three,
        "concurrency_gain": 3.0 * t_one / t_three if t_three > 0 else None,
        "particle_steps_per_s_three_windows": 3.0 * pushes / t_three if t_three > 0 else None,
        "sample": {k: one[k] for k in ("net_collection_current_density_a_per_m2", "total_collisions")},
        "derivative_a_per_m2_per_v": three["current_derivative_a_per_m2_per_v"],
    }), flush=True)


def run_arc(args):
    """Vacuum-arc PIC boundary window: wall time per advance (set the
    cathode current density, seed the probe protons, one PIC step, ledger
    currents), 1e-10 s advances with about 4 field electrons per ion impact."""
    import scdat_amd as S
    from scdat_amd import arc
    name, _ = S.device_info()
    cfg = arc.WindowConfig(surface_potential_v=-200.0, enable_secondary_electron_emission=True,
                           secondary_electron_yield_per_ion=0.5)
    w = arc.BoundaryWindow(cfg)
    dt = 1.0e-10
    j = 4.0 * S.ELEMENTARY_CHARGE / (w.area * dt)
    for _ in range(args.warmup):
        w.advance(j, dt)
    n = max(1, args.steps)
    t0 = time.perf_counter()
    for _ in range(n):
        st = w.advance(j, dt)
    wall = (time.perf_counter() - t0) / n
    led = w.ctx.ledger(6)
    log(f"arc window: {n} advances, {wall * 1e3:.3f} ms each, ledger {led[arc.ZMIN]}")
    print(json.dumps({
        "metric": "vacuum-arc PIC boundary window: wall time per advance",
        "value": wall, "unit": "s", "n_gpus": 1, "steps": n, "warmup": args.warmup,
        "ms_per_step": 1e3 * wall, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic (probe protons of the reference's seeding rule)",
        "config": {"workload": "plate 1e-4 x 1e-4 x 5e-4 m, 1x1x6 hexes, cathode -200 V, J for 4 field electrons "
                               "per impact, secondary yield 0.5, dt 1e-10 s"},
        "device": name,
        "cathode_ion_impacts": led[arc.ZMIN]["absorbed_ion_particles"],
        "emitted_electrons": led[arc.ZMIN]["emitted_electron_particles"],
        "last_status": st.__dict__,
    }), flush=True)
    w.close()


def run_adjoint(args):
    import scdat_amd as S
    from scdat_amd.adjoint import central_difference
    name, _ = S.device_info()
    n_hex = 32 if args.hex == 64 else args.hex
    n_part = (1 << 20) if args.particles == (1 << 24) else args.particles
    steps, interval = args.adjoint_steps, 20
    case = build_case(n_hex, n_part, args.seed)
    m = case["mesh"]

    def ctx_for(precision):
        c = S.Context(precision=precision, dt=case["dt"], pcg_tol=1e-12, sort_interval=0, adaptive_substep=0)
        c.set_mesh(m["xyz"], m["tets"], m["kind"], group=m["group"])
        for g in range(6, 12):
            c.set_boundary(g, S.BC_CONDUCTOR)
        for (q, mass), (x, v, cl) in zip(case["qm"], (case["e"], case["i"])):
            sid = c.add_species(q, mass, case["w"], len(x))
            c.set_particles(sid, x, v, cl)
        c.initialize()
        return c

    p = S.charging_params(see_delta_max=2.0, see_energy_max_ev=400.0,
                          face_yield=[1.0, 0.8, 1.2, 0.9, 1.1, 1.05], photo_current_density=2e-5,
                          photo_temperature_ev=2.0, photo_reference_v=0.0, sunlit_faces=0b000010,
                          surface_conductance_s=1e-9)
    c64 = ctx_for(1)
    t0 = time.perf_counter()
    v_fwd = c64.charging_forward(p, steps)
    t_fwd = time.perf_counter() - t0
    t0 = time.perf_counter()
    v64, g64 = c64.charging_adjoint(p, steps, interval)
    t_adj = time.perf_counter() - t0
    log(f"adjoint fp64: forward {t_fwd:.2f}s, gradient {t_adj:.2f}s, V {v64}, grad {g64}")
    theta = p.vector()
    t0 = time.perf_counter()
    fd = central_difference(c64, theta, steps, rel=1e-4, base=p)
    t_fd = time.perf_counter() - t0
    rel = np.abs(theta * (g64 - fd)) / np.maximum(np.abs(theta * fd), 1e-300)
    log(f"central differences h=1e-4 {t_fd:.1f}s: {fd}; relative differences {rel}")
    fd8 = central_difference(c64, theta, steps, rel=1e-8, base=p)
    rel8 = np.abs(theta * (g64 - fd8)) / np.maximum(np.abs(theta * fd8), 1e-300)
    log(f"central differences h=1e-8: {fd8}; relative differences {rel8}")
    del c64
    c32 = ctx_for(0)
    t0 = time.perf_counter()
    v32, g32 = c32.charging_adjoint(p, steps, interval)
    t32 = time.perf_counter() - t0
    err32 = float(np.abs(theta * (g32 - g64)).max() / np.abs(theta * g64).max())
    names = ["delta_max", "photo_j0", "conductance"] + ["face_yield_%d" % k for k in range(6)]
    res = {
        "metric": "discrete adjoint: wall time of dV_float/d(9 parameters) at 32^3/1M ptcl, 200 steps",
        "value": t_adj,
        "unit": "s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": 0,
        "ms_per_step": 1e3 * t_adj / steps,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 (gradient check); f32 particles + f64 refinement (mixed run)",
        "data": "synthetic Maxwellian plasma (Philox seed %d)" % args.seed,
        "config": {"workload": "charging adjoint, box 1m, floating 0.2m conductor, %d^3 hex, %d particles, "
                               "Te=Ti=1eV, n=1e10, dt=0.2/wpe, %d steps, checkpoint every %d, PCG 1e-12"
                               % (n_hex, n_part, steps, interval)},
        "device": name,
        "floating_potential_v": v64,
        "forward_only_s": t_fwd,
        "forward_matches_adjoint_run": bool(abs(v_fwd - v64) <= 1e-9 * abs(v64)),
        "gradient": dict(zip(names, g64.tolist())),
        "central_difference_h1e-4": dict(zip(names, fd.tolist())),
        "rel_diff_h1e-4": dict(zip(names, rel.tolist())),
        "central_difference_h1e-8": dict(zip(names, fd8.tolist())),
        "rel_diff_h1e-8": dict(zip(names, rel8.tolist())),
        "sensitivity_rel_diff_max_h1e-8": float(rel8.max()),
        "sensitivity_rel_diff_max_best_h": float(np.minimum(rel, rel8).max()),
        "central_difference_note": "both perturbed runs replay the absorption events of the run at theta, so "
                                   "the quotient differences the map the adjoint differentiates; what remains "
                                   "shrinks with h (the cell weights are piecewise linear in position: a "
                                   "perturbed trajectory that enters another tet changes the slope the quotient "
                                   "averages); the conductance's h = 1e-8 step moves V by ~1e-12 V, at the PCG "
                                   "tolerance",
        "central_difference_s": t_fd,
        "mixed_precision": {"floating_potential_v": v32, "gradient_s": t32, "sensitivity_err_vs_fp64": err32},
    }
    print(json.dumps(res), flush=True)


def _geo_oracle(st, threads):
    """The FP64 oracle of a GEO case (CPU baseline leg only)."""
    import oracle
    import scdat_amd as S
    import scdat_amd.geo as G
    cfg = st["cfg"]
    m = st["mesh"]
    o = oracle.Oracle(m["xyz"], m["tets"], m["kind"], group=m["group"], threads=threads)
    for g in G.CONDUCTOR_GROUPS:
        o.set_group(g, S.BC_CONDUCTOR)
    o.set_poisson(0, S.VACUUM_PERMITTIVITY, 1, 1e-8, 20000)
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


def run_geo(args):
    import scdat_amd as S
    import scdat_amd.geo as G
    name, _ = S.device_info()
    t0 = time.time()
    st = G.inputs()
    m = st["mesh"]
    n_nodes, n_tets = len(m["kind"]), len(m["tets"])
    steps = args.geo_steps
    log(f"geo: {st['cfg']['edge']}^3 hexes, {n_tets} tets, {n_nodes} nodes, {st['cfg']['n_particles']} particles, "
        f"dt={st['dt']:.4e}s, inputs in {time.time() - t0:.1f}s")
    t0 = time.time()
    ctx = G.context(st, precision=0, tol=1e-8, sort_interval=args.sort_interval, steps=steps + args.warmup,
                    two_level=args.precond == "two-level")
    chunks = ctx.autotune()
    log(f"geo setup {time.time() - t0:.1f}s (push tile {chunks}); warmup {args.warmup} steps")
    for _ in range(args.warmup):
        ctx.step(1)
    ctx.sync()
    st0, ev0 = ctx.stats(), ctx.event_counts()
    ctx.set_kernel_timing(not args.no_kernel_timing, period=64)
    t1 = time.perf_counter()
    ctx.step(steps)
    ctx.sync()
    wall = time.perf_counter() - t1
    st1, ev1 = ctx.stats(), ctx.event_counts()
    stages = ctx.stage_times()
    push_ms, push_n = ctx.kernel_time(S.KERNEL_PUSH)
    pushes = st1["particle_pushes"] - st0["particle_pushes"]
    led = ctx.ledger(12)
    log(f"geo: {steps} steps in {wall:.3f}s, pushes {pushes}, stages {stages}, stats {st1}")
    # push + gather + deposit bytes per launch: particle SoA read + write
    # 2 x 28 B, the id 4 B for the collision draw, tet map 64 B per tet,
    # nodal E 16 B + integer charge 8 B per node
    per_launch = pushes / max(1, push_n)
    push_bytes = 60.0 * per_launch + 64.0 * n_tets + 24.0 * n_nodes
    gbs = push_bytes / (push_ms * 1e-3) / 1e9 if push_ms > 0 else 0.0
    coll = [ev1["collisions"][k] - ev0["collisions"][k] for k in range(3)]
    res = {
        "metric": "particle-steps/s incl. Poisson, GEO charging case (MCC + SEE + photoemission) at 48^3/8M ptcl",
        "value": pushes / wall,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (push/SpMV) + f64 (refinement, dots, conductor, collisions, emission)",
        "data": "synthetic Maxwellian plasma (Philox seed %d)" % st["cfg"]["seed"],
        "config": {"workload": "GEO charging: box 1m, conductor 0.2m floating, %d^3 hex -> %d tets, %d nodes, "
                               "%d particles (e-/H+), Te=Ti=10keV, n=1e6, O 1e12 (e elastic + ionisation, H+ "
                               "elastic; null-collision), Sternglass SEE 2.0@400eV, photo 20uA/m^2 on +x, "
                               "dt=%.3e s, PCG 1e-8" % (st["cfg"]["edge"], n_tets, n_nodes,
                                                        st["cfg"]["n_particles"], st["dt"])},
        "device": name,
        "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": gbs / HBM_PEAK_GBS, "traffic": None, "avg_launch_ms": push_ms, "launches_timed": push_n,
                     "algorithmic_bytes_per_launch": push_bytes, "kernel": "push_kernel<float,CH>"},
        "stage_ms_total": stages,
        "events": {"collisions_elastic_e": coll[0], "collisions_ionisation": coll[1], "collisions_elastic_h": coll[2],
                   "mcc_candidates": ev1["mcc_candidates"] - ev0["mcc_candidates"],
                   "mcc_majorant_exceeded": ev1["mcc_majorant_exceeded"],
                   "see_emitted": ev1["see_emitted"] - ev0["see_emitted"],
                   "photo_emitted": ev1["photo_emitted"] - ev0["photo_emitted"]},
        "active_particles_end": st1["active_particles"],
        "floating_potential_v": st1["floating_potential"],
        "conductor_absorbed": {"electrons": sum(led[g]["absorbed_electron_particles"] for g in range(6, 12)),
                               "ions": sum(led[g]["absorbed_ion_particles"] for g in range(6, 12))},
        "pcg_failed_solves": st1["pcg_failed_solves"],
    }
    ctx.close()
    if not args.no_cpu_baseline:
        o = _geo_oracle(st, args.cpu_threads)
        o.deposit()
        o.solve()
        o.efield()
        o.step(1)
        cw, cp = 0.0, 0
        for _ in range(args.cpu_steps // 2):
            alive = sum(int((o.get_particles(k)[2] >= 0).sum()) for k in range(3))
            t1 = time.perf_counter()
            o.step(1)
            cw += time.perf_counter() - t1
            cp += alive
        import oracle
        res["cpu_baseline"] = {"value": cp / cw, "unit": "particle-steps/s",
                               "cores": oracle.lib().orc_num_threads(), "kind": "port", "cpu_model": cpu_model(),
                               "sample": "%d full steps of the same case (same inputs, after one warm step), FP64 "
                                         "oracle with OpenMP" % (args.cpu_steps // 2)}
        res["gpu_over_cpu"] = res["value"] / res["cpu_baseline"]["value"]
        log(f"cpu baseline: {args.cpu_steps // 2} steps {cw:.2f}s")
    print(json.dumps(res), flush=True)


def run_fp64(args):
    """The reference-harness case in FP64: 16^3 hexes, 10^5 electrons + 10^5
    protons, 100 collisionless steps, Jacobi-PCG to 1e-10; the FP64 oracle
    runs the same steps from the same inputs (the CPU baseline, and the
    error figures)."""
    import scdat_amd as S
    name, _ = S.device_info()
    steps = args.fp64_steps
    case = build_case(16, 200_000, 20261203)
    m = case["mesh"]
    ctx = S.Context(precision=1, dt=case["dt"], pcg_tol=1e-10, sort_interval=args.sort_interval)
    ctx.set_mesh(m["xyz"], m["tets"], m["kind"], group=m["group"])
    for g in range(6, 12):
        ctx.set_boundary(g, S.BC_CONDUCTOR)
    for (q, mass), (x, v, c) in zip(case["qm"], (case["e"], case["i"])):
        sid = ctx.add_species(q, mass, case["w"], len(x))
        ctx.set_particles(sid, x, v, c)
    ctx.initialize()
    chunks = ctx.autotune()
    ctx.sync()
    st0 = ctx.stats()
    t1 = time.perf_counter()
    ctx.step(steps)
    ctx.sync()
    wall = time.perf_counter() - t1
    st1 = ctx.stats()
    pushes = st1["particle_pushes"] - st0["particle_pushes"]
    _, phi_g, _ = ctx.fields()
    stages = ctx.stage_times()
    log(f"fp64: {steps} steps in {wall:.3f}s, pushes {pushes}, stages {stages}")
    res = {
        "metric": "particle-steps/s incl. Poisson, FP64 reference-harness case 16^3 / 2x10^5 ptcl",
        "value": pushes / wall,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": 0,
        "ms_per_step": 1e3 * wall / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic Maxwellian plasma (Philox seed 20261203)",
        "config": {"workload": "pic box 1m, conductor 0.2m floating, 16^3 hex -> %d tets, %d nodes, 2x10^5 "
                               "particles (e-/H+), Te=Ti=1eV, n=1e10, dt=0.2/wpe, %d steps, Jacobi-PCG 1e-10, FP64"
                               % (len(m["tets"]), len(m["kind"]), steps)},
        "device": name,
        "launch_shape": {"push_particles_per_thread": chunks},
        "stage_ms_total": stages,
        "floating_potential_v": st1["floating_potential"],
    }
    ctx.close()
    if not args.no_cpu_baseline:
        import oracle
        o = oracle.Oracle(m["xyz"], m["tets"], m["kind"], group=m["group"], threads=args.cpu_threads)
        for g in range(6, 12):
            o.set_group(g, 3)
        o.set_poisson(0, S.VACUUM_PERMITTIVITY, 1, 1e-10, 50000)
        o.set_push(case["dt"], 1, 1e-3, 64, 0)
        for (q, mass), (x, v, c) in zip(case["qm"], (case["e"], case["i"])):
            k = o.add_species(q, mass, case["w"])
            o.set_particles(k, x, v, c)
        o.deposit()
        o.solve()
        o.efield()
        cw, cp = 0.0, 0
        for _ in range(steps):
            alive = sum(int((o.get_particles(k)[2] >= 0).sum()) for k in range(2))
            t1 = time.perf_counter()
            o.step(1)
            cw += time.perf_counter() - t1
            cp += alive
        _, phi_o, _ = o.fields()
        vo = o.floating_potential()
        res["cpu_baseline"] = {"value": cp / cw, "unit": "particle-steps/s", "cores": oracle.lib().orc_num_threads(),
                               "kind": "port", "cpu_model": cpu_model(),
                               "sample": "the same %d steps from the same inputs, FP64 oracle with OpenMP" % steps}
        res["gpu_over_cpu"] = res["value"] / res["cpu_baseline"]["value"]
        res["error_vs_fp64_oracle"] = {
            "steps": steps,
            "phi_rel_max": float(np.max(np.abs(phi_g - phi_o)) / max(np.max(np.abs(phi_o)), 1e-300)),
            "floating_potential_rel": abs(st1["floating_potential"] - vo) / max(abs(vo), 1e-300)}
        log(f"cpu baseline: {steps} steps {cw:.2f}s")
    print(json.dumps(res), flush=True)


def copy_calibration(mb):
    import torch
    n = int(mb * 1e6) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_read": 4 * n, "copy_calibration_bytes_written": 4 * n}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default: 500 for the box, the configuration's run; 50 otherwise)")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--hex", type=int, default=64)
    ap.add_argument("--particles", type=int, default=1 << 24)
    ap.add_argument("--seed", type=int, default=20261204)
    ap.add_argument("--sort-interval", type=int, default=20)
    ap.add_argument("--cpu-steps", type=int, default=16)
    ap.add_argument("--cpu-threads", type=int, default=host_threads(),
                    help="OpenMP threads of the CPU baseline (default: OMP_NUM_THREADS, else the CPUs this "
                         "process may run on)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-kernel-timing", action="store_true",
                    help="no per-launch events in the timed region (no roofline figures)")
    ap.add_argument("--copy-calibration", type=float, default=0.0)
    ap.add_argument("--workload", choices=("box", "sweep", "adjoint", "sampler", "arc", "geo", "fp64"),
                    default="box")
    ap.add_argument("--fp64-steps", type=int, default=100)
    ap.add_argument("--geo-steps", type=int, default=2000)
    ap.add_argument("--adjoint-steps", type=int, default=200)
    ap.add_argument("--scenarios", type=int, default=256)
    ap.add_argument("--sweep-seed", type=int, default=20261206)
    ap.add_argument("--row-patterns", type=int, choices=(0, 1), default=1,
                    help="1: the FP32 inner iteration may read the operator as row patterns")
    ap.add_argument("--precond", choices=("jacobi", "two-level"), default="two-level",
                    help="Poisson preconditioner; two-level uses the tiled node order")
    ap.add_argument("--cell-order", type=int, choices=(0, 1), default=0,
                    help="order of the cells along the sorted particles: 0 tet index, 1 Morton")
    ap.add_argument("--graph-steps", type=int, default=50,
                    help="box: after the timed steps, this many more with the launches around the PCG "
                         "iteration replayed from HIP graphs (reported under step_graph; 0: skip)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)
    if args.copy_calibration > 0:
        copy_calibration(args.copy_calibration)
        return
    if args.steps is None:
        args.steps = 500 if args.workload == "box" else 50
    if args.workload == "sweep":
        run_sweep(args)
        return
    if args.workload == "adjoint":
        run_adjoint(args)
        return
    if args.workload == "sampler":
        run_sampler(args)
        return
    if args.workload == "arc":
        run_arc(args)
        return
    if args.workload == "geo":
        run_geo(args)
        return
    if args.workload == "fp64":
        run_fp64(args)
        return

    import scdat_amd as S
    name, ncu = S.device_info()
    case = build_case(args.hex, args.particles, args.seed)
    ctx, g = run_gpu(case, args.steps, args.warmup, args.sort_interval, not args.no_kernel_timing,
                     bool(args.row_patterns), args.precond, args.cell_order, args.graph_steps)

    m = case["mesh"]
    n_nodes, n_tets = len(m["kind"]), len(m["tets"])
    nnz = structural_nnz(m["tets"], n_nodes)
    kt = g["kt"]
    # Algorithmic bytes per launch.
    #  push (one species): particle SoA read + write 2 x 28 B; tet map 64 B
    #    per tet; nodal E 16 B + integer charge 8 B per node
    #  PCG iteration (one fused kernel): r, w, s_old, p, x read and r, w, s,
    #    p, x written, 4 B each per row; the operator either as ELL (column +
    #    value 8 B per structural entry, D^-1 4 B per row) or as row patterns
    #    (2-byte id per row, 4-byte past 65536 patterns, plus the table once:
    #    W offsets and values and D^-1 per pattern)
    #  tile-staged PCG iteration: a 2-byte box pattern id per row, the table
    #    once (16 direction codes, W values of A D^-1 and of A P, D^-1 per
    #    pattern) and 27 neighbour blocks per row block
    #  two-level coarse step (inside the iteration): Ac^-1 once (nc rows of
    #    ncp FP32 columns), zr and zs, the 4 G partials
    push_launch = g["pushes"] / max(1, kt["push"][1]) if kt["push"][1] else 0
    sv = g["solver"]
    npat, width = sv["row_patterns"], sv["ell_width"]
    nblk = -(-n_nodes // 512)
    if sv["tile_staged"]:
        op_bytes = 2.0 * n_nodes + sv["box_patterns"] * (16.0 + 8.0 * width + 4.0) + 27 * 4.0 * nblk
    elif npat:
        id_bytes = 2.0 if npat <= 65536 else 4.0
        op_bytes = id_bytes * n_nodes + npat * ((12.0 if sv["coarse_in_kernel"] else 8.0) * width + 4.0)
    else:
        op_bytes = 8.0 * nnz + 4.0 * n_nodes
    nc = sv["coarse_unknowns"]
    coarse_bytes = (4.0 * nc * (-(-nc // 64) * 64) + 8.0 * nc + 32.0 * nblk) if sv["coarse_in_kernel"] else 0.0
    bytes_k = {
        "push": 56.0 * push_launch + 64.0 * n_tets + 24.0 * n_nodes,
        "pcg_iter": op_bytes + 40.0 * n_nodes + coarse_bytes,
    }
    # the reference's CSR iteration at this size (8 B per structural entry,
    # row pointers, seven vectors of 8 B): the bytes the kernel replaces
    csr_bytes = 8.0 * nnz + 4.0 * (n_nodes + 1) + 56.0 * n_nodes
    pcg_kernel = ("k_cg_iter_box<%d,%s>" % (width, "true" if sv["coarse_in_kernel"] else "false")
                  if sv["tile_staged"] else
                  "k_cg_iter<float,W,false,%s,%d>" % ("true" if npat else "false", sv["pcg_block"]))
    traffic, traffic_src = {}, {}
    if os.path.exists(TRAFFIC_FILE):
        tf = json.load(open(TRAFFIC_FILE))
        kn = tf.get("kernel_names", {})
        for k, v in tf.get("traffic_bytes_per_launch", {}).items():
            if k == "pcg_iter" and kn.get(k) != pcg_kernel:
                continue  # measured on another form of the iteration
            traffic[k] = v
            traffic_src[k] = ("profiles/pmc_traffic_64cube.json: FETCH_SIZE + WRITE_SIZE per launch of %s, "
                              "scaled by a device copy, over the first 20 steps (the whole plasma; commands in "
                              "profiles/pmc_traffic_64cube_16M.md)" % kn.get(k, k))
    # launches that did work in the timed region (PCG: one per inner iteration;
    # every 16th carries timing events)
    iters = g["stats"]["pcg_iterations_total"] - g["stats0"]["pcg_iterations_total"]
    launches = {"push": kt["push"][1], "pcg_iter": iters}
    kernels = {}
    for k, (ms, n) in kt.items():
        gbs = bytes_k[k] / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
        kernels[k] = {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                      "frac": gbs / HBM_PEAK_GBS, "traffic": traffic.get(k), "traffic_source": traffic_src.get(k),
                      "avg_launch_ms": ms, "launches": launches[k], "launches_timed": n,
                      "algorithmic_bytes_per_launch": bytes_k[k], "total_ms": ms * launches[k]}
        if traffic.get(k) and ms > 0:
            kernels[k]["moved_frac"] = traffic[k] / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS
    if kt["pcg_iter"][0] > 0:
        kernels["pcg_iter"]["csr_bytes_per_launch"] = csr_bytes
        kernels["pcg_iter"]["csr_frac"] = csr_bytes / (kt["pcg_iter"][0] * 1e-3) / 1e9 / HBM_PEAK_GBS
    dominant = max(kernels, key=lambda k: kernels[k]["total_ms"])
    names = {"push": "push_kernel<float,CH>", "pcg_iter": pcg_kernel}
    roof = dict(kernels[dominant])
    roof["kernel"] = names[dominant]
    st = g["stats"]
    pcg_ms = g["stages"]["pcg"]
    pcg_bytes = iters * bytes_k["pcg_iter"]
    value = g["pushes"] / g["wall"]
    res = {
        "metric": "particle-steps/s incl. Poisson at 64^3/16M ptcl; CG HBM GB/s vs peak",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * g["wall"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (push/SpMV) + f64 (refinement, dots, conductor)",
        "data": "synthetic Maxwellian plasma (Philox seed %d)" % args.seed,
        "config": {"workload": "pic-mcc box 1m, conductor 0.2m floating, %d^3 hex -> %d tets, "
                               "%d nodes, %d particles (e-/H+), Te=Ti=1eV, n=1e10, dt=0.2/wpe, "
                               "PCG 1e-8, %d steps" % (args.hex, n_tets, n_nodes, args.particles, args.steps)},
        "device": name,
        "roofline": roof,
        "kernels": kernels,
        "cg": {"iterations_per_step": iters / args.steps, "gbs_effective": pcg_bytes / (pcg_ms * 1e-3) / 1e9
               if pcg_ms > 0 else 0.0, "frac_of_hbm_peak": pcg_bytes / (pcg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS
               if pcg_ms > 0 else 0.0, "note": "algorithmic bytes of the PCG iteration kernel per inner iteration over the "
               "whole Poisson stage time (FP64 residuals, launches and host polls included)"},
        "stage_ms_total": g["stages"],
        "step_graph": None if g["graph"] is None else dict(g["graph"], note=(
            "the same context continued for as many steps with the launches around the PCG iteration "
            "replayed from HIP graphs (SCDAT_SHAPE_STEP_GRAPH), then as many eager steps (eager_after); "
            "the particle count falls as the walls absorb, so compare with the eager windows either side")),
        "launch_shape": {"push_particles_per_thread": g["chunks"], "chosen_by": "scdat_autotune",
                         "cell_order": "morton" if args.cell_order == 1 else "tet index"},
        "solver": {"rows": sv["rows"], "ell_width": width, "row_patterns": npat, "pcg_block": sv["pcg_block"],
                   "preconditioner": "jacobi" if sv["preconditioner"] == S.PRECOND_JACOBI else "two-level",
                   "coarse_unknowns": nc, "coarse_in_kernel": sv["coarse_in_kernel"],
                   "tile_staged": sv["tile_staged"], "box_patterns": sv["box_patterns"]},
        "floating_potential_v": st["floating_potential"],
        "pcg_initial_rel_residual": st["pcg_initial_rel_residual"],
        "active_particles_end": st["active_particles"],
        "first_steps": dict(g["first"], note="the first steps of the timed run on their own (the plasma still "
                                              "whole); the walls absorb most electrons over the run"),
    }
    if not args.no_cpu_baseline:
        c = run_cpu_baseline(case, args.cpu_steps, args.cpu_threads)
        res["cpu_baseline"] = {"value": c["pushes"] / c["wall"], "unit": "particle-steps/s",
                               "cores": c["threads"], "kind": "port", "cpu_model": cpu_model(),
                               "sample": "%d full steps of the same workload (same inputs, after one warm "
                                         "step), FP64 oracle with OpenMP, Jacobi-PCG to 1e-8" % args.cpu_steps}
        res["gpu_over_cpu"] = value / res["cpu_baseline"]["value"]
        # this path against the FP64 oracle over the oracle's own steps
        ctx.close()
        chk = box_context(case, args.sort_interval, bool(args.row_patterns), args.precond, args.cell_order)
        chk.step(c["steps"])
        _, phi_g, _ = chk.fields()
        vf_g = chk.stats()["floating_potential"]
        chk.close()
        res["error_vs_fp64_oracle"] = {
            "steps": c["steps"],
            "phi_rel_l2": float(np.linalg.norm(phi_g - c["phi"]) / np.linalg.norm(c["phi"])),
            "floating_potential_v": vf_g, "oracle_floating_potential_v": c["vfloat"],
            "floating_potential_rel": abs(vf_g - c["vfloat"]) / abs(c["vfloat"]),
            "note": "same inputs, this path (mixed precision) and the FP64 oracle from the initial state; "
                    "plasma trajectories decorrelate slowly, so the figure grows with the step count"}
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
