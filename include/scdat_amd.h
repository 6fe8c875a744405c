/*
 * scdat_amd.h - C ABI of the MI355X (gfx950) PIC-MCC cycle.
 *
 * Plain C types only: opaque handles, POD descriptors, caller-owned host
 * buffers copied in and out, int status returns (0 ok, < 0 error, message via
 * scdat_last_error).  No exception crosses this boundary.  One context owns
 * all device memory of one simulation and runs on one HIP stream; contexts
 * share nothing, so several may coexist in one process.
 *
 * Each entry point names the reference interface it replaces
 * (paths relative to the SCDAT source tree).
 */
#ifndef SCDAT_AMD_H
#define SCDAT_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define SCDAT_OK 0
#define SCDAT_ERR_ARG (-1)     /* invalid argument                          */
#define SCDAT_ERR_STATE (-2)   /* call out of order (e.g. step before mesh) */
#define SCDAT_ERR_DEVICE (-3)  /* HIP runtime error / no device             */
#define SCDAT_ERR_NOMEM (-4)   /* allocation failed / capacity exceeded     */
#define SCDAT_ERR_SOLVER (-5)  /* Poisson PCG failed; previous phi kept     */

/* ---- boundary groups ----------------------------------------------------
 * Every boundary face of the mesh carries a group id.  The box mesher uses
 * 0..5 for the outer walls in the order of PICCycle::BoundaryFace
 * (XMin, XMax, YMin, YMax, ZMin, ZMax; Tools/PICcore/include/PICCycle.h:36-44)
 * and 6..11 for the faces of the conductor box (-x,+x,-y,+y,-z,+z).          */
#define SCDAT_MAX_GROUPS 16
#define SCDAT_BC_ABSORB 0     /* AbsorbingBoundaryCondition, p = 1
                                 (Tools/Boundary/src/BoundaryConditions.cpp:360-392) */
#define SCDAT_BC_PERIODIC 1   /* PeriodicBoundaryCondition (:68-130)                 */
#define SCDAT_BC_REFLECT 2    /* ReflectiveBoundaryCondition (specular)              */
#define SCDAT_BC_CONDUCTOR 3  /* absorbed onto the floating conductor: its charge
                                 feeds the floating-potential solve                  */
#define SCDAT_BC_ESCAPE 4     /* no condition set (the default): the particle leaves,
                                 marked ESCAPED, no ledger entry or surface charge
                                 (PICCycle.cpp:1254-1257,1355-1377)                 */
#define SCDAT_BC_HOST 5       /* a caller-side condition (a BoundaryConditionBase
                                 subclass with no device form): the particle leaves
                                 the device at the face crossing and its impact is
                                 kept for the caller (scdat_get_host_impacts), who
                                 runs processParticle and appends what survives or
                                 is emitted; no ledger entry or surface charge on
                                 the device (PICCycle.cpp:1221-1293)                */

/* node kinds */
#define SCDAT_NODE_FREE 0
#define SCDAT_NODE_DIRICHLET 1
#define SCDAT_NODE_CONDUCTOR 2

typedef struct scdat_ctx scdat_ctx;

/* Physical constants of Tools/Basic/include/Constants.h:46-53 (SI). */
#define SCDAT_ELEMENTARY_CHARGE 1.602176634e-19
#define SCDAT_ELECTRON_MASS 9.1093837015e-31
#define SCDAT_PROTON_MASS 1.67262192369e-27
#define SCDAT_VACUUM_PERMITTIVITY 8.8541878128e-12

/* ---- structured mesh generation (host only, no device needed) ----------
 * Replaces GeometryMeshGenerator::generateBoxMesh
 * (Tools/Mesh/src/MeshParsing.cpp:1208-1261): same lexicographic node order,
 * but a conforming 6-tet (Kuhn) split, an optional conductor box cut out of
 * the domain (its faces snapped to grid lines) and boundary-face groups.   */
typedef struct {
  double lo[3], hi[3];   /* domain box (m)                                  */
  int32_t n[3];          /* hex cells per axis                              */
  int32_t conductor;     /* 1: remove [c_lo, c_hi]; its surface is floating */
  double c_lo[3], c_hi[3];
} scdat_box_spec;

int scdat_box_mesh_size(const scdat_box_spec* spec, int64_t* n_nodes, int64_t* n_tets);
/* xyz[3*n_nodes], tets[4*n_tets], node_kind[n_nodes], face_group[4*n_tets]
 * (face k is opposite local vertex k; -1 = interior face).               */
int scdat_box_mesh_build(const scdat_box_spec* spec, double* xyz, int32_t* tets,
                         int32_t* node_kind, int32_t* face_group);

/* Analytic locate on the box mesh built from `spec` (cell = -1 outside the
 * mesh or inside the conductor).  Replaces the O(N_elem) scan of
 * CrossTetrahedron::locateParticle (PICParticleAlgorithms.cpp:853-891) for
 * structured meshes.                                                       */
int scdat_box_locate(const scdat_box_spec* spec, int64_t n, const double* x, int32_t* cell);

/* Philox4x32-10 counter-based generator (Salmon et al., SC'11). */
void scdat_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* Synthetic Maxwellian plasma on the box mesh: positions uniform in the
 * domain minus the conductor, velocities N(drift, vth^2) per component, cells
 * located analytically.  Particle p's state depends only on (seed, stream, p).
 * Replaces the loading loop of Main/main_PICcore.cpp:195-229.              */
int scdat_load_maxwellian(const scdat_box_spec* spec, int64_t n, uint64_t seed, uint32_t stream,
                          double vth, const double* drift, double* x, double* v, int32_t* cell);

/* ---- context configuration ---------------------------------------------
 * Mirrors PICCycle::Parameters (Tools/PICcore/include/PICCycle.h:46-58) plus
 * the precision, Poisson and sorting controls of the device path.          */
typedef struct {
  int32_t precision;        /* 0 mixed: FP32 particles/SpMV, FP64 refinement;
                               1 FP64 everywhere                                  */
  double dt;                /* time_step (s)                                      */
  int32_t adaptive_substep; /* adaptive sub-stepping of CrossTetrahedron
                               (Tools/Particle/src/PICParticleAlgorithms.cpp:750-801) */
  double floor_ratio;       /* sub-step floor as a fraction of the remaining time */
  int32_t max_substeps;     /* sub-step cap per particle step                     */
  int32_t poisson_operator; /* 0 P1 FEM (MultiBoundaryElectricFieldSolver.cpp:1704-1759),
                               1 graph Laplacian (PoissonSolver.cpp:276-327)      */
  double permittivity;      /* F/m                                                */
  int32_t load_sign;        /* +1 physical load (+rho), -1 reference sign         */
  double pcg_tol;           /* relative residual target ||r|| <= tol ||b||        */
  int32_t pcg_max_iter;     /* total inner iterations per solve                   */
  double inner_tol;         /* mixed mode: FP32 inner PCG relative tolerance (1e-5) */
  int32_t sort_interval;    /* most steps between cell sorts of a species (0 =
                               never); a species whose particles cross cells
                               faster is sorted sooner (about every 8 cells of
                               rms travel, at least every 2 steps)              */
  double bfield[3];         /* uniform magnetic field (T); reference B = 0        */
} scdat_config;

void scdat_config_default(scdat_config* cfg);

/* Replaces PICCycle::PICCycle(const Parameters&) (PICCycle.h:60-61). */
int scdat_create(const scdat_config* cfg, scdat_ctx** out);
void scdat_destroy(scdat_ctx* ctx);
const char* scdat_last_error(const scdat_ctx* ctx);
/* device name, CU count; returns SCDAT_ERR_DEVICE if no GPU */
int scdat_device_info(char* name, int32_t name_len, int32_t* n_cu);

/* Replaces PICCycle::setMesh(nodes, elements) (PICCycle.h:63-64) and
 * PoissonSolver(VolMeshPtr) (Tools/FieldSolver/include/PoissonSolver.h:26-40):
 * uploads the mesh, builds tet maps, face neighbours, the Poisson operator and
 * the nodal-gradient operator.  dirichlet may be NULL (all zero);
 * face_group may be NULL (all boundary faces in group 0).                   */
int scdat_set_mesh(scdat_ctx* ctx, int64_t n_nodes, const double* xyz, int64_t n_tets,
                   const int32_t* tets, const int32_t* node_kind, const double* dirichlet,
                   const int32_t* face_group);

/* Replaces PICCycle::setBoundaryCondition(BoundaryFace, BoundaryConditionPtr)
 * (PICCycle.h:67).  For SCDAT_BC_PERIODIC, `partner_group` is the opposite
 * group and shift the translation from this group's faces to the partner's.
 * Must be called before scdat_initialize.                                   */
int scdat_set_boundary(scdat_ctx* ctx, int32_t group, int32_t type, int32_t partner_group,
                       const double* shift);
/* Particles that left through ESCAPE groups, per species, since the start. */
int scdat_get_escaped(scdat_ctx* ctx, int32_t species, int64_t* n);
/* Impacts on SCDAT_BC_HOST groups during the last step (the slow path for
 * BoundaryConditionBase::processParticle, BoundaryConditions.h:116-121, that
 * PICCycle::applyRuntimeBoundaryCondition calls, PICCycle.cpp:1221-1293):
 * species, particle id, group, the crossing point (on the face) and the
 * velocity there.  Writes at most max entries and the count of the step's
 * impacts into *n (more than max: call again with a larger buffer).  Only
 * the last step's impacts are kept, so a caller with such groups steps one
 * step at a time.                                                          */
typedef struct {
  int32_t species, id, group, cell;
  double x[3], v[3];
} scdat_impact;
int scdat_get_host_impacts(scdat_ctx* ctx, int64_t max, scdat_impact* out, int64_t* n);

/* Species (ParticleManager::createElectron/createIon semantics,
 * Tools/Particle/include/ParticleManager.h:821-926): charge q (C), mass m (kg),
 * macro weight w, device capacity in particles.                             */
int scdat_add_species(scdat_ctx* ctx, double q, double m, double w, int64_t capacity,
                      int32_t* species_id);

/* Bulk upload / download of one species (replaces iteration over
 * ParticleContainer, ParticleManager.h:55-156).  cell[i] = containing tet;
 * pass cell = NULL to locate on the device by a walk from a seed tet.
 * Removed particles are compacted out on download.                          */
int scdat_set_particles(scdat_ctx* ctx, int32_t species, int64_t n, const double* x,
                        const double* v, const int32_t* cell);
int scdat_get_particle_count(scdat_ctx* ctx, int32_t species, int64_t* n_alive,
                             int64_t* n_slots);
/* Copies n_slots entries in device (cell-sorted) order; removed particles
 * have cell = -1; id[] is the upload index (or creation order) of each
 * particle.  Any output pointer may be NULL.                                */
int scdat_get_particles(scdat_ctx* ctx, int32_t species, double* x, double* v, int32_t* cell,
                        int32_t* id);

/* Initial conductor charge (C) of the floating conductor. */
int scdat_set_conductor_charge(scdat_ctx* ctx, double q);

/* Replaces PICCycle::initialize (PICCycle.cpp:255-309) and computes the
 * initial field: deposit -> Poisson -> E.                                   */
int scdat_initialize(scdat_ctx* ctx);

/* Replaces PICCycle::executeTimeStep (PICCycle.cpp:311-416), n times:
 * push + gather + walk + deposit with the previous E, boundary events,
 * Poisson (with floating conductor), nodal E.  A solve that misses the
 * tolerance keeps the previous potential and field for that step
 * (PICCycle::solvePoissonEquation, PICCycle.cpp:731-767) and the remaining
 * steps still run; the call then returns SCDAT_ERR_SOLVER, naming the count
 * and the first such step (scdat_stats.pcg_failed_solves counts them all).
 * scdat_initialize and scdat_solve follow the same policy.                  */
int scdat_step(scdat_ctx* ctx, int32_t n);

/* Individual stages, for callers that drive the cycle themselves and tests.
 * scdat_push: one push+deposit pass (rho cleared first);
 * scdat_deposit: deposit current particles without moving them;
 * scdat_solve: PoissonSolver::solve (PoissonSolver.cpp:387-401) + nodal E
 * (:474-529) from the current rho.                                          */
int scdat_push(scdat_ctx* ctx);
int scdat_deposit(scdat_ctx* ctx);
int scdat_solve(scdat_ctx* ctx);
int scdat_compute_efield(scdat_ctx* ctx);
int scdat_sort(scdat_ctx* ctx);

/* Inject a nodal field (V/m, 3 per node) / potential (V per node); the
 * latter mirrors PICCycle::setPotential (PICCycle.h:90-105): it recomputes E,
 * and later solves hold the Dirichlet nodes at the values given here
 * (PICCycle.cpp:674-688 takes a DIRICHLET node's value from potential_).     */
/* Replaces PICCycle::setTimeStep (PICCycle.h:86-92): ignored unless dt > 0. */
int scdat_set_time_step(scdat_ctx* ctx, double dt);

/* Poisson stopping controls between steps: relative tolerance and inner
 * iteration cap per solve (scdat_config.pcg_tol / pcg_max_iter); the
 * tolerance and iteration arguments of LinearSystem::conjugateGradient
 * (Tools/Solver/src/LinearSystem.cpp:258-327).  max_iter <= 0 keeps the cap. */
int scdat_set_pcg_controls(scdat_ctx* ctx, double tol, int32_t max_iter);

int scdat_set_efield(scdat_ctx* ctx, const double* E);
int scdat_set_potential(scdat_ctx* ctx, const double* phi);
/* scdat_set_rho: charge (C) per node, held in the fixed point of the
 * species' unit; in a context with no species (a field solve alone) the
 * unit is 2^-60 of the call's largest |rho|, a power of two.               */
int scdat_set_rho(scdat_ctx* ctx, const double* rho);

/* Replaces PICCycle::getChargeDensity / getPotential / getElectricField
 * (PICCycle.h:73-84).  Any pointer may be NULL.  rho is deposited charge (C)
 * per node.                                                                 */
int scdat_get_fields(scdat_ctx* ctx, double* rho, double* phi, double* E);
/* Fixed-point charge: rho_fx[i] * unit = rho[i] (C). */
int scdat_get_rho_fixed(scdat_ctx* ctx, int64_t* rho_fx, double* unit);
/* FP32 tet maps the device uses: 12 floats per tet (3x3 inverse Jacobian
 * row-major, origin xyz) and the per-species integer charge c_s (rho_fx
 * units per particle per unit barycentric weight 2^-24).                    */
int scdat_get_tet_maps(scdat_ctx* ctx, float* maps);
int scdat_get_species_charge_fixed(scdat_ctx* ctx, int32_t species, int64_t* c_s);

typedef struct {
  int64_t completed_steps;
  double simulation_time;
  double total_charge;         /* sum of deposited rho (C)                  */
  double total_energy;         /* sum 1/2 m w v^2 over alive particles (J) */
  int64_t active_particles;
  double surface_deposited_charge; /* absorbed on all groups since start (C) */
  double conductor_charge;     /* floating conductor charge (C)             */
  double floating_potential;   /* V                                          */
  int32_t pcg_iterations;      /* last solve, inner iterations              */
  int32_t pcg_outer;           /* last solve, refinement passes             */
  double pcg_rel_residual;     /* last solve, FP64 ||b - A x|| / ||b||      */
  int32_t pcg_converged;
  int64_t lost_particles;      /* walk limit exceeded                        */
  int64_t particle_pushes;     /* cumulative particles pushed by scdat_step/
                                  scdat_push (alive at the start of a push)  */
  double pcg_initial_rel_residual; /* last solve, warm-start ||r0|| / ||b||  */
  int64_t pcg_iterations_total;    /* inner PCG iterations of all solves     */
  int64_t pcg_failed_solves;       /* solves that missed the tolerance; each kept
                                      the previous potential and field
                                      (PICCycle.cpp:731-767)                 */
} scdat_stats;

/* Replaces PICCycle::getStatistics (PICCycle.h:135-149). */
int scdat_get_stats(scdat_ctx* ctx, scdat_stats* st);
/* The same statistics, bit for bit, with no allocation of any kind: no
 * counterpart in the reference; for callers that cannot allocate (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Its device reductions use a static per-GPU scratch of
 * 8 + 4096 words (at most 4096 scenarios; more is SCDAT_ERR_ARG); calls
 * that overlap on one GPU race on that scratch.                           */
int scdat_get_stats_noalloc(scdat_ctx* ctx, scdat_stats* st);

typedef struct {
  double absorbed_electron_charge_c;
  double absorbed_ion_charge_c;
  double emitted_electron_charge_c;
  double incident_energy_j;
  int64_t absorbed_electron_particles;
  int64_t absorbed_ion_particles;
  int64_t emitted_electron_particles;
} scdat_ledger_entry;

/* Replaces PICCycle::getSurfaceCurrentLedger (PICCycle.h:166-169); one entry
 * per boundary group (n_groups <= SCDAT_MAX_GROUPS).                        */
int scdat_get_ledger(scdat_ctx* ctx, scdat_ledger_entry* out, int32_t n_groups);

/* ---- Background Monte Carlo collisions -----------------------------------
 * Replace MonteCarloCollisionHandler(seed) + addCollisionProcess +
 * processCollisions (Tools/Interactions/Collisions/include/CollisionAlgorithm.h:
 * 176-226) as driven by executeBackgroundMccStep (CollisionPicAdapter.h:16-75,
 * CollisionPicAdapter.cpp:338-434).  Run inside scdat_step after the push and
 * before the Poisson solve over the particles present at the push; products
 * are deposited into the same step's charge.  Per particle one collision with
 * probability 1 - exp(-sum_p nu_p dt), nu_p = n_p sigma_p(E)|v|, process p
 * with probability nu_p / sum nu (the reference's per-process p^2 draw is not
 * reproduced).  Random numbers: Philox4x32-10 keyed by the seed, counter
 * (particle id, step, species, stream).                                     */
#define SCDAT_MCC_ELASTIC 0         /* sigma_max (E_ref/E)^p, isotropic scatter */
#define SCDAT_MCC_INELASTIC 1       /* elastic form, fixed energy loss          */
#define SCDAT_MCC_IONIZATION 2      /* sigma_max ln(E/I)(1-I/E)/E, e- + ion     */
#define SCDAT_MCC_EXCITATION 3      /* sigma_max (1-1/x)^a x^-b, energy loss    */
#define SCDAT_MCC_CHARGE_EXCHANGE 4 /* elastic form, ion removed (neutralised)  */
#define SCDAT_MCC_RECOMBINATION 5   /* sigma_max (E_ref/max(E_ref,E))^p, removed */
#define SCDAT_MCC_COULOMB 6         /* pi lnL (e^2/4pi eps0)^2/(2E)^2, scatter  */
#define SCDAT_MAX_MCC_PROCESSES 16

typedef struct {
  int32_t type;
  int32_t species;                  /* incident species                      */
  int32_t product_electron_species; /* ionisation products                   */
  int32_t product_ion_species;
  double target_density;            /* background density, m^-3              */
  double threshold_ev;
  double sigma_max_m2;
  double energy_loss_ev;
  double reference_energy_ev;       /* elastic form and recombination        */
  double power_index;
  double alpha, beta;               /* excitation                            */
  double coulomb_log;               /* Coulomb                               */
} scdat_mcc_process;

/* The reference's default parameters of one process type
 * (MonteCarloCollisionHandler::initializeDefaultCollisionParameters,
 * CollisionAlgorithm.cpp:590-646); species fields are left for the caller.  */
int scdat_mcc_process_default(int32_t type, scdat_mcc_process* p);
/* Replaces the process set (n = 0 turns collisions off); processes are
 * counted in this order by scdat_get_event_counts.                         */
int scdat_mcc_configure(scdat_ctx* ctx, const scdat_mcc_process* procs, int32_t n);
/* Background density update (reinitialize_configuration with new effective
 * densities, CollisionPicAdapter.cpp:357-363).                              */
int scdat_mcc_set_density(scdat_ctx* ctx, int32_t process, double density);
/* Seed of every random stream (collisions, emission). */
int scdat_set_seed(scdat_ctx* ctx, uint64_t seed);

/* ---- Surface emission ------------------------------------------------------
 * Secondary electrons from electrons absorbed on a CONDUCTOR or ABSORB face
 * and photoelectrons from the face, per boundary group.  Yields: the
 * reference's secondaryYield (Tools/Material/src/SurfaceMaterialModel.cpp:
 * 163-170, delta_max x e^(1-x)/sqrt(max(0.1, cos theta)), x = E/max(1, E_max))
 * or Sternglass delta_max (E/E_max) e^(2 - 2 sqrt(E/E_max)); photo current
 * J0 Y exp(-max(0, V - V_ref)/T_ph) (:199-208), V = floating potential of the
 * last solve on conductor faces, 0 elsewhere.  floor(delta) secondaries plus
 * one with probability frac(delta) (at most 8), J A dt/(|q| w) photo-
 * electrons per step; emitted with a flux Maxwellian of the given
 * temperature (cosine law about the inward normal) just inside the face.
 * Emission from conductor faces adds -q w per electron to the conductor.   */
#define SCDAT_SEE_OFF 0
#define SCDAT_SEE_REFERENCE 1
#define SCDAT_SEE_STERNGLASS 2
typedef struct {
  int32_t see_model;
  int32_t electron_species;     /* species the emitted electrons join       */
  double see_delta_max;
  double see_energy_max_ev;
  double see_temperature_ev;
  double photo_current_density; /* J0 * Y, A/m^2 (0: no photoemission)    */
  double photo_temperature_ev;
  double photo_reference_v;
} scdat_surface_emission;
int scdat_set_surface_emission(scdat_ctx* ctx, int32_t group, const scdat_surface_emission* e);

/* Cumulative event counts (replaces MonteCarloCollisionHandler::getStatistics,
 * CollisionAlgorithm.h:176-226, and the CollisionStepResult counters).
 * overflow: products dropped at a full species; impacts_dropped: impacts
 * beyond the per-step impact buffer (no emission for them).
 * Collisions use the null-collision method (Vahedi & Surendra 1995): a
 * constant majorant nu_max >= sum_p n_p sigma_p(E)|v| over the energies up
 * to the cap (a speed twice the largest of the previous push) gives every
 * particle one candidate draw with probability 1 - exp(-nu_max dt) inside
 * the push, and only candidates evaluate cross sections, accepted with
 * probability (1 - exp(-nu dt)) / (1 - exp(-nu_max dt)); particles faster
 * than the cap are always candidates, accepted with 1 - exp(-nu dt).
 * mcc_candidates: candidates tested; mcc_majorant_exceeded: candidates whose
 * exact probability exceeded the majorant's (0 when the majorant holds).   */
typedef struct {
  int64_t collisions[SCDAT_MAX_MCC_PROCESSES];
  int64_t created;
  int64_t removed;
  int64_t see_emitted;
  int64_t photo_emitted;
  int64_t overflow;
  int64_t impacts_dropped;
  int64_t mcc_candidates;
  int64_t mcc_majorant_exceeded;
} scdat_event_counts;
int scdat_get_event_counts(scdat_ctx* ctx, scdat_event_counts* out);

/* ---- Live-PIC surface current sampler -------------------------------------
 * Replaces PicMccSurfaceCurrentSampler::sample / sampleWithDerivative
 * ("Toolkit/Surface Charging/include/PicMccSurfaceCurrentSampler.h:20-100",
 * PicMccSurfaceCurrentSampler.cpp:489-670): a PIC-MCC window in a plate gap
 * (surface at z = 0 held at surface_potential_v, z = gap at the reference,
 * periodic sides, absorbing z faces); current densities from the charge
 * collected at z = 0.  probe_delta_v != 0 also runs V +/- max(0.1, |d|)
 * windows concurrently (three contexts) and returns dJ/dV.               */
typedef struct {
  double surface_area_m2;
  double gap_distance_m;
  double surface_aspect_ratio;
  double surface_potential_v;
  double plasma_reference_potential_v;
  double electron_density_m3, electron_temperature_ev;
  double ion_density_m3, ion_temperature_ev, ion_mass_amu;
  double neutral_density_m3;  /* collisions (SurfacePicV1 set) when > 1e6   */
  int32_t enable_mcc;
  int32_t nx, ny, z_layers;    /* plate mesh nx x ny x max(4, z_layers) hexes */
  int32_t particles_per_element;
  int32_t window_steps;
  int32_t precision;           /* 0 mixed, 1 FP64                          */
  uint64_t seed;
  double time_step;            /* 0: gap / (40 max thermal speed), [5e-12, 5e-9] s */
} scdat_sampler_config;

typedef struct {
  int32_t valid;
  double sampled_surface_potential_v;
  double electron_collection_current_density_a_per_m2;
  double ion_collection_current_density_a_per_m2;
  double emitted_electron_current_density_a_per_m2;
  double net_collection_current_density_a_per_m2;
  double current_derivative_a_per_m2_per_v;
  int64_t total_collisions;
  double time_step;
  int64_t macro_electrons, macro_ions;
} scdat_current_sample;

/* PicMccSurfaceSamplerConfig defaults (PicMccSurfaceCurrentSampler.h:20-49). */
void scdat_sampler_config_default(scdat_sampler_config* c);
int scdat_sample_surface_current(const scdat_sampler_config* c, double probe_delta_v, scdat_current_sample* out);
/* The window's inputs, for checkers: plate box spec, macro-particle counts
 * and weight (electrons from the per-element count, ions at the same
 * weight), and the velocity direction step (v_z -> -|v_z|, toward the
 * surface) applied after scdat_load_maxwellian(spec, n, seed, 0 | 1, ...). */
void scdat_sampler_plate(const scdat_sampler_config* c, scdat_box_spec* spec);
void scdat_sampler_counts(const scdat_sampler_config* c, int64_t* n_electrons, int64_t* n_ions, double* weight);
void scdat_sampler_direct(int64_t n, double* v);

/* ---- Surface circuit (host, FP64) -----------------------------------------
 * Replaces SurfaceCircuitCoupling (Tools/Coupling/include/SurfaceCircuitCoupling.h:
 * 13-45,324-361): nodes with capacitance, shunt conductance or a fixed
 * potential, branches with conductance and bias, advanced by backward Euler
 * with a linearised plasma current per node (advanceImplicit,
 * SurfaceCircuitCoupling.cpp:897-1127).  Independent of any device context. */
typedef struct scdat_circuit scdat_circuit;
typedef struct {
  int32_t row, column;   /* node indices of an extra matrix entry (A/V)     */
  double coefficient;
} scdat_circuit_entry;
int scdat_circuit_create(scdat_circuit** out);
void scdat_circuit_destroy(scdat_circuit* c);
const char* scdat_circuit_last_error(const scdat_circuit* c);
/* addNode (SurfaceCircuitCoupling.h:329) */
int scdat_circuit_add_node(scdat_circuit* c, double potential_v, double capacitance_f,
                           double shunt_conductance_s, int32_t fixed, double fixed_value_v, int32_t* index);
/* addBranch (:330) */
int scdat_circuit_add_branch(scdat_circuit* c, int32_t from, int32_t to, double conductance_s, double bias_v,
                             int32_t* index);
/* setNodePotential / setNodeCapacitance / setNodeShuntConductance /
 * setNodeFixedPotential (:332-335) in one call */
int scdat_circuit_set_node(scdat_circuit* c, int32_t node, double potential_v, double capacitance_f,
                           double shunt_conductance_s, int32_t fixed, double fixed_value_v);
/* setBranchConductance (:336) */
int scdat_circuit_set_branch_conductance(scdat_circuit* c, int32_t branch, double conductance_s);
int scdat_circuit_get_potentials(const scdat_circuit* c, double* v);
/* getNodes().size() / getBranches().size() (:338-346) */
int scdat_circuit_size(const scdat_circuit* c, int32_t* n_nodes, int32_t* n_branches);

/* ---- Batched scenarios ----------------------------------------------------
 * One context holds n independent charging scenarios whose meshes are
 * concatenated into the context's mesh: scenario s owns the nodes
 * [node_start[s], node_start[s+1]) (each start a multiple of
 * SCDAT_SEGMENT_ALIGN; nodes that no tetrahedron uses are padding, held at
 * 0 V) and the tetrahedra on them.  Every step advances all scenarios with one
 * launch per kernel: one push over the particles of all scenarios, one PCG
 * whose scalars, convergence test and floating conductor are per scenario.
 * Scenario s has its own time step, weight multiplier (every species' weight
 * times weight_scale[s]), random key (seed[s]) and surface emission; its
 * results equal those of the same scenario run alone in its own context.
 * Any array may be NULL (cfg.dt, 1, the context seed).  Call after
 * scdat_set_mesh and before scdat_set_particles; particle ids then count
 * within each scenario.  Background collisions (each scenario with its own
 * target densities, scdat_set_scenario_mcc_density) and photoemission (per
 * scenario through scdat_set_scenario_emission) run per scenario; conductor
 * patches and field emission need a single scenario.  This batched form has no counterpart in
 * the reference, which runs one PICCycle per case
 * (PicMccSurfaceCurrentSampler.cpp:519-586 builds one per window).        */
#define SCDAT_SEGMENT_ALIGN 512
int scdat_set_scenarios(scdat_ctx* ctx, int32_t n, const int64_t* node_start, const double* weight_scale,
                        const double* time_step, const uint64_t* seed);
/* Surface emission of one boundary group of one scenario
 * (scdat_set_surface_emission sets the group in every scenario). */
int scdat_set_scenario_emission(scdat_ctx* ctx, int32_t scenario, int32_t group, const scdat_surface_emission* e);
/* Target density of background-collision process `process` (an index into
 * the scdat_mcc_configure list) in one scenario; every scenario starts at the
 * process's own density, and scdat_mcc_configure resets them.  The
 * scenario's majorant, collision probabilities and random key are its own,
 * so its collisions equal those of the scenario run alone with that
 * density (executeBackgroundMccStep per case, CollisionPicAdapter.cpp:338-434). */
int scdat_set_scenario_mcc_density(scdat_ctx* ctx, int32_t scenario, int32_t process, double density);
typedef struct {
  int64_t completed_steps;
  double simulation_time, time_step, weight_scale;
  double floating_potential, conductor_charge;
  int32_t pcg_iterations, pcg_converged;   /* of the last solve */
  double pcg_rel_residual;
  int64_t first_node, n_nodes;
} scdat_scenario_stats;
int scdat_get_scenario_stats(scdat_ctx* ctx, int32_t scenario, scdat_scenario_stats* out);
/* getSurfaceCurrentLedger of one scenario (scdat_get_ledger sums them). */
int scdat_get_scenario_ledger(scdat_ctx* ctx, int32_t scenario, scdat_ledger_entry* out, int32_t n_groups);

/* ---- Field-emission cathode (vacuum-arc coupling) ----------------------------
 * FieldEmissionBoundaryCondition (Tools/Boundary/src/FieldEmissionBoundary
 * Condition.cpp:23-178), as the vacuum-arc toolkit sets it on the cathode
 * (SurfaceDischargeArcAlgorithm.cpp:441-458): an ion hitting the group is
 * absorbed and adds J A dt / e expected field electrons and (when enabled)
 * yield * w_ion / w_secondary expected secondaries to the group's running
 * accumulators; their integer parts, at most max_*_per_event, are emitted
 * along the face's inward normal at the given speeds, 1e-10 m apart off the
 * impact point, into electron_species (which carries their weight).  Within
 * a step, impacts are taken in particle-id order (the reference's is its
 * container order).  Non-positive settings take the constructor's
 * defaults.                                                                 */
typedef struct {
  int32_t electron_species;
  double emission_current_density_a_per_m2;
  double emission_area_m2;
  double emitted_electron_speed_m_per_s;
  double emitted_electron_weight;
  int32_t max_emitted_particles_per_event;
  int32_t enable_secondary_electron_emission;
  double secondary_electron_yield_per_ion;
  double secondary_electron_speed_m_per_s;
  double secondary_electron_weight;
  int32_t max_secondary_particles_per_event;
} scdat_field_emission;
/* FieldEmissionBoundaryCondition::Parameters defaults
 * (FieldEmissionBoundaryCondition.h:17-31). */
void scdat_field_emission_default(scdat_field_emission* p);
int scdat_set_field_emission(scdat_ctx* ctx, int32_t group, const scdat_field_emission* e);
/* setEmissionCurrentDensity (FieldEmissionBoundaryCondition.cpp:57-61) */
int scdat_set_emission_current_density(scdat_ctx* ctx, int32_t group, double j_a_per_m2);
/* Appends particles to a species between steps (ParticleManager::createIon /
 * createElectron + PICCycle::addParticleLocation,
 * SurfaceDischargeArcAlgorithm.cpp:527-557); cell NULL: located on the host;
 * ids continue the species' count.                                         */
int scdat_append_particles(scdat_ctx* ctx, int32_t species, int64_t n, const double* x, const double* v,
                           const int32_t* cell);

/* ---- Mesh import and result export (host, no device) -----------------------
 * Readers for the formats of GmshLoader / VTKLoader / STLLoader
 * (Tools/Mesh/src/MeshParsing.cpp:651-1091): Gmsh ASCII 2.x and 4.x
 * (triangles and tetrahedra with their physical tags), VTK legacy ASCII
 * (CELL_TYPES 5 and 10), STL ASCII (triangles, vertices merged on a 1e-12
 * grid).  SCDAT_MESH_AUTO picks the reader from the extension (msh, vtk,
 * stl).  Errors: SCDAT_ERR_ARG with the message in scdat_io_last_error.   */
#define SCDAT_MESH_AUTO 0
#define SCDAT_MESH_GMSH 1
#define SCDAT_MESH_VTK 2
#define SCDAT_MESH_STL 3
typedef struct scdat_mesh_data scdat_mesh_data;
const char* scdat_io_last_error(void);
int scdat_mesh_read(const char* path, int32_t format, scdat_mesh_data** out);
void scdat_mesh_free(scdat_mesh_data* m);
int scdat_mesh_sizes(const scdat_mesh_data* m, int64_t* n_nodes, int64_t* n_tets, int64_t* n_tris);
/* xyz[3 n], tets[4 n_tets], tet_tag[n_tets], tris[3 n_tris], tri_tag[n_tris]; any may be NULL */
int scdat_mesh_arrays(const scdat_mesh_data* m, double* xyz, int32_t* tets, int32_t* tet_tag, int32_t* tris,
                      int32_t* tri_tag);
/* face_group[4 n_tets] for scdat_set_mesh: -1 on interior faces; on a
 * boundary face, groups[k] when the triangle with the same nodes has
 * physical tag tags[k], else default_group.                               */
int scdat_mesh_face_groups(const scdat_mesh_data* m, int32_t n_map, const int32_t* tags, const int32_t* groups,
                           int32_t default_group, int32_t* face_group, int64_t* n_matched);
/* Tetrahedral mesh with point fields (1 or 3 components) as a legacy ASCII
 * VTK unstructured grid.                                                   */
int scdat_write_vtk(const char* path, int64_t n_nodes, const double* xyz, int64_t n_tets, const int32_t* tets,
                    int32_t n_fields, const char* const* names, const int32_t* components, const double* const* data);
/* The columnar output contract of ResultExporter::exportDataSet
 * (Tools/Output/src/ResultExporter.cpp:176-230): csv_path, and beside it
 * <stem>.vtk (poly-data points along the axis) and <stem>.h5 (the text
 * archive), each with a .metadata.json sidecar; series in name order.      */
int scdat_export_columns(const char* csv_path, const char* module, const char* axis_name, int64_t n_rows,
                         const double* axis, int32_t n_series, const char* const* names, const double* const* series);
/* Layer profile of main_PICcore.cpp:230-310 (collectLayerProfile): nodes
 * grouped by their coordinate along `axis` (0..2, rounded to 1e-9 m); per
 * layer the mean potential and mean field component along the axis, and
 * from the particles (binned by llround(clamp(x, 0, length) / (length /
 * n_cells))) the electron and ion densities (weight / (area length /
 * n_cells)) and the electron temperature (2/3 of the weighted mean kinetic
 * energy, eV).  scdat_layer_count gives the number of layers.              */
int scdat_layer_count(scdat_ctx* ctx, int32_t axis, int64_t* n_layers);
int scdat_layer_profile(scdat_ctx* ctx, int32_t axis, double length, int32_t n_cells, double area, double* position,
                        double* potential, double* field, double* electron_density, double* ion_density,
                        double* electron_temperature_ev);

/* ---- Internal node order -----------------------------------------------------
 * SCDAT_ORDER_GIVEN (default) keeps the caller's node order;
 * SCDAT_ORDER_MORTON renumbers the nodes internally in Morton order of their
 * coordinates, so that consecutive 512-node row blocks are compact boxes;
 * SCDAT_ORDER_TILED, for nodes on a lattice (uniform spacing per axis, such
 * as scdat_box_mesh_build's, with or without the conductor box), makes every
 * 512-node row block an 8 x 8 x 8 tile of lattice points (the index-0 layers
 * fill the tiles' holes and follow the tiles); on other meshes it is Morton.
 * The two-level preconditioner's aggregates are the row blocks, so it needs
 * MORTON or TILED; with TILED and the mixed-precision solve its coarse step
 * runs inside the iteration kernel.  On a structured box mesh numbered
 * (i, j, k) the caller's order is faster for the Jacobi iteration than
 * Morton (neighbour gathers of consecutive rows coalesce: 15 us against
 * 20 us per iteration at 64^3).  Every entry point takes and returns node
 * arrays in the caller's order.  Batched scenarios always use the caller's
 * order.  Call before scdat_set_mesh.                                      */
#define SCDAT_ORDER_MORTON 0
#define SCDAT_ORDER_GIVEN 1
#define SCDAT_ORDER_TILED 2
int scdat_set_node_order(scdat_ctx* ctx, int32_t order);

/* ---- Poisson preconditioner ------------------------------------------------
 * SCDAT_PRECOND_JACOBI (default): diagonal (MultiBoundaryElectricFieldSolver
 * .cpp:1432-1513), one fused kernel per iteration.  SCDAT_PRECOND_TWO_LEVEL:
 * M^-1 = D^-1 + P Ac^-1 P^T, P the free nodes of each 512-node block of the
 * internal order (use SCDAT_ORDER_TILED or SCDAT_ORDER_MORTON, so that
 * blocks are compact), Ac = P^T A P inverted once at scdat_set_mesh.  It
 * needs fewer iterations (64^3 box: about 0.4x).  In mixed precision with
 * SCDAT_ORDER_TILED each iteration is still one kernel (every block applies
 * the rows of Ac^-1 of the aggregates it touches); otherwise a coarse
 * kernel runs after every iteration.  Meshes under 4096 free nodes or over
 * 2048 blocks, and batched scenarios, always use Jacobi.  No counterpart in
 * the reference.                                                           */
#define SCDAT_PRECOND_JACOBI 0
#define SCDAT_PRECOND_TWO_LEVEL 1
int scdat_set_preconditioner(scdat_ctx* ctx, int32_t kind);
typedef struct {
  int64_t rows, nonzeros;   /* operator rows (all nodes), stored non-zeros     */
  int32_t ell_width;        /* ELL slots per row                              */
  int32_t preconditioner;   /* in use for the next solve                       */
  int32_t coarse_unknowns;  /* two-level: aggregates (row blocks); 0          */
  int32_t row_patterns;     /* FP32 inner operator as row patterns (0: ELL)   */
  int32_t pcg_block;        /* rows (= threads) of a PCG row block            */
  int32_t coarse_in_kernel; /* two-level coarse step inside the iteration kernel */
  int32_t tile_staged;      /* iteration stages tiles in LDS (box patterns)   */
  int32_t box_patterns;     /* distinct rows of the tile-staged iteration     */
} scdat_solver_info;
int scdat_get_solver_info(scdat_ctx* ctx, scdat_solver_info* out);

/* ---- Launch shapes -----------------------------------------------------------
 * SCDAT_SHAPE_PUSH_CHUNKS: particles per thread of the push kernel, whose
 * deposit tile is chunks x 256 consecutive (cell-sorted) particles: 1, 2, 4
 * (default) or 8.  scdat_autotune (after scdat_initialize) times the push at
 * every shape on cell-sorted scratch copies of all species, every push of
 * one sort cycle of each species, in the current field, and
 * keeps the shape with the least summed average; the simulation state is
 * untouched.  Every shape gives the same bits.  No counterpart in the
 * reference.                                                               */
#define SCDAT_SHAPE_PUSH_CHUNKS 0
/* SCDAT_SHAPE_ROW_PATTERNS: 1 (default) lets the FP32 inner iteration read
 * the operator as per-row pattern ids into a table of distinct rows (column
 * offsets and values bit-identical), when ids and table take at most half
 * the ELL's bytes (structured meshes in their own node order); 0 keeps the
 * ELL slots; 2 uses 4-byte pattern ids even where 2 bytes would do (the
 * form a mesh with more than 65536 distinct rows gets).  Same bits in every
 * form.                                                                    */
#define SCDAT_SHAPE_ROW_PATTERNS 1
/* SCDAT_SHAPE_PCG_BLOCK: rows (= threads) of a PCG row block.  0 (default):
 * 512, or 256 for a single scenario under 131072 rows with the Jacobi
 * preconditioner; 256 or 512 fix it (batched scenarios and the two-level
 * preconditioner always use 512).  The block partition fixes the order of
 * the solve's reductions: a batched case equals its stand-alone run bit for
 * bit when both use 512.                                                   */
#define SCDAT_SHAPE_PCG_BLOCK 2
/* SCDAT_SHAPE_TILE_STAGING: 1 (default) lets the mixed-precision iteration
 * in SCDAT_ORDER_TILED stage each 512-row tile's neighbour values in LDS (a
 * 10^3 box per tile, read by direction codes) instead of gathering W
 * neighbours per row, where every tile fills its row block and every
 * operator entry joins lattice neighbours; 0 keeps the row patterns.  Same
 * bits either way.                                                        */
#define SCDAT_SHAPE_TILE_STAGING 3
/* SCDAT_SHAPE_CELL_ORDER: order of the cells along the sorted particle
 * arrays.  0 (default): tet index order, in which a block's particles read
 * consecutive tet records; 1: the cells' centroids along a Morton curve
 * (inside each scenario), so that a stretch of the push covers a compact
 * region while its particles drift between sorts.  Equal cells stay
 * contiguous either way; fields, particles and ledger counts are the same
 * bit for bit (the ledger's FP64 energy sums are atomics in block order:
 * equal to rounding).                                                     */
#define SCDAT_SHAPE_CELL_ORDER 4
/* SCDAT_SHAPE_STEP_GRAPH: 1 records the launches of scdat_step around the
 * Poisson iteration (push, collisions, emission, right-hand side; floating
 * potential, phi and E, sorts) as HIP graphs, one per launch configuration
 * met in the call, and replays them; slot counts and the step number are
 * then kept on the device.  Same kernels, arguments and order as the eager
 * loop, so the same results.  0 (default): eager launches.  Contexts with a
 * surface circuit or a field-emission cathode (host-side stages) always
 * step eagerly.  No counterpart in the reference.                          */
#define SCDAT_SHAPE_STEP_GRAPH 5
/* SCDAT_SHAPE_LATTICE_PUSH: 1 lets the FP32 push locate the end tet of a
 * sub-step that left its start tet on the grid, when the mesh is a
 * rectilinear grid of Kuhn tetrahedra (the box mesher's), the end point is
 * at least 1e-3 (barycentric) inside a tet and every hexahedron between
 * start and end is whole with no boundary face; every other sub-step walks
 * tet faces.  The end tet is the walk's, so results are the same bits.  0
 * (default): every sub-step walks; scdat_autotune picks whichever pushes
 * faster.  Reads back 1 only when the push uses it (FP32, lattice mesh).   */
#define SCDAT_SHAPE_LATTICE_PUSH 6
/* SCDAT_SHAPE_DEPOSIT_BUCKETS: 1 (default) sends each push block's reduced
 * charge table to per-512-row buckets that one block per row range sums
 * after the pushes (one atomic per block and range, plain stores for the
 * entries); 0: each block adds its table to the charge with one 64-bit
 * global atomic per node.  Integer sums: the same charge either way.      */
#define SCDAT_SHAPE_DEPOSIT_BUCKETS 7
int scdat_set_launch_shape(scdat_ctx* ctx, int32_t what, int32_t value);
int scdat_get_launch_shape(scdat_ctx* ctx, int32_t what, int32_t* value);
int scdat_autotune(scdat_ctx* ctx);

/* ---- Discrete adjoint of the charging cycle -------------------------------
 * Objective: the floating conductor's potential after `steps` collisionless
 * steps from the initialized state (before any scdat_step).  Forward model
 * per step: B = 0 push (one sub-step), FP64 charge deposition, and the
 * conductor charge
 *   Q_{n+1} = Q_n + sum over conductor hits (q w + [electrons] e w delta)
 *             + dt (J_ph(V_n) A_sunlit - G V_n),
 * delta = see_delta_max x face_yield[face] x the Sternglass shape
 * (E/E_max) exp(2 - 2 sqrt(E/E_max)) of the impact energy, J_ph =
 * J0 exp(-max(0, V - V_ref) / T_ph) (SurfaceMaterialModel.cpp:199-208) on
 * the sunlit faces, G the conductance from the conductor to the 0 V walls:
 * secondary and photoelectrons leave as surface current, so the yields
 * enter continuously.  The adjoint runs the transposed steps backwards
 * (Poisson backward = one PCG solve with the same SPD operator, push
 * backward from states recomputed per checkpoint segment) and returns
 * dV/d(delta_max, J0, G, face_yield[0..5]).  The context's fields and
 * particles are left as they were.  No counterpart in the reference, which
 * integrates surface charging without gradients.                         */
typedef struct {
  double see_delta_max;          /* Sternglass delta_max                          */
  double see_energy_max_ev;      /* Sternglass E_max                              */
  double face_yield[6];          /* delta_max multiplier of conductor face k (group 6 + k) */
  double photo_current_density;  /* J0 (A/m^2)                                    */
  double photo_temperature_ev, photo_reference_v;
  int32_t sunlit_faces;          /* bit k: conductor face k is sunlit             */
  double surface_conductance_s;  /* conductor -> outer walls                      */
} scdat_charging_params;
#define SCDAT_GRAD_DELTA_MAX 0
#define SCDAT_GRAD_PHOTO_J 1
#define SCDAT_GRAD_CONDUCTANCE 2
#define SCDAT_GRAD_FACE_YIELD 3 /* .. 8 */
#define SCDAT_N_GRAD 9
/* delta_max 2 at 400 eV, yields 1, 20 uA/m^2 at 2 eV on the XMax face, G 0 */
void scdat_charging_params_default(scdat_charging_params* p);
int scdat_charging_forward(scdat_ctx* ctx, const scdat_charging_params* p, int32_t steps, double* v_final);
/* The same forward with its absorption events held fixed, for finite
 * differences of the map the adjoint differentiates (each particle is
 * absorbed at a fixed step through a fixed group; P1 gather and deposit are
 * continuous across tet faces, so absorption is the map's only jump).
 * SCDAT_EVENTS_RECORD runs the forward and keeps, per particle, the step and
 * group of its absorption; SCDAT_EVENTS_REPLAY runs it with those events
 * imposed: a particle is absorbed at its recorded step and group whatever
 * its walk, and a particle the record kept whose walk leaves the mesh is
 * relocated by a walk from the cell the recorded run reached to its
 * unconstrained position (a residual overshoot beyond a boundary face is
 * projected back onto the face's plane).  Both rules are continuous in the
 * parameters and the identity where the events agree, so replaying at the
 * recorded parameters reproduces the recorded run.  A record keeps 4 bytes
 * per particle and step of device memory.                                 */
#define SCDAT_EVENTS_FREE 0
#define SCDAT_EVENTS_RECORD 1
#define SCDAT_EVENTS_REPLAY 2
int scdat_charging_forward_events(scdat_ctx* ctx, const scdat_charging_params* p, int32_t steps, int32_t events,
                                  double* v_final);
/* grad[SCDAT_N_GRAD]; states are kept every checkpoint_interval steps. */
int scdat_charging_adjoint(scdat_ctx* ctx, const scdat_charging_params* p, int32_t steps, int32_t checkpoint_interval,
                           double* v_final, double* grad);

/* ---- Conductor patches on a surface circuit ------------------------------
 * The conductor's nodes split into patches (node_patch[n], -1 off the
 * conductor; NULL: the six faces, boundary groups 6..11 -> patches 0..5,
 * and the body, patch 6 = nodes on more than one face, n_patches = 7).
 * Each patch is held at the potential of circuit node patch_circuit_node[k];
 * after every push the charge collected by its groups (group_patch[g], NULL:
 * group 6 + k -> patch k) over dt is that node's plasma current and the
 * circuit takes one backward-Euler step (SurfaceCircuitCoupling::
 * advanceImplicit, SurfaceCircuitCoupling.cpp:897-1127) before the field:
 * phi = phi_0 + sum_k V_k phi_k, phi_k the unit solution of patch k (solved
 * at scdat_initialize).  This replaces the Gauss-law floating conductor; the
 * circuit is caller-owned and must outlive the context's use.  n_patches = 0
 * returns to the floating conductor.                                        */
#define SCDAT_MAX_PATCHES 16
int scdat_set_conductor_circuit(scdat_ctx* ctx, int32_t n_patches, const int32_t* node_patch, scdat_circuit* circuit,
                                const int32_t* patch_circuit_node, const int32_t* group_patch);
/* dI/dV per patch (A/V) for the circuit's linearised plasma current. */
int scdat_set_patch_didv(scdat_ctx* ctx, const double* didv);
int scdat_get_patch_potentials(scdat_ctx* ctx, double* v);
/* Unit-potential solution of one patch (FP64, n_nodes), after initialize. */
int scdat_get_patch_basis(scdat_ctx* ctx, int32_t patch, double* phi);
/* advanceImplicit(dt, linearization, max_delta) (:348-353).  Per-node arrays
 * of length n_nodes, any may be NULL (zeros); max_delta_v = INFINITY for no
 * step limit.  Outputs (any NULL): new potentials, linearised plasma
 * currents I + dI/dV (V_new - V_old), branch currents G (V_from - V_to - bias),
 * largest |dV|.                                                              */
int scdat_circuit_advance(scdat_circuit* c, double dt, const double* plasma_current, const double* plasma_didv,
                          const double* additional_rhs, const double* additional_diag,
                          const scdat_circuit_entry* entries, int32_t n_entries, double max_delta_v,
                          double* node_potentials, double* plasma_currents, double* branch_currents,
                          double* max_delta_out);

/* Per-stage device time of the last scdat_step call, milliseconds:
 * [push, rhs, pcg, field, sort].                                            */
int scdat_get_stage_times(scdat_ctx* ctx, double* ms5);

/* Average device time (ms) of the fused push kernel over the last
 * scdat_step call, measured with HIP events on the context's stream, and the
 * number of launches it averages.                                           */
int scdat_get_push_kernel_time(scdat_ctx* ctx, double* avg_ms, int64_t* launches);

/* Per-kernel timing from the kernels' own dispatch timestamps (HIP events
 * attached to the launches on the context's stream; off by default, enabling
 * resets the sums).  period 0: off; n > 0: every push launch and every n-th
 * PCG iteration launch (counted across solves, skipping the first five of
 * each batch) is timed; a timed launch leaves about 5 us gaps on both sides,
 * so a period of 32-64 keeps the measured loop close to the untimed one.
 * which: 0 push, 1 PCG iteration (fused SpMV + vector update); averages over
 * timed launches that did work since timing was enabled.                   */
#define SCDAT_KERNEL_PUSH 0
#define SCDAT_KERNEL_PCG 1
int scdat_set_kernel_timing(scdat_ctx* ctx, int32_t period);
int scdat_get_kernel_time(scdat_ctx* ctx, int32_t which, double* avg_ms, int64_t* launches);

/* Synchronise the context's stream. */
int scdat_sync(scdat_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif /* SCDAT_AMD_H */
