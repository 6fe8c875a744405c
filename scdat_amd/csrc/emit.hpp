// Launch interface of the background-collision and surface-emission kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "device_common.hpp"

namespace scdat {

constexpr int kMaxProc = 16;
constexpr int kMaxEmit = 8;  // secondaries per impact
// counter slots: [0, kMaxProc) collisions per process, then:
enum CountSlot {
  C_CREATED = kMaxProc, C_REMOVED, C_SEE, C_PHOTO, C_OVERFLOW, C_EVENT_DROP,
  C_CAND,   // collision candidates tested (null-collision pre-selection)
  C_MAJOR,  // candidates whose exact probability exceeded the majorant's
  C_COUNT
};

// Null-collision majorant of one species (Vahedi & Surendra, Comput. Phys.
// Commun. 87 (1995) 179): nu_max bounds sum_p n_p sigma_p(E) sqrt(2E/m) for
// E up to the cap energy of a speed twice the largest speed of the previous
// push (v^2 cap = kMccCapFactor x max v^2).  A particle is a candidate with
// probability P_null = 1 - exp(-nu_max dt), or always when it is faster than
// the cap; a candidate collides with probability (1 - exp(-nu dt)) / P_c
// (P_c = P_null, or 1 for a forced candidate), so every particle collides
// with probability exactly 1 - exp(-nu dt) while only candidates evaluate
// cross sections.
constexpr int kMccMajorantPoints = 1024;   // log-spaced energies over [1e-8, 1] x cap
constexpr double kMccMajorantMargin = 1.05;
constexpr double kMccCapFactor = 4.0;
// ctl[species]: P_null, v^2 cap, nu_max, unused
constexpr int kMccCtl = 4;
constexpr uint32_t kMccForced = 0x80000000u;  // candidate entry flag: faster than the cap

enum MccType { MCC_ELASTIC = 0, MCC_INELASTIC, MCC_IONIZATION, MCC_EXCITATION, MCC_CHARGE_EXCHANGE,
               MCC_RECOMBINATION, MCC_COULOMB };

struct MccProcDev {
  int type, e_species, i_species, slot;
  double n, thr, smax, loss, eref, pidx, alpha, beta, clog;
};

// Append target: one species' SoA.  New slots are base + atomicAdd(append).
template <typename T>
struct SpeciesOut {
  T *x, *y, *z, *vx, *vy, *vz;
  int32_t* cell;
  int32_t* id;
  int64_t base, cap;
  const int64_t* base_dev;  // captured steps: the slot count on the device (null: base)
  unsigned long long* append;
  double m;
  int64_t cs;  // integer charge per unit weight (deposit)
};

// Impact of an electron on a face with secondary emission, recorded by the push.
struct SeeEvent {
  double x[3], v[3];
  int32_t id, cell, group_face, species;  // group_face = group | face << 8
};

template <typename T>
struct MccArgs {
  T *x, *y, *z, *vx, *vy, *vz;
  int32_t* cell;
  const int32_t* id;
  int64_t n;
  // candidates flagged by the push: slot | kMccForced, their count, and the
  // species' majorant controls
  const uint32_t* cand;
  const unsigned int* n_cand;
  const double* ctl;
  int species, np;
  double mass, dt;
  uint32_t step, key0, key1;
  MccProcDev proc[kMaxProc];
  SpeciesOut<T> out[kMaxSpecies];
  const TetMap<T>* tets;
  unsigned long long* rho;
  unsigned long long* counts;
  // batched scenarios (null: one scenario): segment of each tet, target
  // density of every process slot per segment [segment][kMaxProc], FP64 time
  // step and random key (2 words) per segment; ctl then holds kMccCtl
  // doubles per segment
  const int32_t* tet_seg;
  const double* seg_n;
  const double* seg_dt;
  const uint32_t* seg_key;
  const uint32_t* step_dev;  // captured steps: the step number on the device (null: step)
};

struct SurfEmitDev {
  int see_model, e_species;
  double dmax, emax, t_see, photo_j, t_ph, v_ref;
};

template <typename T>
struct EmitArgs {
  SurfEmitDev g[kMaxGroups];
  SpeciesOut<T> out[kMaxSpecies];
  double q_abs_w[kMaxSpecies];  // |q| w of each species
  double mass[kMaxSpecies];
  const TetMap<T>* tets;
  const double* xyz;  // node coordinates (FP64)
  const int32_t* tet4;
  unsigned long long* rho;
  unsigned long long* counts;
  unsigned long long* emit_count;  // [kMaxGroups][kMaxSpecies]
  uint32_t step, key0, key1;
  double dt;
  // impacts
  const SeeEvent* events;
  const unsigned long long* n_events;
  int64_t event_cap;
  // photoemission of one group: faces [f0, f1) of the face table
  const int32_t* face_tet;
  const int32_t* face_k;
  const double* face_cum;  // cumulative area within the group
  int64_t f0, f1;
  int group;
  const double* vfloat;  // floating potential of the last solve
  int is_conductor;
  // batched scenarios (null: one scenario): segment of each tet, emission
  // settings [segment][kMaxGroups], random key (2 words) per segment;
  // emit_count is then [segment][kMaxGroups][kMaxSpecies]
  const int32_t* tet_seg;
  const SurfEmitDev* gseg;
  const uint32_t* seg_key;
  uint32_t electron_mask;  // species whose impacts are secondary-emission events
  // photoemission per scenario (grid y = scenario): faces [seg_face_off[s],
  // seg_face_off[s + 1]) of this group, cumulative areas restarting at each
  // scenario; FP64 time step and weight multiplier per scenario; the
  // floating potential of scenario s at vfloat + s * vfloat_stride doubles
  const int64_t* seg_face_off;
  const double* seg_dt;
  const double* seg_wscale;
  int64_t vfloat_stride;
  const uint32_t* step_dev;  // captured steps: the step number on the device (null: step)
};

// Majorant of one species from the largest |v|^2 of the previous push; also
// clears that maximum and the candidate count for the coming push.
// With scenarios (seg_n non-null) one block per scenario: densities
// seg_n[s][slot], time step seg_dt[s], maximum v2max[s], controls
// ctl + s kMccCtl.
struct MccMajorantArgs {
  MccProcDev proc[kMaxProc];
  int np;
  double mass, dt;
  unsigned long long* v2max;  // FP64 bits of max |v|^2 (non-negative: ordered as integers)
  unsigned int* n_cand;
  double* ctl;                // kMccCtl doubles
  const double* seg_n;
  const double* seg_dt;
  int n_seg;
};
void launch_mcc_majorant(const MccMajorantArgs& a, hipStream_t s);
// max |v|^2 over the alive particles of one species into *v2max (atomic
// max); with tet_seg non-null into v2max[segment of the particle's cell]
template <typename T>
void launch_v2max(const T* vx, const T* vy, const T* vz, const int32_t* cell, int64_t n, unsigned long long* v2max,
                  const int32_t* tet_seg, hipStream_t s);
template <typename T>
void launch_mcc(const MccArgs<T>& a, hipStream_t s);
template <typename T>
void launch_see(const EmitArgs<T>& a, int64_t max_events, hipStream_t s);
// max_count: upper bound of the group's emissions this step (grid size);
// n_seg > 1: one grid row per scenario (a.seg_face_off set)
template <typename T>
void launch_photo(const EmitArgs<T>& a, int64_t max_count, int n_seg, hipStream_t s);
void launch_fill_u64(unsigned long long* p, int64_t n, hipStream_t s);
// Captured steps: every species' slot count n_dev[k] += append[k], capped at
// cap[k] (the appends past the capacity were counted as overflow), for
// k < n_species; then the append counters are cleared.
struct CountArgs {
  int64_t cap[kMaxSpecies];
  int n_species;
};
void launch_commit_counts(int64_t* n_dev, unsigned long long* append, const CountArgs& a, hipStream_t s);
// Captured steps: the device step number advances by one.
void launch_step_advance(uint32_t* step_dev, hipStream_t s);
// n particles (x, v: 3n FP64, cell, id) appended to o and deposited
template <typename T>
void launch_inject(const SpeciesOut<T>& o, int64_t n, const double* x, const double* v, const int32_t* cell,
                   const int32_t* id, const TetMap<T>* tets, unsigned long long* rho, unsigned long long* counts,
                   hipStream_t s);

}  // namespace scdat
