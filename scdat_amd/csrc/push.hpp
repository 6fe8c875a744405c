// Launch interface of the fused particle kernel.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "device_common.hpp"
#include "emit.hpp"

namespace scdat {

template <typename T>
struct PushArgs {
  // mesh
  const TetMap<T>* tets;
  const WalkMap<T>* wmap;  // FP32: walk records (map + face neighbours) per tet
  const int4* nb;       // face neighbours (face k opposite vertex k), <0 boundary
  const int4* groups;   // boundary group per face
  const int4* partner;  // periodic partner tet per face
  const Vec4<T>* E;     // nodal E
  unsigned long long* rho;  // int64 fixed-point nodal charge
  uint32_t periodic_mask;  // bit g set: group g periodic
  uint32_t reflect_mask;   // bit g set: group g specular
  const T* gshift;         // [kMaxGroups][3] periodic translation (device)
  // species SoA
  T *x, *y, *z, *vx, *vy, *vz;
  int32_t* cell;
  int64_t n;
  // captured steps: slot count and step number on the device (null: n and
  // mcc_step); n then only sizes the grid
  const int64_t* n_dev;
  const uint32_t* step_dev;
  int64_t* n_pushed;  // captured steps: receives the count this push covered
  int species;
  T qm;
  int64_t cs;           // integer charge per particle per unit weight 2^-24
  double mass, weight, energy_weight;
  // push controls
  T dt;
  int adaptive;
  T floor_ratio;
  int max_sub;
  int max_walk;
  T tol;
  T B[3];
  T bmag;
  int has_b;
  // diagnostics
  unsigned long long* ledger_count;  // [kMaxGroups][kMaxSpecies]
  double* ledger_energy;             // [kMaxGroups][kMaxSpecies]
  unsigned long long* lost;
  unsigned long long* pushes;        // cumulative particles pushed
  double* ke_partial;                // per block
  int32_t* alive_partial;            // per block
  // secondary emission: impacts on groups in see_mask are recorded
  const int32_t* id;
  uint32_t see_mask;
  SeeEvent* events;
  unsigned long long* n_events;
  int64_t event_cap;
  // batched scenarios: segment of each tet, time step and weight multiplier
  // per segment; ledger slots [segment][kMaxGroups][kMaxSpecies] (null: one
  // scenario, a.dt and a.weight)
  const int32_t* tet_seg;
  const T* seg_dt;
  const double* seg_wscale;
  int chunks;  // particles per thread of the push kernel: 1, 2, 4 or 8
  // Background collisions (null: none this push).  Every particle alive
  // after the push draws its null-collision candidate test, Philox keyed by
  // (id, step, species); candidates (and particles faster than the cap) are
  // appended to mcc_cand, and the largest |v|^2 goes to mcc_v2max.
  const double* mcc_ctl;  // P_null, v^2 cap (emit.hpp)
  uint32_t* mcc_cand;
  unsigned int* mcc_ncand;
  unsigned long long* mcc_v2max;
  uint32_t mcc_step, mcc_key0, mcc_key1;
  // with scenarios (tet_seg): kMccCtl controls, |v|^2 maximum and random
  // key (2 words) per scenario, at mcc_ctl + s kMccCtl, mcc_v2max + s and
  // mcc_seg_key + 2 s
  const uint32_t* mcc_seg_key;
};

// blocks of one push launch over n particle slots with `chunks` particles
// per thread (one statistics partial each)
int push_blocks(int64_t n, int chunks);
template <typename T>
// with start/stop non-null they receive the kernel's dispatch timestamps
void launch_push(const PushArgs<T>& a, hipStream_t s, hipEvent_t start = nullptr,
                 hipEvent_t stop = nullptr);
template <typename T>
void launch_deposit(const PushArgs<T>& a, hipStream_t s);

}  // namespace scdat
