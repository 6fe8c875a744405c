// Launch interface of the fused particle kernel.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "device_common.hpp"
#include "emit.hpp"

namespace scdat {

// Lattice location for the FP32 push.  On a mesh of Kuhn tetrahedra over a
// rectilinear grid (the box mesher's: six tets per hexahedron around its
// v000-v111 diagonal), the tet holding a point follows from arithmetic: the
// hexahedron from the grid line runs of each axis, the tet from the
// descending order of the point's fractional coordinates u in it (its
// barycentric coordinates are 1 - u_p0, u_p0 - u_p1, u_p1 - u_p2, u_p2).
// A sub-step whose end point left its start tet is located this way when
// (a) every barycentric coordinate of the end point is at least kLatEps, so
// the end tet is the one the face walk reaches (no face is within rounding
// of the point), and (b) the end hexahedron's clearance d exceeds the
// Chebyshev distance from the start hexahedron, so every hexahedron between
// them is whole and the segment crosses no boundary face.  Otherwise the
// sub-step walks.  Results are those of the walk, bit for bit.
constexpr int kLatSeg = 4;        // uniform-spacing runs of grid lines per axis
constexpr int kLatAxis = 1024;    // hexahedra per axis (10-bit indices)
constexpr float kLatEps = 1e-3f;  // smallest barycentric coordinate located on the grid
struct LatticeDev {
  const uint32_t* tet_hex;  // per tet: its hexahedron i | j << 10 | k << 20 (null: no lattice)
  // 6 per hexahedron (i slowest, k fastest), by Kuhn permutation:
  // (tet + 1) | clearance << 29, 0 where the tet is missing; clearance 0 (the
  // hexahedron is not whole) to 7 (whole hexahedra to Chebyshev distance 6)
  const uint32_t* hex_tet;
  int nh[3];                // hexahedra per axis
  int nseg[3];              // uniform-spacing runs per axis
  float sx[3][kLatSeg];     // first line of each run
  float sinv[3][kLatSeg];   // 1 / spacing of each run
  int si[3][kLatSeg + 1];   // index of each run's first line; si[r][nseg[r]] = nh[r]
  float lo[3], hi[3];       // first and last line of each axis
};

// Deposit buckets.  A push block reduces its particles' charge into an LDS
// table keyed by node; the table then goes to the nodes' 512-row ranges:
// each range has a bucket of (node, charge) entries, a block reserves room
// in a bucket with one atomic per range it touches and writes its entries
// with plain stores, and after the pushes one block per range sums its
// bucket in LDS and adds the sums to its own rows of the charge.  An entry
// that finds its bucket full (or a block touching more than kDepRanges
// ranges) goes to the charge with a global atomic.  Integer sums, so the
// charge is the same bits in any case.
constexpr int kDepRangeRows = 512;
constexpr int kDepRanges = 16;
struct alignas(16) DepEntry {
  int32_t node, pad;
  unsigned long long val;
};
struct DepBuckets {
  unsigned int* count;  // [ranges] entries appended since the last reduction (null: no buckets)
  DepEntry* ent;        // [ranges][cap]
  int cap;
};

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
  LatticeDev lat;  // FP32 only (lat.tet_hex null: every sub-step walks)
  DepBuckets dep;  // dep.count null: blocks add their tables with global atomics
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
// Sums the deposit buckets into rho (n_ranges = rows rounded up to
// kDepRangeRows) and empties them.
void launch_dep_reduce(const DepBuckets& b, unsigned long long* rho, int64_t n_rows, hipStream_t s);

}  // namespace scdat
