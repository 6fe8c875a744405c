// Launch interface of the discrete-adjoint kernels of the charging cycle.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "device_common.hpp"
#include "field.hpp"
#include "push.hpp"

namespace scdat {

// per-block partial sums: forward [0] charge absorbed by the conductor,
// [1 + k] secondary-emission charge of boundary group 6 + k; backward
// [0] d/d delta_max, [1 + k] d/d yield of group 6 + k
constexpr int kAdjPartials = 8;
constexpr int kAdjFaceGroup0 = 6;  // conductor faces are groups 6..11

// Particle state of one step (all species concatenated).  cell >= 0 alive;
// -1 gone before this step; -2 - g removed during the step that produced
// this state, through boundary group g.
template <typename T>
struct AdjState {
  T *x, *y, *z, *vx, *vy, *vz;
  int32_t* cell;
};

template <typename T>
struct AdjArgs {
  PushArgs<T> mesh;  // tets, nb, groups, walk controls, E of the step
  AdjState<T> s0, s1;  // state n -> state n + 1
  const int32_t* sp;   // species of each particle
  int64_t n;
  T qm[kMaxSpecies];
  double qw[kMaxSpecies], mass[kMaxSpecies];
  int electron[kMaxSpecies];
  T dt;
  uint32_t cond_mask;         // groups whose hits charge the floating conductor
  double dmax[kMaxGroups];    // Sternglass delta_max x face yield (0: none)
  double emax;                // Sternglass E_max (eV)
  double* rho;                // forward: FP64 nodal charge
  double* partial;            // [blocks][kAdjPartials]
  // Absorption events of the forward (the map the adjoint differentiates
  // holds them fixed): mode 0 free; 1 record: events[i] = step << 8 | group
  // of the step that absorbs particle i (-1 never); 2 replay a record: a
  // particle is absorbed exactly at its recorded step and group, and one
  // the record kept whose walk leaves the mesh is projected back onto the
  // boundary face's plane and walks the tangential remainder.
  // The record also keeps each particle's cell after every step (path[step
  // x n + i]), where a replayed particle that the record kept is relocated.
  int event_mode;
  int step;
  int32_t* events;
  int32_t* path;
  // backward
  double* xb;                 // 3n position adjoints (in: step n+1, out: step n)
  double* vb;                 // 3n velocity adjoints
  const double* rhobar;       // adjoint of the nodal charge of step n + 1
  double* Ebar;               // 3 n_nodes adjoint of E of step n (accumulated)
  double qbar;                // adjoint of the conductor charge of step n + 1
  double yield[kMaxGroups];   // face multipliers (dmax = delta_max x yield)
  double delta_max;
};

int adj_blocks(int64_t n);
template <typename T>
void launch_adj_forward(const AdjArgs<T>& a, hipStream_t s);
template <typename T>
void launch_adj_backward(const AdjArgs<T>& a, hipStream_t s);
// phibar += L^T Ebar for the least-squares field E = -L phi (weights of A)
template <typename T>
void launch_adj_field_t(const EllDev& A, const double* Ebar, double* phibar, hipStream_t s);
// per-block partials of a . b over n entries (adj_blocks(n) of them)
void launch_adj_dot(int64_t n, const double* a, const double* b, double* partial, hipStream_t s);
// v[idx[j]] += scale * val[j]
void launch_adj_scatter(int64_t m, const int32_t* idx, const double* val, double scale, double* v, hipStream_t s);
// out[i] = y[i] on free nodes, cond on conductor nodes, 0 on Dirichlet nodes
void launch_adj_rhobar(int64_t n, const int32_t* kind, const double* y, double cond, double* out, hipStream_t s);

}  // namespace scdat
