// Field gather and straight-segment tet walk shared by the push kernels.
#pragma once
#include "device_common.hpp"
#include "push.hpp"

namespace scdat {

// P1 gather of nodal E at barycentric lam over the tet's nodes nd.
template <typename T>
__device__ __forceinline__ void gather_E(const PushArgs<T>& a, const int32_t nd[4], const T lam[4], T E[3]) {
  E[0] = E[1] = E[2] = T(0);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const Vec4<T> e = a.E[nd[k]];
    E[0] = fma_t(lam[k], e.x, E[0]);
    E[1] = fma_t(lam[k], e.y, E[1]);
    E[2] = fma_t(lam[k], e.z, E[2]);
  }
}
template <typename T>
__device__ __forceinline__ void gather_E(const PushArgs<T>& a, const TetMap<T>& m, const T lam[4], T E[3]) {
  gather_E(a, m.node, lam, E);
}

// The walk reads a tet's map and its face neighbours either from the tet map
// and the neighbour array, or from the walk record that holds both.
template <typename T>
__device__ __forceinline__ void load_map(const PushArgs<T>& a, int t, TetMap<T>& m) { m = a.tets[t]; }
template <typename T>
__device__ __forceinline__ void load_map(const PushArgs<T>& a, int t, WalkMap<T>& m) { m = a.wmap[t]; }
template <typename T>
__device__ __forceinline__ int4 face_nb(const PushArgs<T>& a, int t, const TetMap<T>&) { return a.nb[t]; }
template <typename T>
__device__ __forceinline__ int4 face_nb(const PushArgs<T>&, int, const WalkMap<T>& m) {
  return make_int4(m.nb[0], m.nb[1], m.nb[2], m.nb[3]);
}

// Straight-segment walk from x0 (inside cell `cur`) to x1.  Returns the final
// cell, -1 when the particle left through an absorbing/conductor face (group,
// hit point, exit face and its tet set), -2 when the walk limit was exceeded.
template <typename T, typename M>
__device__ __forceinline__ int walk(const PushArgs<T>& a, int cur, T x0[3], T x1[3], T v[3], int& group,
                    T hit[3], M& m, int& face, int& last) {
  const T tol = a.tol;
  for (int it = 0; it < a.max_walk; ++it) {
    T l1[4];
    bary(m, x1, l1);
    if (l1[0] >= -tol && l1[1] >= -tol && l1[2] >= -tol && l1[3] >= -tol) return cur;
    T l0[4];
    bary(m, x0, l0);
    int kb = -1;
    T tb = T(2);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (l1[k] < -tol && l1[k] < l0[k]) {
        const T t = l0[k] / (l0[k] - l1[k]);
        if (t < tb) {
          tb = t;
          kb = k;
        }
      }
    }
    if (kb < 0) {
      T mn = T(0);
#pragma unroll
      for (int k = 0; k < 4; ++k)
        if (l1[k] < mn) {
          mn = l1[k];
          kb = k;
        }
      tb = T(0);
    }
    tb = tb < T(0) ? T(0) : (tb > T(1) ? T(1) : tb);
    const int4 nb4 = face_nb(a, cur, m);
    const int n = kb == 0 ? nb4.x : (kb == 1 ? nb4.y : (kb == 2 ? nb4.z : nb4.w));
    if (n >= 0) {
      cur = n;
      load_map(a, cur, m);
      continue;
    }
    const int4 g4 = a.groups[cur];
    const int g = kb == 0 ? g4.x : (kb == 1 ? g4.y : (kb == 2 ? g4.z : g4.w));
#pragma unroll
    for (int r = 0; r < 3; ++r) hit[r] = x0[r] + tb * (x1[r] - x0[r]);
    const uint32_t gbit = (g >= 0 && g < kMaxGroups) ? (1u << g) : 0u;
    if (a.periodic_mask & gbit) {
      const int4 p4 = a.partner[cur];
      const int pt = kb == 0 ? p4.x : (kb == 1 ? p4.y : (kb == 2 ? p4.z : p4.w));
      if (pt >= 0) {
#pragma unroll
        for (int r = 0; r < 3; ++r) {
          const T sh = a.gshift[3 * g + r];
          x0[r] = hit[r] + sh;
          x1[r] += sh;
        }
        cur = pt;
        load_map(a, cur, m);
        continue;
      }
    } else if (a.reflect_mask & gbit) {
      // inward normal of face kb = grad lambda_kb (rows of the map)
      T nrm[3];
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        const T g1 = m.a[r], g2 = m.a[3 + r], g3 = m.a[6 + r];
        nrm[r] = kb == 0 ? -(g1 + g2 + g3) : (kb == 1 ? g1 : (kb == 2 ? g2 : g3));
      }
      const T inl = T(1) / sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
#pragma unroll
      for (int r = 0; r < 3; ++r) nrm[r] *= inl;
      const T dd = (x1[0] - hit[0]) * nrm[0] + (x1[1] - hit[1]) * nrm[1] + (x1[2] - hit[2]) * nrm[2];
      const T vn = v[0] * nrm[0] + v[1] * nrm[1] + v[2] * nrm[2];
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        x1[r] -= T(2) * dd * nrm[r];
        v[r] -= T(2) * vn * nrm[r];
        x0[r] = hit[r];
      }
      continue;
    }
    group = g;
    face = kb;
    last = cur;
    return -1;
  }
  return -2;
}

}  // namespace scdat
