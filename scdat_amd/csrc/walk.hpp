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

// Grid cell of x along axis r of the lattice and the fractional coordinate
// in it, from the run's first line and spacing; false outside the grid or
// within kLatEps of a grid line.
__device__ __forceinline__ bool lat_axis(const LatticeDev& L, int r, float x, int& i, float& u) {
  if (!(x >= L.lo[r] && x < L.hi[r])) return false;
  // the run holding x, selected with compile-time indices (no per-lane
  // indexing of the argument tables)
  float x0 = L.sx[r][0], inv = L.sinv[r][0];
  int j0 = L.si[r][0], j1 = L.si[r][1];
#pragma unroll
  for (int k = 1; k < kLatSeg; ++k) {
    if (k < L.nseg[r] && x >= L.sx[r][k]) {
      x0 = L.sx[r][k];
      inv = L.sinv[r][k];
      j0 = L.si[r][k];
      j1 = L.si[r][k + 1];
    }
  }
  const float t = (x - x0) * inv;
  const float ft = floorf(t);
  i = j0 + (int)ft;
  u = t - ft;
  return i >= j0 && i < j1 && u >= kLatEps && u <= 1.0f - kLatEps;
}

// End tet of a sub-step that left its start tet, from the start tet's
// hexahedron th to xn on the lattice; -3 when the sub-step must walk.
// hx_out receives the end hexahedron (packed like th).
__device__ __forceinline__ int lat_locate(const LatticeDev& L, uint32_t th, const float xn[3], uint32_t& hx_out) {
  int i[3];
  float u[3];
  if (!lat_axis(L, 0, xn[0], i[0], u[0]) || !lat_axis(L, 1, xn[1], i[1], u[1]) ||
      !lat_axis(L, 2, xn[2], i[2], u[2]))
    return -3;
  // Kuhn permutation (p0, p1, p2), u[p0] >= u[p1] >= u[p2], in the order
  // (012, 021, 102, 120, 201, 210); near-equal coordinates walk
  const float d01 = u[0] - u[1], d02 = u[0] - u[2], d12 = u[1] - u[2];
  if (fabsf(d01) < kLatEps || fabsf(d02) < kLatEps || fabsf(d12) < kLatEps) return -3;
  const bool a01 = d01 > 0.0f, a02 = d02 > 0.0f, a12 = d12 > 0.0f;
  const int p = (a01 && a02) ? (a12 ? 0 : 1) : ((!a01 && a12) ? (a02 ? 2 : 3) : (a01 ? 4 : 5));
  const int64_t hx = ((int64_t)i[0] * L.nh[1] + i[1]) * L.nh[2] + i[2];
  const uint32_t e = L.hex_tet[6 * hx + p];
  const int di = abs(i[0] - (int)(th & 1023u)), dj = abs(i[1] - (int)((th >> 10) & 1023u)),
            dk = abs(i[2] - (int)((th >> 20) & 1023u));
  const int dm = di > dj ? (di > dk ? di : dk) : (dj > dk ? dj : dk);
  if (e == 0u || dm >= (int)(e >> 29)) return -3;
  hx_out = (uint32_t)i[0] | (uint32_t)i[1] << 10 | (uint32_t)i[2] << 20;
  return (int)(e & 0x1FFFFFFFu) - 1;
}

}  // namespace scdat
