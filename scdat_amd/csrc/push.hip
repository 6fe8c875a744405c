// Fused particle kernel of the PIC-MCC cycle: P1 gather of nodal E, Boris
// push with the reference's adaptive sub-stepping, straight-segment tet walk
// with boundary events, and fixed-point charge deposition.
//
// Replaces, per particle, CrossTetrahedron::integrateTrajectory
// (Tools/Particle/src/PICParticleAlgorithms.cpp:536-740), the P1 gather
// (:107-181), deposition at the end position (:325-369, NearestPoint) and the
// boundary-event handling of PICCycle (Tools/PICcore/src/PICCycle.cpp:1221-1379).
//
// Layout: particles are SoA (x,y,z,vx,vy,vz,cell), kept sorted by cell, so a
// wavefront's 64 lanes mostly share a handful of tets.  Deposition reduces the
// integer barycentric weights over runs of equal cell inside the wavefront
// (segmented scan on 32-bit integers), then one 64-bit atomic per (run, node)
// lands in the node accumulator.  Integer addition is associative, so the
// result is independent of launch order and bit-exact.
#include "device_common.hpp"
#include "push.hpp"

namespace scdat {

template <typename T>
__device__ inline void gather_E(const PushArgs<T>& a, const TetMap<T>& m, const T lam[4], T E[3]) {
  E[0] = E[1] = E[2] = T(0);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const Vec4<T> e = a.E[m.node[k]];
    E[0] = fma_t(lam[k], e.x, E[0]);
    E[1] = fma_t(lam[k], e.y, E[1]);
    E[2] = fma_t(lam[k], e.z, E[2]);
  }
}

// Straight-segment walk from x0 (inside cell `cur`) to x1.  Returns the final
// cell, -1 when the particle left through an absorbing/conductor face (group
// and hit point set), -2 when the walk limit was exceeded.
template <typename T>
__device__ int walk(const PushArgs<T>& a, int cur, T x0[3], T x1[3], T v[3], int& group,
                    T hit[3], TetMap<T>& m) {
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
    const int4 nb4 = a.nb[cur];
    const int n = kb == 0 ? nb4.x : (kb == 1 ? nb4.y : (kb == 2 ? nb4.z : nb4.w));
    if (n >= 0) {
      cur = n;
      m = a.tets[cur];
      continue;
    }
    const int4 g4 = a.groups[cur];
    const int g = kb == 0 ? g4.x : (kb == 1 ? g4.y : (kb == 2 ? g4.z : g4.w));
#pragma unroll
    for (int r = 0; r < 3; ++r) hit[r] = x0[r] + tb * (x1[r] - x0[r]);
    const int ty = (g >= 0 && g < kMaxGroups) ? a.gtype[g] : BC_ABSORB;
    if (ty == BC_PERIODIC) {
      const int4 p4 = a.partner[cur];
      const int pt = kb == 0 ? p4.x : (kb == 1 ? p4.y : (kb == 2 ? p4.z : p4.w));
      if (pt >= 0) {
#pragma unroll
        for (int r = 0; r < 3; ++r) {
          x0[r] = hit[r] + a.gshift[g][r];
          x1[r] += a.gshift[g][r];
        }
        cur = pt;
        m = a.tets[cur];
        continue;
      }
    } else if (ty == BC_REFLECT) {
      T nrm[3];
      if (kb == 0) {
#pragma unroll
        for (int r = 0; r < 3; ++r) nrm[r] = -(m.a[r] + m.a[3 + r] + m.a[6 + r]);
      } else {
#pragma unroll
        for (int r = 0; r < 3; ++r) nrm[r] = m.a[3 * (kb - 1) + r];
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
    return -1;
  }
  return -2;
}

template <typename T>
__global__ void __launch_bounds__(kBlock) push_kernel(PushArgs<T> a) {
  const uint32_t lb = xcd_remap(blockIdx.x, gridDim.x);
  const int64_t i = (int64_t)lb * kBlock + threadIdx.x;
  const int lane = threadIdx.x & 63;
  int key = -1;
  int32_t f[4] = {0, 0, 0, 0};
  double ke = 0.0;
  int alive = 0;
  int c = (i < a.n) ? a.cell[i] : -1;
  if (c >= 0) {
    T x[3] = {a.x[i], a.y[i], a.z[i]};
    T v[3] = {a.vx[i], a.vy[i], a.vz[i]};
    const T v0[3] = {v[0], v[1], v[2]};
    TetMap<T> m = a.tets[c];
    T remaining = a.dt;
    int sub = 0;
    int dead = 0;
    int group = -1;
    T hit[3] = {x[0], x[1], x[2]};
    // Sub-step loop, PICParticleAlgorithms.cpp:593-701.
    while (remaining > T(1e-20) && sub < a.max_sub) {
      ++sub;
      T lam[4];
      bary(m, x, lam);
      T E[3];
      gather_E(a, m, lam, E);
      const T h = a.adaptive ? adaptive_dt(v0, E, a.bmag, a.qm, remaining, a.floor_ratio) : remaining;
      boris(v, E, a.B, a.has_b != 0, a.qm, h);
      T xs[3] = {x[0], x[1], x[2]};
      T xn[3] = {fma_t(v[0], h, x[0]), fma_t(v[1], h, x[1]), fma_t(v[2], h, x[2])};
      const int r = walk(a, c, xs, xn, v, group, hit, m);
      if (r < 0) {
        dead = r;
        break;
      }
      c = r;
      x[0] = xn[0];
      x[1] = xn[1];
      x[2] = xn[2];
      remaining -= h;
    }
    if (dead) {
      if (dead == -2) {
        atomicAdd(a.lost, 1ull);
      } else {
        const int g = (group >= 0 && group < kMaxGroups) ? group : 0;
        const double k = 0.5 * (double)a.mass *
                         ((double)v[0] * v[0] + (double)v[1] * v[1] + (double)v[2] * v[2]);
        atomicAdd(&a.ledger_count[g * kMaxSpecies + a.species], 1ull);
        atomicAdd(&a.ledger_energy[g * kMaxSpecies + a.species], k * a.energy_weight);
      }
      a.cell[i] = -1;
      a.x[i] = hit[0];
      a.y[i] = hit[1];
      a.z[i] = hit[2];
    } else {
      T lam[4];
      bary(m, x, lam);
      lambda_fixed(lam, f);
      key = c;
      alive = 1;
      ke = 0.5 * (double)a.mass * (double)a.weight *
           ((double)v[0] * v[0] + (double)v[1] * v[1] + (double)v[2] * v[2]);
      a.x[i] = x[0];
      a.y[i] = x[1];
      a.z[i] = x[2];
      a.cell[i] = c;
    }
    a.vx[i] = v[0];
    a.vy[i] = v[1];
    a.vz[i] = v[2];
  }

  // ---- wave-segmented reduction of integer weights by cell -----------------
  const int prev = __shfl_up(key, 1, 64);
  const bool head = lane == 0 || prev != key;
  const uint64_t heads = __ballot(head);
  const uint64_t le = lane == 63 ? ~0ull : ((1ull << (lane + 1)) - 1ull);
  const int start = 63 - __builtin_clzll(heads & le);
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int32_t o0 = __shfl_up(f[0], d, 64);
    const int32_t o1 = __shfl_up(f[1], d, 64);
    const int32_t o2 = __shfl_up(f[2], d, 64);
    const int32_t o3 = __shfl_up(f[3], d, 64);
    if (lane - d >= start) {
      f[0] += o0;
      f[1] += o1;
      f[2] += o2;
      f[3] += o3;
    }
  }
  const int next = __shfl_down(key, 1, 64);
  const bool tail = lane == 63 || next != key;
  if (tail && key >= 0) {
    const int4 nd = *reinterpret_cast<const int4*>(a.tets[key].node);
    atomicAdd(&a.rho[nd.x], (unsigned long long)(a.cs * (int64_t)f[0]));
    atomicAdd(&a.rho[nd.y], (unsigned long long)(a.cs * (int64_t)f[1]));
    atomicAdd(&a.rho[nd.z], (unsigned long long)(a.cs * (int64_t)f[2]));
    atomicAdd(&a.rho[nd.w], (unsigned long long)(a.cs * (int64_t)f[3]));
  }

  // ---- per-block kinetic energy and alive count (deterministic partials) ----
  __shared__ double lds[kBlock / 64];
  const double ks = block_sum(ke, lds);
  const double as = block_sum((double)alive, lds);
  if (threadIdx.x == 0) {
    a.ke_partial[lb] = ks;
    a.alive_partial[lb] = (int32_t)as;
  }
}

// Deposit only (no motion): integer weights of the current positions.
template <typename T>
__global__ void __launch_bounds__(kBlock) deposit_kernel(PushArgs<T> a) {
  const uint32_t lb = xcd_remap(blockIdx.x, gridDim.x);
  const int64_t i = (int64_t)lb * kBlock + threadIdx.x;
  const int lane = threadIdx.x & 63;
  int key = -1;
  int32_t f[4] = {0, 0, 0, 0};
  const int c = (i < a.n) ? a.cell[i] : -1;
  if (c >= 0) {
    const TetMap<T> m = a.tets[c];
    const T x[3] = {a.x[i], a.y[i], a.z[i]};
    T lam[4];
    bary(m, x, lam);
    lambda_fixed(lam, f);
    key = c;
  }
  const int prev = __shfl_up(key, 1, 64);
  const bool head = lane == 0 || prev != key;
  const uint64_t heads = __ballot(head);
  const uint64_t le = lane == 63 ? ~0ull : ((1ull << (lane + 1)) - 1ull);
  const int start = 63 - __builtin_clzll(heads & le);
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int32_t o0 = __shfl_up(f[0], d, 64);
    const int32_t o1 = __shfl_up(f[1], d, 64);
    const int32_t o2 = __shfl_up(f[2], d, 64);
    const int32_t o3 = __shfl_up(f[3], d, 64);
    if (lane - d >= start) {
      f[0] += o0;
      f[1] += o1;
      f[2] += o2;
      f[3] += o3;
    }
  }
  const int next = __shfl_down(key, 1, 64);
  const bool tail = lane == 63 || next != key;
  if (tail && key >= 0) {
    const int4 nd = *reinterpret_cast<const int4*>(a.tets[key].node);
    atomicAdd(&a.rho[nd.x], (unsigned long long)(a.cs * (int64_t)f[0]));
    atomicAdd(&a.rho[nd.y], (unsigned long long)(a.cs * (int64_t)f[1]));
    atomicAdd(&a.rho[nd.z], (unsigned long long)(a.cs * (int64_t)f[2]));
    atomicAdd(&a.rho[nd.w], (unsigned long long)(a.cs * (int64_t)f[3]));
  }
}

template <typename T>
void launch_push(const PushArgs<T>& a, hipStream_t s) {
  if (a.n <= 0) return;
  const uint32_t g = (uint32_t)((a.n + kBlock - 1) / kBlock);
  hipLaunchKernelGGL(push_kernel<T>, dim3(g), dim3(kBlock), 0, s, a);
}
template <typename T>
void launch_deposit(const PushArgs<T>& a, hipStream_t s) {
  if (a.n <= 0) return;
  const uint32_t g = (uint32_t)((a.n + kBlock - 1) / kBlock);
  hipLaunchKernelGGL(deposit_kernel<T>, dim3(g), dim3(kBlock), 0, s, a);
}
template void launch_push<float>(const PushArgs<float>&, hipStream_t);
template void launch_push<double>(const PushArgs<double>&, hipStream_t);
template void launch_deposit<float>(const PushArgs<float>&, hipStream_t);
template void launch_deposit<double>(const PushArgs<double>&, hipStream_t);

}  // namespace scdat
