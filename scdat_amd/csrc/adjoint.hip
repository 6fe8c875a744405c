// Discrete adjoint of the collisionless charging cycle on gfx950.
//
// Forward step n (FP64 charge, particle storage in T):
//   v1 = v0 + (q/m) E_n(x0) dt           (Boris with B = 0, PICParticleAlgorithms.cpp:803-851)
//   x1 = x0 + v1 dt, tet walk; hits on conductor groups add q w to the
//   floating conductor's charge and, for electrons, the Sternglass secondary
//   charge e w delta(E_hit) leaves it; survivors deposit q w lambda(x1)
//   Q_{n+1} = Q_n + hits + dt (J_ph(V_n) A_sunlit - G V_n)       (host)
//   phi_0 = A^-1 b(rho), V_{n+1} = (Q_{n+1} + sum_C rho - kc.phi_0) / C,
//   phi = phi_0 + V phi_1, E_{n+1} = -L phi                       (field kernels)
// The backward kernels here run the transposed chain of one step: the
// charge adjoint to positions (grad lambda of the final tet), the secondary
// charge to impact velocities and to the yield parameters, the drift and the
// E kick back to step n, and the kick's field adjoint scattered to the nodes.
// Deposits and scatters use FP64 atomics (order-dependent at the last bit).
#include "adjoint.hpp"
#include "walk.hpp"

namespace scdat {

namespace {

constexpr double kQe = 1.602176634e-19;

template <typename T>
__device__ inline void grad_lambda(const TetMap<T>& m, double g[4][3]) {
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    g[1][r] = (double)m.a[r];
    g[2][r] = (double)m.a[3 + r];
    g[3][r] = (double)m.a[6 + r];
    g[0][r] = -(g[1][r] + g[2][r] + g[3][r]);
  }
}

// Sternglass shape s(E) = r exp(2 - 2 sqrt r), r = E / E_max, and ds/dE.
__device__ inline void sternglass(double E, double emax, double& s, double& ds) {
  if (!(E > 0.0)) {
    s = 0.0;
    ds = 0.0;
    return;
  }
  const double r = E / emax, sr = sqrt(r), ex = exp(2.0 - 2.0 * sr);
  s = r * ex;
  ds = ex * (1.0 - sr) / emax;
}

__device__ inline void publish_partials(double p[kAdjPartials], double* out) {
  __shared__ double lds[kBlock / 64];
#pragma unroll
  for (int k = 0; k < kAdjPartials; ++k) {
    const double t = block_sum(p[k], lds);
    if (threadIdx.x == 0) out[(size_t)blockIdx.x * kAdjPartials + k] = t;
  }
}

// Centroid of a tet from its affine map: the point with lambda_1..3 = 1/4,
// x = x0 + J (1/4, 1/4, 1/4), J the inverse of the stored map a.
template <typename T>
__device__ inline void tet_centroid(const TetMap<T>& m, T c[3]) {
  const double a00 = m.a[0], a01 = m.a[1], a02 = m.a[2], a10 = m.a[3], a11 = m.a[4], a12 = m.a[5], a20 = m.a[6],
               a21 = m.a[7], a22 = m.a[8];
  const double det = a00 * (a11 * a22 - a12 * a21) - a01 * (a10 * a22 - a12 * a20) + a02 * (a10 * a21 - a11 * a20);
  const double q = 0.25 / det;
  // J (1,1,1) / 4 = adj(a) (1,1,1) / (4 det)
  const double y0 = (a11 * a22 - a12 * a21) - (a01 * a22 - a02 * a21) + (a01 * a12 - a02 * a11);
  const double y1 = -(a10 * a22 - a12 * a20) + (a00 * a22 - a02 * a20) - (a00 * a12 - a02 * a10);
  const double y2 = (a10 * a21 - a11 * a20) - (a00 * a21 - a01 * a20) + (a00 * a11 - a01 * a10);
  c[0] = (T)((double)m.x0[0] + q * y0);
  c[1] = (T)((double)m.x0[1] + q * y1);
  c[2] = (T)((double)m.x0[2] + q * y2);
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_adj_forward(AdjArgs<T> a) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  double part[kAdjPartials];
#pragma unroll
  for (int k = 0; k < kAdjPartials; ++k) part[k] = 0.0;
  if (i < a.n) {
    const int c = a.s0.cell[i];
    T x[3] = {a.s0.x[i], a.s0.y[i], a.s0.z[i]};
    T v[3] = {a.s0.vx[i], a.s0.vy[i], a.s0.vz[i]};
    int c1 = -1;
    if (c >= 0) {
      const int s = a.sp[i];
      TetMap<T> m = a.mesh.tets[c];
      T lam[4];
      bary(m, x, lam);
      T E[3];
      gather_E(a.mesh, m, lam, E);
      const T B0[3] = {T(0), T(0), T(0)};
      boris(v, E, B0, false, a.qm[s], a.dt);
      T xs[3] = {x[0], x[1], x[2]};
      T xn[3] = {fma_t(v[0], a.dt, x[0]), fma_t(v[1], a.dt, x[1]), fma_t(v[2], a.dt, x[2])};
      int group = -1, face = 0, last = c;
      T hit[3] = {x[0], x[1], x[2]};
      int r = walk(a.mesh, c, xs, xn, v, group, hit, m, face, last);
      if (a.event_mode == 2) {
        const int32_t ev = a.events[i];
        if (ev >= 0 && (ev >> 8) == a.step) {
          // absorbed here in the recorded run: absorbed here, same group
          if (r >= 0) {
#pragma unroll
            for (int q = 0; q < 3; ++q) hit[q] = xn[q];
          }
          group = ev & 0xFF;
          r = -1;
        } else if (r == -1) {
          // kept in the recorded run: relocate from the cell the recorded run
          // reached (a straight path that clips a convex edge of the boundary
          // would otherwise be stopped on the far side of the edge); a
          // residual overshoot beyond a face is projected back onto the
          // face's plane and the tangential remainder walked (a face, an
          // edge, a corner).  Continuous, and the identity while the walk
          // stays inside.
          const int cb = a.path[(int64_t)a.step * a.n + i];
          const TetMap<T> mb = a.mesh.tets[cb >= 0 ? cb : last];
          T st[3];
          tet_centroid(mb, st);
          int cur = cb >= 0 ? cb : last;
          m = mb;
          {
            T xe[3] = {xn[0], xn[1], xn[2]};
            r = walk(a.mesh, cur, st, xe, v, group, hit, m, face, last);
          }
          for (int pass = 0; pass < 3 && r == -1; ++pass) {
            T nrm[3];  // inward normal of the crossed face of `last`
#pragma unroll
            for (int q = 0; q < 3; ++q) {
              const T g1 = m.a[q], g2 = m.a[3 + q], g3 = m.a[6 + q];
              nrm[q] = face == 0 ? -(g1 + g2 + g3) : (face == 1 ? g1 : (face == 2 ? g2 : g3));
            }
            const T inl = T(1) / sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
            T d = T(0);
#pragma unroll
            for (int q = 0; q < 3; ++q) {
              nrm[q] *= inl;
              d += (xn[q] - hit[q]) * nrm[q];  // < 0: beyond the face
            }
#pragma unroll
            for (int q = 0; q < 3; ++q) {
              xn[q] -= d * nrm[q];
              st[q] = hit[q];
            }
            cur = last;
            T xe[3] = {xn[0], xn[1], xn[2]};
            r = walk(a.mesh, cur, st, xe, v, group, hit, m, face, last);
          }
          if (r < 0) {
            // not reached in three projections: stay at the last crossing
            r = cur;
            m = a.mesh.tets[cur];
#pragma unroll
            for (int q = 0; q < 3; ++q) xn[q] = st[q];
          }
        }
      }
      if (a.event_mode == 1) {
        if (r == -1) a.events[i] = (a.step << 8) | (group & 0xFF);
        a.path[(int64_t)a.step * a.n + i] = r;
      }
      if (r >= 0) {
        c1 = r;
        x[0] = xn[0];
        x[1] = xn[1];
        x[2] = xn[2];
        T l1[4];
        bary(m, xn, l1);
#pragma unroll
        for (int k = 0; k < 4; ++k) unsafeAtomicAdd(&a.rho[m.node[k]], a.qw[s] * (double)l1[k]);
      } else if (r == -1) {
        const int g = (group >= 0 && group < kMaxGroups) ? group : 0;
        c1 = -2 - g;
        x[0] = hit[0];
        x[1] = hit[1];
        x[2] = hit[2];
        if ((a.cond_mask >> g) & 1u) {
          part[0] += a.qw[s];
          const int f = g - kAdjFaceGroup0;
          if (a.electron[s] && f >= 0 && f < 6 && a.dmax[g] > 0.0) {
            const double v2 = (double)v[0] * v[0] + (double)v[1] * v[1] + (double)v[2] * v[2];
            double sh, ds;
            sternglass(0.5 * a.mass[s] * v2 / kQe, a.emax, sh, ds);
            part[1 + f] += fabs(a.qw[s]) * a.dmax[g] * sh;
          }
        }
      }
    }
    a.s1.x[i] = x[0];
    a.s1.y[i] = x[1];
    a.s1.z[i] = x[2];
    a.s1.vx[i] = v[0];
    a.s1.vy[i] = v[1];
    a.s1.vz[i] = v[2];
    a.s1.cell[i] = c1;
  }
  publish_partials(part, a.partial);
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_adj_backward(AdjArgs<T> a) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  double part[kAdjPartials];
#pragma unroll
  for (int k = 0; k < kAdjPartials; ++k) part[k] = 0.0;
  const int c0 = i < a.n ? a.s0.cell[i] : -1;
  if (c0 >= 0) {
    const int c1 = a.s1.cell[i], s = a.sp[i];
    const double dt = (double)a.dt, qmdt = (double)a.qm[s] * dt;
    double xb[3], vb[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      xb[r] = a.xb[3 * i + r];
      vb[r] = a.vb[3 * i + r];
    }
    if (c1 >= 0) {
      // deposit of step n+1: rho_k += q w lambda_k(x1)
      const TetMap<T> m1 = a.mesh.tets[c1];
      double g[4][3];
      grad_lambda(m1, g);
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const double rb = a.rhobar[m1.node[k]] * a.qw[s];
#pragma unroll
        for (int r = 0; r < 3; ++r) xb[r] = fma(rb, g[k][r], xb[r]);
      }
      // drift x1 = x0 + v1 dt
#pragma unroll
      for (int r = 0; r < 3; ++r) vb[r] = fma(xb[r], dt, vb[r]);
    } else {
#pragma unroll
      for (int r = 0; r < 3; ++r) xb[r] = vb[r] = 0.0;
      const int g = -2 - c1;
      const int f = g - kAdjFaceGroup0;
      if (c1 <= -2 && ((a.cond_mask >> g) & 1u) && a.electron[s] && f >= 0 && f < 6 && a.dmax[g] > 0.0) {
        // secondary charge |q| w dmax s(E(v1)) of the impact
        const double v1[3] = {(double)a.s1.vx[i], (double)a.s1.vy[i], (double)a.s1.vz[i]};
        const double v2 = v1[0] * v1[0] + v1[1] * v1[1] + v1[2] * v1[2];
        double sh, ds;
        sternglass(0.5 * a.mass[s] * v2 / kQe, a.emax, sh, ds);
        const double ew = fabs(a.qw[s]);
        const double kv = a.qbar * ew * a.dmax[g] * ds * a.mass[s] / kQe;  // dE/dv = m v / e
#pragma unroll
        for (int r = 0; r < 3; ++r) vb[r] = kv * v1[r];
        part[0] += a.qbar * ew * a.yield[g] * sh;
        part[1 + f] += a.qbar * ew * a.delta_max * sh;
      }
    }
    // kick v1 = v0 + (q/m) E_n(x0) dt
    const TetMap<T> m0 = a.mesh.tets[c0];
    const T x0[3] = {a.s0.x[i], a.s0.y[i], a.s0.z[i]};
    T lam[4];
    bary(m0, x0, lam);
    double g[4][3];
    grad_lambda(m0, g);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int node = m0.node[k];
      const Vec4<T> e = a.mesh.E[node];
      const double d = qmdt * ((double)e.x * vb[0] + (double)e.y * vb[1] + (double)e.z * vb[2]);
#pragma unroll
      for (int r = 0; r < 3; ++r) xb[r] = fma(d, g[k][r], xb[r]);
      const double l = qmdt * (double)lam[k];
#pragma unroll
      for (int r = 0; r < 3; ++r) unsafeAtomicAdd(&a.Ebar[3 * (int64_t)node + r], l * vb[r]);
    }
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      a.xb[3 * i + r] = xb[r];
      a.vb[3 * i + r] = vb[r];
    }
  }
  publish_partials(part, a.partial);
}

// E_i = -sum_s w_is (phi_j - phi_i):  phibar_j -= w_is . Ebar_i,
// phibar_i += sum_s w_is . Ebar_i.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_adj_field_t(EllDev A, const double* __restrict__ Ebar,
                                                        double* phibar) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= A.n) return;
  const double ex = Ebar[3 * i], ey = Ebar[3 * i + 1], ez = Ebar[3 * i + 2];
  if (ex == 0.0 && ey == 0.0 && ez == 0.0) return;
  const T* gx;
  const T* gy;
  const T* gz;
  if constexpr (sizeof(T) == 4) {
    gx = reinterpret_cast<const T*>(A.gx32);
    gy = reinterpret_cast<const T*>(A.gy32);
    gz = reinterpret_cast<const T*>(A.gz32);
  } else {
    gx = reinterpret_cast<const T*>(A.gx64);
    gy = reinterpret_cast<const T*>(A.gy64);
    gz = reinterpret_cast<const T*>(A.gz64);
  }
  double self = 0.0;
  for (int k = 1; k < A.w; ++k) {
    const int64_t o = (int64_t)k * A.n + i;
    const double wd = (double)gx[o] * ex + (double)gy[o] * ey + (double)gz[o] * ez;
    if (wd == 0.0) continue;
    unsafeAtomicAdd(&phibar[A.col[o]], -wd);
    self += wd;
  }
  unsafeAtomicAdd(&phibar[i], self);
}

__global__ void __launch_bounds__(kBlock) k_adj_dot(int64_t n, const double* a, const double* b, double* partial) {
  __shared__ double lds[kBlock / 64];
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const double v = i < n ? a[i] * b[i] : 0.0;
  const double t = block_sum(v, lds);
  if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

__global__ void __launch_bounds__(kBlock) k_adj_scatter(int64_t m, const int32_t* idx, const double* val,
                                                        double scale, double* v) {
  const int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (j < m) v[idx[j]] += scale * val[j];  // idx holds distinct nodes
}

__global__ void __launch_bounds__(kBlock) k_adj_rhobar(int64_t n, const int32_t* kind, const double* y,
                                                       double cond, double* out) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const int k = kind[i];
  out[i] = k == 0 ? y[i] : (k == 2 ? cond : 0.0);
}

}  // namespace

int adj_blocks(int64_t n) { return (int)((n + kBlock - 1) / kBlock); }

template <typename T>
void launch_adj_forward(const AdjArgs<T>& a, hipStream_t s) {
  if (a.n <= 0) return;
  hipLaunchKernelGGL(k_adj_forward<T>, dim3(adj_blocks(a.n)), dim3(kBlock), 0, s, a);
}
template <typename T>
void launch_adj_backward(const AdjArgs<T>& a, hipStream_t s) {
  if (a.n <= 0) return;
  hipLaunchKernelGGL(k_adj_backward<T>, dim3(adj_blocks(a.n)), dim3(kBlock), 0, s, a);
}
template <typename T>
void launch_adj_field_t(const EllDev& A, const double* Ebar, double* phibar, hipStream_t s) {
  hipLaunchKernelGGL(k_adj_field_t<T>, dim3(adj_blocks(A.n)), dim3(kBlock), 0, s, A, Ebar, phibar);
}
void launch_adj_dot(int64_t n, const double* a, const double* b, double* partial, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_adj_dot, dim3(adj_blocks(n)), dim3(kBlock), 0, s, n, a, b, partial);
}
void launch_adj_scatter(int64_t m, const int32_t* idx, const double* val, double scale, double* v, hipStream_t s) {
  if (m <= 0) return;
  hipLaunchKernelGGL(k_adj_scatter, dim3(adj_blocks(m)), dim3(kBlock), 0, s, m, idx, val, scale, v);
}
void launch_adj_rhobar(int64_t n, const int32_t* kind, const double* y, double cond, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_adj_rhobar, dim3(adj_blocks(n)), dim3(kBlock), 0, s, n, kind, y, cond, out);
}

#define SCDAT_ADJ_INST(T)                                                              \
  template void launch_adj_forward<T>(const AdjArgs<T>&, hipStream_t);                 \
  template void launch_adj_backward<T>(const AdjArgs<T>&, hipStream_t);                \
  template void launch_adj_field_t<T>(const EllDev&, const double*, double*, hipStream_t);
SCDAT_ADJ_INST(float)
SCDAT_ADJ_INST(double)
#undef SCDAT_ADJ_INST

}  // namespace scdat
