// Poisson solve and nodal field of the PIC-MCC cycle on gfx950.
//
// Operator: P1 FEM stiffness (or the reference graph Laplacian) over all
// nodes in slot-major ELL, Dirichlet and conductor rows eliminated
// symmetrically.  The floating conductor of
// Tools/FieldSolver/src/MultiBoundaryElectricFieldSolver.cpp:1172-1217 (node
// collapse onto an anchor dof) is solved exactly by superposition instead:
//   phi = phi_0 + V_f phi_1,  phi_0: conductor held at 0,  phi_1: conductor at
//   1 V with no charge (solved once),
//   V_f = (Q + sum_C load - sum_C (K phi_0)) / sum_C (K phi_1)   (Gauss law on
//   the conductor), which is the anchor row of the collapsed system.
// This keeps every row of the PCG operator at the mesh's own width.
//
// Solver: Jacobi PCG (MultiBoundary :1432-1513, ParallelSolver.cpp:33-106).
// Mixed mode runs the inner PCG in FP32 on a unit-norm residual and refines
// in FP64 (r = b - A x with FP64 values, x += |r| d) until
// ||r|| <= tol ||b||.  All dot products accumulate in FP64, reduced in a fixed
// order (wave shuffles, block partials summed by the last block), so a solve
// is bitwise reproducible.  Per iteration: one fused SpMV kernel (which also
// forms p = z + beta p on the fly) and one fused update kernel.
#include "device_common.hpp"
#include "field.hpp"

namespace scdat {

namespace {

template <typename T>
__device__ inline const T* vals(const EllDev& A);
template <>
__device__ inline const float* vals<float>(const EllDev& A) { return A.v32; }
template <>
__device__ inline const double* vals<double>(const EllDev& A) { return A.v64; }
template <typename T>
__device__ inline const T* dinv(const EllDev& A);
template <>
__device__ inline const float* dinv<float>(const EllDev& A) { return A.dinv32; }
template <>
__device__ inline const double* dinv<double>(const EllDev& A) { return A.dinv64; }

inline uint32_t grid_for(int64_t n) { return (uint32_t)((n + kBlock - 1) / kBlock); }

// Reductions.  Every reducing kernel publishes one partial per block; the
// totals are formed either by a one-block finalize kernel or, inside the PCG
// iteration, by every block of the *next* kernel summing the same partials in
// the same fixed order.  No device-scope fence or atomic is needed (kernel
// boundaries order the memory), and every block obtains bit-identical
// scalars, so a solve is bitwise reproducible.
constexpr int kPB = kBlock;           // PCG block: 4 waves, two rows per thread
constexpr int kRowsPerBlock = 2 * kPB;
enum Slot { S_BB = 0, S_CL, S_RR, S_RZ0, S_RZ1, S_RR0, S_RR1, S_PQ, S_COUNT };

__device__ inline double* slot(const FieldBufs& f, int s, int G) { return f.partial + (size_t)s * G; }

// Sum of G partials, identical in every block; result broadcast to all threads.
__device__ inline double sum_partials(const double* p, int G, double* lds) {
  double x = 0.0;
  for (int j = threadIdx.x; j < G; j += blockDim.x) x += p[j];
  __shared__ double bc;
  const double t = block_sum(x, lds);
  if (threadIdx.x == 0) bc = t;
  __syncthreads();
  return bc;
}

__device__ inline void publish(double v, double* p, double* lds) {
  const double t = block_sum(v, lds);
  if (threadIdx.x == 0) p[blockIdx.x] = t;
}

__global__ void __launch_bounds__(kPB) k_rhs(EllDev A, FieldBufs f,
                                                 const unsigned long long* rho, double unit,
                                                 int load_sign) {
  __shared__ double lds[kPB / 64];
  const int G = gridDim.x;
  double bb = 0.0, cl = 0.0;
  const int64_t r0 = (int64_t)blockIdx.x * kRowsPerBlock;
  for (int64_t i = r0 + threadIdx.x; i < r0 + kRowsPerBlock && i < A.n; i += kPB) {
    const int k = f.kind[i];
    const double load = (double)load_sign * unit * (double)(long long)rho[i];
    double b;
    if (k == 0) b = load + f.lift[i];
    else if (k == 1) b = f.dscale[i] * f.dval[i];
    else {
      b = 0.0;
      cl += load;
    }
    f.b64[i] = b;
    bb += b * b;
  }
  publish(bb, slot(f, S_BB, G), lds);
  publish(cl, slot(f, S_CL, G), lds);
}

__global__ void __launch_bounds__(kPB) k_fin_rhs(FieldBufs f, int G) {
  __shared__ double lds[kPB / 64];
  const double bb = sum_partials(slot(f, S_BB, G), G, lds);
  const double cl = sum_partials(slot(f, S_CL, G), G, lds);
  if (threadIdx.x == 0) {
    PcgScalars* sc = f.sc;
    sc->bnorm2 = bb;
    sc->cond_load = cl;
    sc->outer_done = 0;
    sc->inner_done = 0;
    sc->outer_it = 0;
    sc->total_it = 0;
    sc->failed = 0;
  }
}

__global__ void __launch_bounds__(kPB) k_resid64(EllDev A, FieldBufs f) {
  if (f.sc->outer_done) return;
  __shared__ double lds[kPB / 64];
  const int G = gridDim.x;
  double rr = 0.0;
  const int64_t r0 = (int64_t)blockIdx.x * kRowsPerBlock;
  for (int64_t i = r0 + threadIdx.x; i < r0 + kRowsPerBlock && i < A.n; i += kPB) {
    double s = 0.0;
    for (int k = 0; k < A.w; ++k) {
      const int64_t o = (int64_t)k * A.n + i;
      s = fma(A.v64[o], f.x64[A.col[o]], s);
    }
    const double r = f.b64[i] - s;
    f.r64[i] = r;
    rr += r * r;
  }
  publish(rr, slot(f, S_RR, G), lds);
}

__global__ void __launch_bounds__(kPB) k_fin_resid(FieldBufs f, int G) {
  if (f.sc->outer_done) return;
  __shared__ double lds[kPB / 64];
  const double rr = sum_partials(slot(f, S_RR, G), G, lds);
  if (threadIdx.x == 0) {
    PcgScalars* sc = f.sc;
    const double rn = sqrt(rr);
    const double bn = sqrt(sc->bnorm2);
    sc->rnorm = rn;
    sc->inv_rnorm = rn > 0.0 ? 1.0 / rn : 0.0;
    const bool ok = rn <= sc->tol * (bn > 0.0 ? bn : 1.0) || rn == 0.0;
    const bool give_up = sc->outer_it >= sc->max_outer || sc->total_it >= sc->max_it;
    if (ok || give_up) {
      sc->outer_done = 1;
      sc->failed = ok ? 0 : 1;
    }
    sc->inner_done = sc->outer_done;
    sc->inner_it = 0;
    // inner target: the configured floor, or just enough to meet the outer goal
    const double need = 0.3 * sc->tol * (bn > 0.0 ? bn : 1.0) / (rn > 0.0 ? rn : 1.0);
    sc->inner_tol_eff = need > sc->inner_tol ? need : sc->inner_tol;
  }
}

template <typename T>
__global__ void __launch_bounds__(kPB) k_inner_init(EllDev A, FieldBufs f, InnerBufs<T> ib) {
  if (f.sc->outer_done) return;
  __shared__ double lds[kPB / 64];
  const int G = gridDim.x;
  const double inv = f.sc->inv_rnorm;
  double rz = 0.0, rr = 0.0;
  const int64_t r0 = (int64_t)blockIdx.x * kRowsPerBlock;
  for (int64_t i = r0 + threadIdx.x; i < r0 + kRowsPerBlock && i < A.n; i += kPB) {
    const T r = (T)(f.r64[i] * inv);
    const T z = dinv<T>(A)[i] * r;
    ib.r[i] = r;
    ib.z[i] = z;
    ib.d[i] = T(0);
    ib.p0[i] = T(0);
    ib.p1[i] = T(0);
    rz += (double)r * (double)z;
    rr += (double)r * (double)r;
  }
  publish(rz, slot(f, S_RZ0, G), lds);
  publish(rr, slot(f, S_RR0, G), lds);
}

// Scalars of inner iteration `it`, identical in every block: rz_it, rr_it
// (from the previous update or the init), rz_{it-1} for beta, and the stop
// decision.
struct IterScalars {
  double rz, rz_prev, rr;
  bool done;
};

__device__ inline IterScalars iter_scalars(const FieldBufs& f, int it, int G, double* lds,
                                           bool need_prev) {
  IterScalars s;
  const int cur = it & 1;
  s.rz = sum_partials(slot(f, cur ? S_RZ1 : S_RZ0, G), G, lds);
  s.rr = sum_partials(slot(f, cur ? S_RR1 : S_RR0, G), G, lds);
  s.rz_prev = (need_prev && it > 0) ? sum_partials(slot(f, cur ? S_RZ0 : S_RZ1, G), G, lds) : 0.0;
  const PcgScalars* sc = f.sc;
  const double tol = sc->inner_tol_eff;
  s.done = sc->outer_done || (it > 0 && s.rr <= tol * tol) || (sc->total_it + it >= sc->max_it) ||
           !(s.rz == s.rz);
  return s;
}

// q = A p_new with p_new = z + beta p_old formed on the fly; pq published.
// W > 0: compile-time ELL width, fully unrolled; the 2W coalesced slot loads
// are issued before the scalar prologue so its latency hides behind them.
// W = 0: runtime width.
template <typename T, int W>
__global__ void __launch_bounds__(kPB) k_spmv(EllDev A, FieldBufs f, const T* __restrict__ z,
                                               const T* __restrict__ pold, T* __restrict__ pnew,
                                               T* __restrict__ q, int it) {
  __shared__ double lds[kPB / 64];
  const int G = gridDim.x;
  const T* __restrict__ v = vals<T>(A);
  const int32_t* __restrict__ col = A.col;
  const int64_t n = A.n;
  const int64_t r0 = (int64_t)blockIdx.x * kRowsPerBlock + threadIdx.x;
  constexpr int WW = W > 0 ? W : 1;
  int32_t cj[2][WW];
  T vj[2][WW];
  if constexpr (W > 0) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int64_t i = r0 + h * kPB;
      const bool ok = i < n;
#pragma unroll
      for (int k = 0; k < W; ++k) {
        cj[h][k] = ok ? col[(int64_t)k * n + i] : 0;
        vj[h][k] = ok ? v[(int64_t)k * n + i] : T(0);
      }
    }
  }
  const IterScalars s = iter_scalars(f, it, G, lds, true);
  if (s.done) {
    if (blockIdx.x == 0 && threadIdx.x == 0) f.sc->inner_done = 1;
    return;
  }
  const T beta = it > 0 && s.rz_prev != 0.0 ? (T)(s.rz / s.rz_prev) : T(0);
  double pq = 0.0;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int64_t i = r0 + h * kPB;
    if (i >= n) continue;
    T acc = T(0);
    if constexpr (W > 0) {
#pragma unroll
      for (int k = 0; k < W; ++k) acc = fma_t(vj[h][k], fma_t(beta, pold[cj[h][k]], z[cj[h][k]]), acc);
    } else {
      for (int k = 0; k < A.w; ++k) {
        const int64_t o = (int64_t)k * n + i;
        const int j = col[o];
        acc = fma_t(v[o], fma_t(beta, pold[j], z[j]), acc);
      }
    }
    const T pi = fma_t(beta, pold[i], z[i]);
    pnew[i] = pi;
    q[i] = acc;
    pq += (double)pi * (double)acc;
  }
  publish(pq, slot(f, S_PQ, G), lds);
}

template <typename T>
__global__ void __launch_bounds__(kPB) k_update(EllDev A, FieldBufs f, InnerBufs<T> ib,
                                                 const T* __restrict__ pnew, int it) {
  __shared__ double lds[kPB / 64];
  const int G = gridDim.x;
  const int64_t r0 = (int64_t)blockIdx.x * kRowsPerBlock + threadIdx.x;
  T pn[2], dd[2], rv[2], qv[2], di[2];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int64_t i = r0 + h * kPB;
    const bool ok = i < A.n;
    pn[h] = ok ? pnew[i] : T(0);
    dd[h] = ok ? ib.d[i] : T(0);
    rv[h] = ok ? ib.r[i] : T(0);
    qv[h] = ok ? ib.q[i] : T(0);
    di[h] = ok ? dinv<T>(A)[i] : T(0);
  }
  // update(it) overwrites the other rz/rr parity, so it must not read it
  const IterScalars s = iter_scalars(f, it, G, lds, false);
  if (s.done) return;
  const double pq = sum_partials(slot(f, S_PQ, G), G, lds);
  if (!(fabs(pq) > 1e-300)) {
    if (blockIdx.x == 0 && threadIdx.x == 0) f.sc->inner_done = 1;
    return;
  }
  const T alpha = (T)(s.rz / pq);
  double rz = 0.0, rr = 0.0;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int64_t i = r0 + h * kPB;
    if (i >= A.n) continue;
    const T d = fma_t(alpha, pn[h], dd[h]);
    const T r = rv[h] - alpha * qv[h];
    const T z = di[h] * r;
    ib.d[i] = d;
    ib.r[i] = r;
    ib.z[i] = z;
    rz += (double)r * (double)z;
    rr += (double)r * (double)r;
  }
  const int nxt = (it + 1) & 1;
  publish(rz, slot(f, nxt ? S_RZ1 : S_RZ0, G), lds);
  publish(rr, slot(f, nxt ? S_RR1 : S_RR0, G), lds);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    f.sc->inner_it = it + 1;
    f.sc->rz = s.rz;
    f.sc->alpha = s.rz / pq;
  }
}

// Closes an inner solve: records its iteration count and applies
// x += |r| d (skipped when the outer loop has already converged).
template <typename T>
__global__ void __launch_bounds__(kPB) k_apply(EllDev A, FieldBufs f, InnerBufs<T> ib) {
  if (f.sc->outer_done) return;
  const double rn = f.sc->rnorm;
  const int64_t r0 = (int64_t)blockIdx.x * kRowsPerBlock;
  for (int64_t i = r0 + threadIdx.x; i < r0 + kRowsPerBlock && i < A.n; i += kPB)
    f.x64[i] = fma(rn, (double)ib.d[i], f.x64[i]);
}

__global__ void k_count_outer(FieldBufs f) {
  if (f.sc->outer_done) return;
  f.sc->outer_it += 1;
  f.sc->total_it += f.sc->inner_it;
}

// Floating potential from Gauss's law on the conductor (one block).
__global__ void __launch_bounds__(kBlock) k_float(FieldBufs f, FloatArgs fa) {
  __shared__ double lds[kBlock / 64];
  double s = 0.0;
  for (int64_t j = threadIdx.x; j < f.n_kc; j += blockDim.x) s += f.kc_val[j] * f.x64[f.kc_idx[j]];
  const double tot = block_sum(s, lds);
  if (threadIdx.x == 0) {
    PcgScalars* sc = f.sc;
    double q = fa.q_init;
    for (int g = 0; g < kMaxGroups; ++g) {
      if (!((fa.cond_mask >> g) & 1u)) continue;
      for (int sp = 0; sp < kMaxSpecies; ++sp) {
        q += (double)fa.ledger_count[g * kMaxSpecies + sp] * fa.qw[sp];
        if (fa.emit_count) q -= (double)fa.emit_count[g * kMaxSpecies + sp] * fa.qw[sp];
      }
    }
    sc->qcond = q;
    sc->vfloat = sc->ccap > 0.0 ? (q + sc->cond_load - tot) / sc->ccap : 0.0;
  }
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_phi(int64_t n, FieldBufs f, T* phiT, double* phi64) {
  const uint32_t lb = xcd_remap(blockIdx.x, gridDim.x);
  const int64_t i = (int64_t)lb * kBlock + threadIdx.x;
  if (i >= n) return;
  const double p = f.phi1 ? fma(f.sc->vfloat, f.phi1[i], f.x64[i]) : f.x64[i];
  phi64[i] = p;
  phiT[i] = (T)p;
}

// E_i = -sum_j w_ij (phi_j - phi_i): the least-squares nodal gradient of
// PoissonSolver::computeElectricFieldGradient (PoissonSolver.cpp:474-529,
// 749-787) with the 3x3 normal equations solved once per node at setup.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_efield(EllDev A, const T* __restrict__ phi,
                                                    Vec4<T>* __restrict__ E,
                                                    unsigned long long* rho_clear) {
  const uint32_t lb = xcd_remap(blockIdx.x, gridDim.x);
  const int64_t i = (int64_t)lb * kBlock + threadIdx.x;
  if (i >= A.n) return;
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
  const T pi = phi[i];
  T ex = T(0), ey = T(0), ez = T(0);
  for (int k = 1; k < A.w; ++k) {
    const int64_t o = (int64_t)k * A.n + i;
    const T d = phi[A.col[o]] - pi;
    ex = fma_t(gx[o], d, ex);
    ey = fma_t(gy[o], d, ey);
    ez = fma_t(gz[o], d, ez);
  }
  Vec4<T> e;
  e.x = -ex;
  e.y = -ey;
  e.z = -ez;
  e.w = pi;
  E[i] = e;
  if (rho_clear) rho_clear[i] = 0ull;
}

}  // namespace

static inline int grid_rows(int64_t n) { return (int)((n + kRowsPerBlock - 1) / kRowsPerBlock); }

void launch_rhs(const EllDev& A, const FieldBufs& f, const unsigned long long* rho_fx,
                double unit, int load_sign, hipStream_t s) {
  const int G = grid_rows(A.n);
  hipLaunchKernelGGL(k_rhs, dim3(G), dim3(kPB), 0, s, A, f, rho_fx, unit, load_sign);
  hipLaunchKernelGGL(k_fin_rhs, dim3(1), dim3(kPB), 0, s, f, G);
}
void launch_resid64(const EllDev& A, const FieldBufs& f, hipStream_t s) {
  const int G = grid_rows(A.n);
  hipLaunchKernelGGL(k_resid64, dim3(G), dim3(kPB), 0, s, A, f);
  hipLaunchKernelGGL(k_fin_resid, dim3(1), dim3(kPB), 0, s, f, G);
}
template <typename T>
void launch_inner_init(const EllDev& A, const FieldBufs& f, const InnerBufs<T>& ib,
                       hipStream_t s) {
  hipLaunchKernelGGL(k_inner_init<T>, dim3(grid_rows(A.n)), dim3(kPB), 0, s, A, f, ib);
}
template <typename T>
void launch_inner_iter(const EllDev& A, const FieldBufs& f, const InnerBufs<T>& ib, int it,
                       hipStream_t s, hipEvent_t* ev) {
  const T* pold = (it & 1) ? ib.p1 : ib.p0;
  T* pnew = (it & 1) ? ib.p0 : ib.p1;
  const int G = grid_rows(A.n);
  const dim3 g(G), b(kPB);
  hipEvent_t e0 = ev ? ev[0] : nullptr, e1 = ev ? ev[1] : nullptr;
  switch (A.w) {
#define SCDAT_W(w) \
  case w:          \
    hipExtLaunchKernelGGL((k_spmv<T, w>), g, b, 0, s, e0, e1, 0, A, f, ib.z, pold, pnew, ib.q, it); \
    break;
    SCDAT_W(7) SCDAT_W(8) SCDAT_W(9) SCDAT_W(10) SCDAT_W(11) SCDAT_W(12) SCDAT_W(13) SCDAT_W(14)
    SCDAT_W(15) SCDAT_W(16) SCDAT_W(17) SCDAT_W(18) SCDAT_W(19) SCDAT_W(20)
#undef SCDAT_W
    default:
      hipExtLaunchKernelGGL((k_spmv<T, 0>), g, b, 0, s, e0, e1, 0, A, f, ib.z, pold, pnew, ib.q, it);
  }
  hipExtLaunchKernelGGL(k_update<T>, dim3(G), dim3(kPB), 0, s, ev ? ev[2] : nullptr,
                        ev ? ev[3] : nullptr, 0, A, f, ib, (const T*)pnew, it);
}
template <typename T>
void launch_apply(const EllDev& A, const FieldBufs& f, const InnerBufs<T>& ib, hipStream_t s) {
  hipLaunchKernelGGL(k_apply<T>, dim3(grid_rows(A.n)), dim3(kPB), 0, s, A, f, ib);
  hipLaunchKernelGGL(k_count_outer, dim3(1), dim3(1), 0, s, f);
}
size_t partial_doubles(int64_t n) { return (size_t)S_COUNT * (size_t)grid_rows(n) + 16; }
void launch_float(const FieldBufs& f, const FloatArgs& fa, hipStream_t s) {
  hipLaunchKernelGGL(k_float, dim3(1), dim3(kBlock), 0, s, f, fa);
}
template <typename T>
void launch_phi_efield(const EllDev& A, const FieldBufs& f, T* phiT, double* phi64, Vec4<T>* E,
                       unsigned long long* rho_clear, hipStream_t s) {
  hipLaunchKernelGGL(k_phi<T>, dim3(grid_for(A.n)), dim3(kBlock), 0, s, A.n, f, phiT, phi64);
  hipLaunchKernelGGL(k_efield<T>, dim3(grid_for(A.n)), dim3(kBlock), 0, s, A, (const T*)phiT, E,
                     rho_clear);
}
template <typename T>
void launch_efield_only(const EllDev& A, const T* phiT, Vec4<T>* E, hipStream_t s) {
  hipLaunchKernelGGL(k_efield<T>, dim3(grid_for(A.n)), dim3(kBlock), 0, s, A, phiT, E,
                     (unsigned long long*)nullptr);
}

#define SCDAT_INST(T)                                                                          \
  template void launch_inner_init<T>(const EllDev&, const FieldBufs&, const InnerBufs<T>&,     \
                                     hipStream_t);                                             \
  template void launch_inner_iter<T>(const EllDev&, const FieldBufs&, const InnerBufs<T>&, int, \
                                     hipStream_t, hipEvent_t*);                                \
  template void launch_apply<T>(const EllDev&, const FieldBufs&, const InnerBufs<T>&,          \
                                hipStream_t);                                                  \
  template void launch_phi_efield<T>(const EllDev&, const FieldBufs&, T*, double*, Vec4<T>*,   \
                                     unsigned long long*, hipStream_t);                        \
  template void launch_efield_only<T>(const EllDev&, const T*, Vec4<T>*, hipStream_t);
SCDAT_INST(float)
SCDAT_INST(double)
#undef SCDAT_INST

}  // namespace scdat
