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

// Two-value block reduction published to partial[lb] and partial[g + lb];
// returns true in the last block, whose thread 0 then holds the totals.
__device__ inline bool reduce2(double a, double b, double* partial, unsigned int* ticket,
                               uint32_t lb, double* ta, double* tb) {
  __shared__ double lds[kBlock / 64];
  const double sa = block_sum(a, lds);
  const double sb = block_sum(b, lds);
  const uint32_t g = gridDim.x;
  if (threadIdx.x == 0) {
    partial[lb] = sa;
    partial[g + lb] = sb;
  }
  if (!last_block(ticket)) return false;
  double xa = 0.0, xb = 0.0;
  for (uint32_t j = threadIdx.x; j < g; j += blockDim.x) {
    xa += partial[j];
    xb += partial[g + j];
  }
  // fixed-order tree over the per-thread strided sums
  const double ra = block_sum(xa, lds);
  const double rb = block_sum(xb, lds);
  if (threadIdx.x == 0) {
    *ta = ra;
    *tb = rb;
    *ticket = 0u;
  }
  return true;
}

__global__ void __launch_bounds__(kBlock) k_rhs(EllDev A, FieldBufs f,
                                                 const unsigned long long* rho, double unit,
                                                 int load_sign) {
  const uint32_t lb = xcd_remap(blockIdx.x, gridDim.x);
  const int64_t i = (int64_t)lb * kBlock + threadIdx.x;
  double bb = 0.0, cl = 0.0;
  if (i < A.n) {
    const int k = f.kind[i];
    const double load = (double)load_sign * unit * (double)(long long)rho[i];
    double b;
    if (k == 0) b = load + f.lift[i];
    else if (k == 1) b = f.dscale[i] * f.dval[i];
    else {
      b = 0.0;
      cl = load;
    }
    f.b64[i] = b;
    bb = b * b;
  }
  double ta, tb;
  if (reduce2(bb, cl, f.partial, &f.sc->ticket[0], lb, &ta, &tb) && threadIdx.x == 0) {
    f.sc->bnorm2 = ta;
    f.sc->cond_load = tb;
    f.sc->outer_done = 0;
    f.sc->inner_done = 0;
    f.sc->outer_it = 0;
    f.sc->total_it = 0;
    f.sc->failed = 0;
  }
}

__global__ void __launch_bounds__(kBlock) k_resid64(EllDev A, FieldBufs f) {
  if (f.sc->outer_done) return;
  const uint32_t lb = xcd_remap(blockIdx.x, gridDim.x);
  const int64_t i = (int64_t)lb * kBlock + threadIdx.x;
  double rr = 0.0;
  if (i < A.n) {
    double s = 0.0;
    for (int k = 0; k < A.w; ++k) {
      const int64_t o = (int64_t)k * A.n + i;
      s = fma(A.v64[o], f.x64[A.col[o]], s);
    }
    const double r = f.b64[i] - s;
    f.r64[i] = r;
    rr = r * r;
  }
  double ta, tb;
  if (reduce2(rr, 0.0, f.partial, &f.sc->ticket[1], lb, &ta, &tb) && threadIdx.x == 0) {
    PcgScalars* sc = f.sc;
    const double rn = sqrt(ta);
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
__global__ void __launch_bounds__(kBlock) k_inner_init(EllDev A, FieldBufs f, InnerBufs<T> ib) {
  if (f.sc->outer_done) return;
  const uint32_t lb = xcd_remap(blockIdx.x, gridDim.x);
  const int64_t i = (int64_t)lb * kBlock + threadIdx.x;
  double rz = 0.0;
  if (i < A.n) {
    const T r = (T)(f.r64[i] * f.sc->inv_rnorm);
    const T z = dinv<T>(A)[i] * r;
    ib.r[i] = r;
    ib.z[i] = z;
    ib.d[i] = T(0);
    ib.p0[i] = T(0);
    ib.p1[i] = T(0);
    rz = (double)r * (double)z;
  }
  double ta, tb;
  if (reduce2(rz, 0.0, f.partial, &f.sc->ticket[2], lb, &ta, &tb) && threadIdx.x == 0) {
    f.sc->rz = ta;
    f.sc->beta = 0.0;
  }
}

// q = A p_new with p_new = z + beta p_old formed on the fly; pq reduced.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_spmv(EllDev A, FieldBufs f, const T* __restrict__ z,
                                                  const T* __restrict__ pold, T* __restrict__ pnew,
                                                  T* __restrict__ q) {
  if (f.sc->inner_done) return;
  const uint32_t lb = xcd_remap(blockIdx.x, gridDim.x);
  const int64_t i = (int64_t)lb * kBlock + threadIdx.x;
  const T beta = (T)f.sc->beta;
  double pq = 0.0;
  if (i < A.n) {
    const T* v = vals<T>(A);
    T s = T(0);
    for (int k = 0; k < A.w; ++k) {
      const int64_t o = (int64_t)k * A.n + i;
      const int j = A.col[o];
      s = fma_t(v[o], fma_t(beta, pold[j], z[j]), s);
    }
    const T pi = fma_t(beta, pold[i], z[i]);
    pnew[i] = pi;
    q[i] = s;
    pq = (double)pi * (double)s;
  }
  double ta, tb;
  if (reduce2(pq, 0.0, f.partial, &f.sc->ticket[3], lb, &ta, &tb) && threadIdx.x == 0) {
    PcgScalars* sc = f.sc;
    if (!(fabs(ta) > 1e-300)) {
      sc->inner_done = 1;
      sc->alpha = 0.0;
    } else {
      sc->pq = ta;
      sc->alpha = sc->rz / ta;
    }
  }
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_update(EllDev A, FieldBufs f, InnerBufs<T> ib,
                                                    const T* __restrict__ pnew) {
  if (f.sc->inner_done) return;
  const uint32_t lb = xcd_remap(blockIdx.x, gridDim.x);
  const int64_t i = (int64_t)lb * kBlock + threadIdx.x;
  const T alpha = (T)f.sc->alpha;
  double rz = 0.0, rr = 0.0;
  if (i < A.n) {
    const T d = fma_t(alpha, pnew[i], ib.d[i]);
    const T r = ib.r[i] - alpha * ib.q[i];
    const T z = dinv<T>(A)[i] * r;
    ib.d[i] = d;
    ib.r[i] = r;
    ib.z[i] = z;
    rz = (double)r * (double)z;
    rr = (double)r * (double)r;
  }
  double ta, tb;
  if (reduce2(rz, rr, f.partial, &f.sc->ticket[0], lb, &ta, &tb) && threadIdx.x == 0) {
    PcgScalars* sc = f.sc;
    sc->beta = sc->rz != 0.0 ? ta / sc->rz : 0.0;
    sc->rz = ta;
    sc->rr = tb;
    sc->inner_it += 1;
    sc->total_it += 1;
    if (tb <= sc->inner_tol_eff * sc->inner_tol_eff || sc->total_it >= sc->max_it) sc->inner_done = 1;
  }
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_apply(EllDev A, FieldBufs f, InnerBufs<T> ib) {
  if (f.sc->outer_done) return;
  const uint32_t lb = xcd_remap(blockIdx.x, gridDim.x);
  const int64_t i = (int64_t)lb * kBlock + threadIdx.x;
  if (i < A.n) f.x64[i] = fma(f.sc->rnorm, (double)ib.d[i], f.x64[i]);
  if (blockIdx.x == 0 && threadIdx.x == 0) f.sc->outer_it += 1;
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

void launch_rhs(const EllDev& A, const FieldBufs& f, const unsigned long long* rho_fx,
                double unit, int load_sign, hipStream_t s) {
  hipLaunchKernelGGL(k_rhs, dim3(grid_for(A.n)), dim3(kBlock), 0, s, A, f, rho_fx, unit,
                     load_sign);
}
void launch_resid64(const EllDev& A, const FieldBufs& f, hipStream_t s) {
  hipLaunchKernelGGL(k_resid64, dim3(grid_for(A.n)), dim3(kBlock), 0, s, A, f);
}
template <typename T>
void launch_inner_init(const EllDev& A, const FieldBufs& f, const InnerBufs<T>& ib,
                       hipStream_t s) {
  hipLaunchKernelGGL(k_inner_init<T>, dim3(grid_for(A.n)), dim3(kBlock), 0, s, A, f, ib);
}
template <typename T>
void launch_inner_iter(const EllDev& A, const FieldBufs& f, const InnerBufs<T>& ib, int parity,
                       hipStream_t s) {
  const T* pold = parity ? ib.p1 : ib.p0;
  T* pnew = parity ? ib.p0 : ib.p1;
  hipLaunchKernelGGL(k_spmv<T>, dim3(grid_for(A.n)), dim3(kBlock), 0, s, A, f, ib.z, pold, pnew,
                     ib.q);
  hipLaunchKernelGGL(k_update<T>, dim3(grid_for(A.n)), dim3(kBlock), 0, s, A, f, ib,
                     (const T*)pnew);
}
template <typename T>
void launch_apply(const EllDev& A, const FieldBufs& f, const InnerBufs<T>& ib, hipStream_t s) {
  hipLaunchKernelGGL(k_apply<T>, dim3(grid_for(A.n)), dim3(kBlock), 0, s, A, f, ib);
}
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
                                     hipStream_t);                                             \
  template void launch_apply<T>(const EllDev&, const FieldBufs&, const InnerBufs<T>&,          \
                                hipStream_t);                                                  \
  template void launch_phi_efield<T>(const EllDev&, const FieldBufs&, T*, double*, Vec4<T>*,   \
                                     unsigned long long*, hipStream_t);                        \
  template void launch_efield_only<T>(const EllDev&, const T*, Vec4<T>*, hipStream_t);
SCDAT_INST(float)
SCDAT_INST(double)
#undef SCDAT_INST

}  // namespace scdat
