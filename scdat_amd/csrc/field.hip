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
// order (wave shuffles, block partials summed by the next kernel), so a solve
// is bitwise reproducible.  Per iteration: one kernel (Chronopoulos-Gear CG,
// below).
#include <algorithm>

#include "device_common.hpp"
#include "field.hpp"

namespace scdat {

namespace {

template <typename T>
__device__ inline const T* vhat(const EllDev& A);
template <>
__device__ inline const float* vhat<float>(const EllDev& A) { return A.vh32; }
template <>
__device__ inline const double* vhat<double>(const EllDev& A) { return A.vh64; }
template <typename T>
__device__ inline const T* dinv(const EllDev& A);
template <>
__device__ inline const float* dinv<float>(const EllDev& A) { return A.dinv32; }
template <>
__device__ inline const double* dinv<double>(const EllDev& A) { return A.dinv64; }

template <typename T>
__device__ inline const T* dco(const EllDev& A);
template <>
__device__ inline const float* dco<float>(const EllDev& A) { return A.dco32; }
template <>
__device__ inline const double* dco<double>(const EllDev& A) { return A.dco64; }

inline uint32_t grid_for(int64_t n) { return (uint32_t)((n + kBlock - 1) / kBlock); }

// Reductions.  Every reducing kernel publishes one partial per block; the
// totals are formed either by a one-block finalize kernel or, inside the PCG
// iteration, by every block of the *next* kernel summing the same partials in
// the same fixed order.  No device-scope fence or atomic is needed (kernel
// boundaries order the memory), and every block obtains bit-identical
// scalars, so a solve is bitwise reproducible.
// Row blocks of the PCG kernels: one row per thread, PB (256 or 512) rows
// per block, chosen per context (EllDev::pb); kPB, the larger, is the
// block of the one-block finishing kernels.
constexpr int kPB = kPcgBlock;
constexpr int kRPT = 1;
enum Slot { S_BB = 0, S_CL, S_RR, S_G0, S_G1, S_D0, S_D1, S_RR0, S_RR1, S_WC, S_WC1, S_COUNT };

// slot s of the partials of a G-block grid; slots start 256-byte aligned
// Nothing to do for this solve: it has converged (or given up), or it waits
// for the host to resume a pass.
__device__ inline bool idle(const PcgScalars* s) { return s->outer_done || s->stall; }

__device__ inline double* slot(const FieldBufs& f, int s, int G) {
  return f.partial + (size_t)s * (size_t)((G + 31) & ~31);
}

// Row blocks [b0, b1) of the segment that block `blk` belongs to, and its
// scalars.  Sums over a segment run over its own blocks in the same order
// as a one-segment solve of that segment alone, so a batched scenario
// reproduces its stand-alone solve bit for bit.
struct SegRange {
  int s, b0, b1;
};
__device__ inline SegRange seg_of_block(const FieldBufs& f, int blk, int G) {
  if (!f.blk_seg) return {0, 0, G};
  const int s = f.blk_seg[blk];
  return {s, f.seg_b0[s], f.seg_b0[s + 1]};
}
// segment handled by block `blk` of a one-block-per-segment kernel
__device__ inline SegRange seg_by_index(const FieldBufs& f, int s, int G) {
  if (!f.blk_seg) return {0, 0, G};
  return {s, f.seg_b0[s], f.seg_b0[s + 1]};
}

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

template <int PB>
__global__ void __launch_bounds__(PB) k_rhs(EllDev A, FieldBufs f,
                                                 const unsigned long long* rho, double unit,
                                                 int load_sign) {
  __shared__ double lds[PB / 64];
  const int G = gridDim.x;
  if (f.seg_unit) unit = f.seg_unit[seg_of_block(f, blockIdx.x, G).s];
  double bb = 0.0, cl = 0.0;
  const int64_t r0 = (int64_t)blockIdx.x * PB;
  for (int64_t i = r0 + threadIdx.x; i < r0 + PB && i < A.n; i += PB) {
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

// Right-hand side from an FP64 nodal vector: src as deposited charge (scaled
// like k_rhs), or, homogeneous, the load on free rows and zero on
// constrained rows (no lift): the adjoint solve A y = P load(src).
template <int PB>
__global__ void __launch_bounds__(PB) k_rhs_real(EllDev A, FieldBufs f, const double* src, double scale,
                                                  int load_sign, int homogeneous) {
  __shared__ double lds[PB / 64];
  const int G = gridDim.x;
  double bb = 0.0, cl = 0.0;
  const int64_t r0 = (int64_t)blockIdx.x * PB;
  for (int64_t i = r0 + threadIdx.x; i < r0 + PB && i < A.n; i += PB) {
    const int k = f.kind[i];
    const double load = (double)load_sign * scale * src[i];
    double b;
    if (homogeneous) {
      b = k == 0 ? load : 0.0;
    } else {
      if (k == 0) b = load + f.lift[i];
      else if (k == 1) b = f.dscale[i] * f.dval[i];
      else {
        b = 0.0;
        cl += load;
      }
    }
    f.b64[i] = b;
    bb += b * b;
  }
  publish(bb, slot(f, S_BB, G), lds);
  publish(cl, slot(f, S_CL, G), lds);
}

// one block per segment
__global__ void __launch_bounds__(kPB) k_fin_rhs(FieldBufs f, int G) {
  __shared__ double lds[kPB / 64];
  const SegRange sg = seg_by_index(f, blockIdx.x, G);
  const double bb = sum_partials(slot(f, S_BB, G) + sg.b0, sg.b1 - sg.b0, lds);
  const double cl = sum_partials(slot(f, S_CL, G) + sg.b0, sg.b1 - sg.b0, lds);
  if (threadIdx.x == 0) {
    PcgScalars* sc = f.sc + sg.s;
    sc->bnorm2 = bb;
    sc->cond_load = cl;
    sc->outer_done = 0;
    sc->inner_done = 0;
    sc->outer_it = 0;
    sc->total_it = 0;
    sc->failed = 0;
    sc->stall = 0;
  }
}

template <int PB>
__global__ void __launch_bounds__(PB) k_resid64(EllDev A, FieldBufs f) {
  const int G = gridDim.x;
  if (idle(f.sc + seg_of_block(f, blockIdx.x, G).s)) return;
  __shared__ double lds[PB / 64];
  double rr = 0.0;
  const int64_t r0 = (int64_t)blockIdx.x * PB;
  for (int64_t i = r0 + threadIdx.x; i < r0 + PB && i < A.n; i += PB) {
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

// one block per segment
__global__ void __launch_bounds__(kPB) k_fin_resid(FieldBufs f, int G) {
  const SegRange sg = seg_by_index(f, blockIdx.x, G);
  if (idle(f.sc + sg.s)) return;
  __shared__ double lds[kPB / 64];
  const double rr = sum_partials(slot(f, S_RR, G) + sg.b0, sg.b1 - sg.b0, lds);
  if (threadIdx.x == 0) {
    PcgScalars* sc = f.sc + sg.s;
    const double rn = sqrt(rr);
    const double bn = sqrt(sc->bnorm2);
    sc->rnorm = rn;
    sc->inv_rnorm = rn > 0.0 ? 1.0 / rn : 0.0;
    const bool ok = rn <= sc->tol * (bn > 0.0 ? bn : 1.0) || rn == 0.0;
    if (sc->outer_it == 0) sc->init_rnorm = rn;
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

// Inner solve: Jacobi-preconditioned CG in the Chronopoulos-Gear form
// (A. T. Chronopoulos, C. W. Gear, J. Comput. Appl. Math. 25 (1989) 153-168),
// which needs one global reduction per iteration instead of two:
//   p_i = u_i + b_i p_{i-1},  s_i = w_i + b_i s_{i-1}  (s = A p),
//   x += a_i p_i,  r_{i+1} = r_i - a_i s_i,  u = D^-1 r,  w = A u,
//   g = (r,u), d = (w,u),  b_{i+1} = g_{i+1}/g_i,
//   a_{i+1} = g_{i+1} / (d_{i+1} - b_{i+1} g_{i+1} / a_i).
// It is the same Krylov iteration as the PCG of
// MultiBoundaryElectricFieldSolver.cpp:1432-1513 in exact arithmetic.  One
// kernel runs a whole iteration: each row forms r_{i+1} of its ELL neighbours
// from (r_i, w_i, s_{i-1}) with the same fused operations the owner uses, so
// w_{i+1} = (A D^-1) r_{i+1} needs no second launch.  r, w and s are
// double-buffered by iteration parity; p and x are owner-only, in place.
static_assert(kPcgBlock == kCoarseBlockRows, "the two-level aggregates are the 512-row blocks");
constexpr int kCoarseRowsPerWave = 2;  // rows of Ac^-1 per wave in the coarse kernel

// Restriction of the two-level preconditioner: block b's partial of P^T v is
// the sum of v over its free rows, published like the dot products.
__device__ inline void publish_restrict(double v, double* out, double* lds) {
  const double t = block_sum(v, lds);
  if (threadIdx.x == 0) out[blockIdx.x] = t;
}

template <typename T, int PB>
__global__ void __launch_bounds__(PB) k_inner_init(EllDev A, FieldBufs f, InnerBufs<T> ib) {
  const PcgScalars* sc = f.sc + seg_of_block(f, blockIdx.x, gridDim.x).s;
  if (idle(sc)) return;
  const double inv = sc->inv_rnorm;
  const int64_t r0 = (int64_t)blockIdx.x * PB;
  double rv = 0.0;
  for (int64_t i = r0 + threadIdx.x; i < r0 + PB && i < A.n; i += PB) {
    const T ri = (T)(f.r64[i] * inv);
    ib.r[0][i] = ri;
    ib.s[1][i] = T(0);
    ib.p[i] = T(0);
    ib.x[i] = T(0);
    if (A.nc && dco<T>(A)[i] != T(0)) rv += (double)ri;
  }
  if (A.nc) {
    __shared__ double lds[PB / 64];
    publish_restrict(rv, slot(f, S_WC, gridDim.x), lds);
  }
}

// Coarse step of the two-level preconditioner (a few blocks, after the
// kernel that published the restriction partials R = P^T v):
//   mode 0 (after k_inner_init): zr[0] = Ac^-1 R, zs[1] = 0;
//   mode 1 (after k_cg_start):   zw[0] = Ac^-1 R;
//   mode 2 (after iteration it): zs_i = zw_i + b zs_{i-1}, zr_{i+1} = zr_i -
//     a zs_i (the recurrences the iteration kernel evaluates on the fly, in
//     the same fused operations) and zw_{i+1} = Ac^-1 R.
template <typename T>
__global__ void __launch_bounds__(kPB) k_coarse(EllDev A, FieldBufs f, InnerBufs<T> ib, int mode, int it, int G) {
  const PcgScalars* sc = f.sc;
  if (idle(sc) || (mode == 2 && sc->inner_done)) return;
  __shared__ T wc[kMaxCoarse];
  const int nc = A.nc;
  const double* part = slot(f, S_WC, G);
  for (int c = threadIdx.x; c < nc; c += kPB) wc[c] = (T)part[c];
  __syncthreads();
  const int cur = it & 1, nxt = cur ^ 1;
  T a = T(0), b = T(0);
  if (mode == 2) {
    const double g = sc->hist[cur];
    a = (T)sc->hist[2 + cur];
    b = it > 0 ? (T)(g / sc->hist[1 - cur]) : T(0);
  }
  const T* aci;
  if constexpr (sizeof(T) == 4) aci = reinterpret_cast<const T*>(A.aci32);
  else aci = reinterpret_cast<const T*>(A.aci64);
  T* out = mode == 0 ? ib.zr[0] : (mode == 1 ? ib.zw[0] : ib.zw[nxt]);
  constexpr int NW = kPB / 64;
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  // kCoarseRowsPerWave rows per wave, their loads interleaved
  const int c0 = (blockIdx.x * NW + wv) * kCoarseRowsPerWave;
  T t[kCoarseRowsPerWave];
#pragma unroll
  for (int q = 0; q < kCoarseRowsPerWave; ++q) t[q] = T(0);
  for (int k = lane; k < nc; k += 64) {
    const T x = wc[k];
#pragma unroll
    for (int q = 0; q < kCoarseRowsPerWave; ++q)
      if (c0 + q < nc) t[q] = fma_t(aci[(size_t)(c0 + q) * nc + k], x, t[q]);
  }
#pragma unroll
  for (int q = 0; q < kCoarseRowsPerWave; ++q)
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) t[q] += __shfl_xor(t[q], off, 64);
  for (int q = 0; q < kCoarseRowsPerWave; ++q) {
    const int c = c0 + q;
    if (c >= nc) break;
    if (lane == 0) {
      out[c] = t[q];
      if (mode == 0) ib.zs[1][c] = T(0);
      if (mode == 2) {
        const T zs = fma_t(b, ib.zs[nxt][c], ib.zw[cur][c]);
        ib.zs[cur][c] = zs;
        ib.zr[nxt][c] = fma_t(-a, zs, ib.zr[cur][c]);
      }
    }
  }
}

template <int PB>
__device__ inline void publish3(double a, double b, double c, double* pa, double* pb, double* pc,
                                double* lds) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_xor(a, off, 64);
    b += __shfl_xor(b, off, 64);
    c += __shfl_xor(c, off, 64);
  }
  constexpr int NW = PB / 64;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  __syncthreads();
  if (lane == 0) {
    lds[w] = a;
    lds[NW + w] = b;
    lds[2 * NW + w] = c;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double x = 0.0, y = 0.0, z = 0.0;
#pragma unroll
    for (int k = 0; k < NW; ++k) {
      x += lds[k];
      y += lds[NW + k];
      z += lds[2 * NW + k];
    }
    pa[blockIdx.x] = x;
    pb[blockIdx.x] = y;
    pc[blockIdx.x] = z;
  }
}

// element at a 32-bit byte offset from a uniform base (the offset stays one
// 32-bit register: the load takes the base from scalar registers)
__device__ inline float at_byte(const float* __restrict__ base, uint32_t byte_off) {
  return *reinterpret_cast<const float*>(reinterpret_cast<const char*>(base) + byte_off);
}

template <typename T>
__device__ inline T r_next(T r, T w, T s, T a, T b) {
  return fma_t(-a, fma_t(b, s, w), r);
}

// w_0 = (A D^-1) r_0 and the first dot products.
template <typename T, int PB>
__global__ void __launch_bounds__(PB) k_cg_start(EllDev A, FieldBufs f, InnerBufs<T> ib) {
  const int G = gridDim.x;
  if (idle(f.sc + seg_of_block(f, blockIdx.x, G).s)) return;
  __shared__ double lds[3 * (PB / 64)];
  const T* __restrict__ v = vhat<T>(A);
  const T* __restrict__ r = ib.r[0];
  double g = 0.0, d = 0.0, rr = 0.0, wv = 0.0;
  const int64_t r0 = (int64_t)blockIdx.x * PB;
  for (int64_t i = r0 + threadIdx.x; i < r0 + PB && i < A.n; i += PB) {
    T acc = T(0);
    // w_0 = A u_0, u_0 = D^-1 r_0 + P zr_0: column j carries r_j + D_j zr_0[block of j]
    for (int k = 0; k < A.w; ++k) {
      const int64_t o = (int64_t)k * A.n + i;
      const int j = A.col[o];
      const T uj = A.nc ? fma_t(dco<T>(A)[j], ib.zr[0][j / kCoarseBlockRows], r[j]) : r[j];
      acc = fma_t(v[o], uj, acc);
    }
    const T ri = r[i];
    T u = dinv<T>(A)[i] * ri;
    if (A.nc && dco<T>(A)[i] != T(0)) {
      u += ib.zr[0][blockIdx.x];
      wv += (double)acc;
    }
    ib.w[0][i] = acc;
    g += (double)ri * (double)u;
    d += (double)acc * (double)u;
    rr += (double)ri * (double)ri;
  }
  publish3<PB>(g, d, rr, slot(f, S_G0, G), slot(f, S_D0, G), slot(f, S_RR0, G), lds);
  if (A.nc) publish_restrict(wv, slot(f, S_WC, G), lds);
}

// k_cg_start in the tiled order: u_j = r_j (+ D_j zr_0 of j's block,
// two-level) staged over the tile's 10^3 box in LDS, each row reading its
// neighbours by direction code from the box pattern table.  The values, the
// slot order and the fused operations are k_cg_start's, so both give the
// same bits.
template <int W, bool TL>
__global__ void __launch_bounds__(kPcgBlock) k_cg_start_box(EllDev A, FieldBufs f, InnerBufs<float> ib) {
  constexpr int PB = kPcgBlock;
  static_assert(2 * PB >= 1000, "two box points per thread");
  __shared__ double lds[3 * (PB / 64)];
  __shared__ float box[1000];  // u_j at (x, y, z) in [-1, 8]^3: (x+1) + 10 (y+1) + 100 (z+1)
  const int G = gridDim.x;
  if (idle(f.sc)) return;
  const int64_t n = A.n;
  const int bt = (int)blockIdx.x;
  const int64_t i = (int64_t)bt * PB + threadIdx.x;
  const float* __restrict__ r = ib.r[0];
  const float* __restrict__ dc = A.dco32;
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    const int q = (int)threadIdx.x + u * PB;
    if (q >= 1000) continue;
    const int bx = q % 10 - 1, by = (q / 10) % 10 - 1, bz = q / 100 - 1;
    const int code = ((bx >> 3) + 1) + 3 * ((by >> 3) + 1) + 9 * ((bz >> 3) + 1);
    const int nr = A.nbb[(size_t)bt * 27 + code];
    const int nb = nr >= 0 ? nr : bt;
    const int64_t row = (int64_t)nb * PB + (bx & 7) + 8 * (by & 7) + 64 * (bz & 7);
    float v = 0.0f;
    if ((code == 13 || nb != bt) && row < n) v = TL ? fmaf(dc[row], ib.zr[0][nb], r[row]) : r[row];
    box[q] = v;
  }
  __syncthreads();
  double g = 0.0, d = 0.0, rr = 0.0, wv = 0.0;
  if (i < n) {
    const int pj = (int)A.bpat[i];
    const int x = threadIdx.x & 7, y = (threadIdx.x >> 3) & 7, z = threadIdx.x >> 6;
    const int cpos = (x + 1) + 10 * (y + 1) + 100 * (z + 1);
    const uint4 dq = *reinterpret_cast<const uint4*>(A.bd + (size_t)pj * 16);
    const uint32_t dw[4] = {dq.x, dq.y, dq.z, dq.w};
    const float* pv = A.bv + (size_t)pj * W;
    float acc = 0.0f;
#pragma unroll 5
    for (int k = 0; k < W; ++k) {
      const int dcode = (int)((dw[k >> 2] >> (8 * (k & 3))) & 0xffu);
      const int dx = (dcode & 3) - 1, dy = ((dcode >> 2) & 3) - 1, dz = (dcode >> 4) - 1;
      acc = fmaf(pv[k], box[cpos + dx + 10 * dy + 100 * dz], acc);
    }
    const float ri = r[i];
    float u = A.dinv32[i] * ri;
    if (TL && dc[i] != 0.0f) {
      u += ib.zr[0][bt];
      wv += (double)acc;
    }
    ib.w[0][i] = acc;
    g += (double)ri * (double)u;
    d += (double)acc * (double)u;
    rr += (double)ri * (double)ri;
  }
  publish3<PB>(g, d, rr, slot(f, S_G0, G), slot(f, S_D0, G), slot(f, S_RR0, G), lds);
  if (TL) publish_restrict(wv, slot(f, S_WC, G), lds);
}

// One iteration.  W > 0: compile-time ELL width, fully unrolled, with the 2W
// coalesced slot loads issued before the scalar prologue; W = 0: runtime width.
// TL: two-level preconditioner (coarse terms from the coarse vectors of
// this pass; restriction partials of w_{i+1} for the coarse kernel).
// PAT (FP32, W > 0, no TL): the row's slots and D^-1 come from its pattern;
// the pattern id is loaded before the prologue and the table entries (hot
// in L1/L2: neighbouring rows mostly share a pattern) right after it.
// (Hoisting more loads above the prologue -- x, p, the scalars, unrolled
// partials -- raised VGPRs and took the iteration from 11.7 to 12.3 us.)
template <typename T, int W, bool TL, bool PAT, int PB>
__global__ void __launch_bounds__(PB) k_cg_iter(EllDev A, FieldBufs f, InnerBufs<T> ib, int it) {
  __shared__ double lds[3 * (PB / 64)];
  __shared__ double tot[3];
  const int G = gridDim.x;
  const int cur = it & 1, nxt = cur ^ 1;
  const T* __restrict__ v = vhat<T>(A);
  const int32_t* __restrict__ col = A.col;
  const int64_t n = A.n;
  const int64_t r0 = (int64_t)blockIdx.x * PB + threadIdx.x;
  constexpr int WW = W > 0 ? W : 1;
  int32_t cj[kRPT][WW];
  T vj[kRPT][WW];
  int pj[kRPT];
  if constexpr (PAT) {
#pragma unroll
    for (int h = 0; h < kRPT; ++h) {
      const int64_t i = r0 + h * PB;
      pj[h] = i < n ? (A.pat ? (int)A.pat[i] : (int)A.pat32[i]) : 0;
    }
  } else if constexpr (W > 0) {
#pragma unroll
    for (int h = 0; h < kRPT; ++h) {
      const int64_t i = r0 + h * PB;
      const bool ok = i < n;
#pragma unroll
      for (int k = 0; k < W; ++k) {
        cj[h][k] = ok ? col[(int64_t)k * n + i] : 0;
        vj[h][k] = ok ? v[(int64_t)k * n + i] : T(0);
      }
    }
  }
  // g_i, d_i, rr_i from the previous kernel's block partials of this segment
  const SegRange sg = seg_of_block(f, blockIdx.x, G);
  {
    const double* pg = slot(f, cur ? S_G1 : S_G0, G) + sg.b0;
    const double* pd = slot(f, cur ? S_D1 : S_D0, G) + sg.b0;
    const double* pr = slot(f, cur ? S_RR1 : S_RR0, G) + sg.b0;
    const int nb = sg.b1 - sg.b0;
    double a = 0.0, b = 0.0, c = 0.0;
    for (int j = threadIdx.x; j < nb; j += PB) {
      a += pg[j];
      b += pd[j];
      c += pr[j];
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      a += __shfl_xor(a, off, 64);
      b += __shfl_xor(b, off, 64);
      c += __shfl_xor(c, off, 64);
    }
    constexpr int NW = PB / 64;
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) {
      lds[wv] = a;
      lds[NW + wv] = b;
      lds[2 * NW + wv] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      double x = 0.0, y = 0.0, z = 0.0;
#pragma unroll
      for (int k = 0; k < NW; ++k) {
        x += lds[k];
        y += lds[NW + k];
        z += lds[2 * NW + k];
      }
      tot[0] = x;
      tot[1] = y;
      tot[2] = z;
    }
    __syncthreads();
  }
  const double g = tot[0], d = tot[1], rr = tot[2];
  if constexpr (PAT) {
#pragma unroll
    for (int h = 0; h < kRPT; ++h) {
      const int64_t i = r0 + h * PB;
      const int32_t* po = A.poff + (size_t)pj[h] * W;
      const float* pv = A.pval32 + (size_t)pj[h] * W;
#pragma unroll
      for (int k = 0; k < W; ++k) {
        cj[h][k] = (int32_t)i + po[k];
        vj[h][k] = (T)pv[k];
      }
    }
  }
  PcgScalars* sc = f.sc + sg.s;
  const double tol = sc->inner_tol_eff;
  double beta = 0.0, alpha;
  if (it > 0) {
    beta = g / sc->hist[1 - cur];
    alpha = g / (d - beta * g / sc->hist[2 + (1 - cur)]);
  } else {
    alpha = g / d;
  }
  // inner_done: an earlier launch of this batch already stopped the solve;
  // the partials of this parity are then stale and must not be used
  if (idle(sc)) return;  // converged, or a stalled solve waits for the host
  const bool done = sc->inner_done || (it > 0 && rr <= tol * tol) ||
                    (sc->total_it + it >= sc->max_it) || !(alpha > 0.0) || !(alpha < 1e300);
  if (done) {
    if ((int)blockIdx.x == sg.b0 && threadIdx.x == 0) sc->inner_done = 1;
    return;
  }
  const T a = (T)alpha, b = (T)beta;
  const T* __restrict__ rc = ib.r[cur];
  const T* __restrict__ wc = ib.w[cur];
  const T* __restrict__ sp = ib.s[nxt];
  double g1 = 0.0, d1 = 0.0, rr1 = 0.0, wrest = 0.0;
  // zr_{i+1}[c] from this pass's coarse vectors (same operations as k_coarse)
  auto zr1 = [&](int c) { return fma_t(-a, fma_t(b, ib.zs[nxt][c], ib.zw[cur][c]), ib.zr[cur][c]); };
#pragma unroll
  for (int h = 0; h < kRPT; ++h) {
    const int64_t i = r0 + h * PB;
    if (i >= n) continue;
    T acc = T(0);
    // two-level: column j carries D_j u_j = r_j + D_j zr_{i+1}[block of j]
    if constexpr (W > 0) {
#pragma unroll
      for (int k = 0; k < W; ++k) {
        const int j = cj[h][k];
        T uj = r_next(rc[j], wc[j], sp[j], a, b);
        if constexpr (TL) uj = fma_t(dco<T>(A)[j], zr1(j / kCoarseBlockRows), uj);
        acc = fma_t(vj[h][k], uj, acc);
      }
    } else {
      for (int k = 0; k < A.w; ++k) {
        const int64_t o = (int64_t)k * n + i;
        const int j = col[o];
        T uj = r_next(rc[j], wc[j], sp[j], a, b);
        if constexpr (TL) uj = fma_t(dco<T>(A)[j], zr1(j / kCoarseBlockRows), uj);
        acc = fma_t(v[o], uj, acc);
      }
    }
    const T ri = rc[i], wi = wc[i], so = sp[i];
    const T si = fma_t(b, so, wi);
    const T rn = fma_t(-a, si, ri);
    T di;
    if constexpr (PAT) di = (T)A.pdinv32[pj[h]];
    else di = dinv<T>(A)[i];
    T ui = di * ri, uc = T(0);
    if constexpr (TL) {
      // u = D^-1 r + P zr on free rows (this row's aggregate is its block)
      if (dco<T>(A)[i] != T(0)) {
        ui += ib.zr[cur][blockIdx.x];
        uc = zr1(blockIdx.x);
        wrest = (double)acc;
      }
    }
    const T pi = fma_t(b, ib.p[i], ui);
    ib.x[i] = fma_t(a, pi, ib.x[i]);
    ib.p[i] = pi;
    ib.s[cur][i] = si;
    ib.r[nxt][i] = rn;
    ib.w[nxt][i] = acc;
    const T un = di * rn + uc;
    g1 += (double)rn * (double)un;
    d1 += (double)acc * (double)un;
    rr1 += (double)rn * (double)rn;
  }
  publish3<PB>(g1, d1, rr1, slot(f, nxt ? S_G1 : S_G0, G), slot(f, nxt ? S_D1 : S_D0, G),
           slot(f, nxt ? S_RR1 : S_RR0, G), lds);
  if ((int)blockIdx.x == sg.b0 && threadIdx.x == 0) {
    sc->hist[cur] = g;
    sc->hist[2 + cur] = alpha;
    sc->inner_it = it + 1;
  }
  if constexpr (TL) publish_restrict(wrest, slot(f, S_WC, G), lds);
}

// One iteration with the two-level preconditioner and the coarse solve
// fused into it (FP32, row patterns, 512-row blocks = the aggregates).
// The restriction R = P^T w_i of the previous kernel arrives with the dot
// products as block partials (by parity).  Each block applies the rows of
// Ac^-1 of the aggregates it touches (zw_i = Ac^-1 R, the same operations in
// every block, so all blocks hold the same bits), advances zs_i = zw_i +
// b zs_{i-1} and zr_{i+1} = zr_i - a zs_i for them, and the owner stores its
// own.  The row then forms
//   w_{i+1} = A u_{i+1} = (A D^-1) r_{i+1} + (A P) zr_{i+1}
// from its pattern (column offsets, A D^-1, A P), with r_{i+1} of the
// neighbours formed on the fly as in the Jacobi iteration.  One launch per
// iteration, like Jacobi: no coarse kernel between iterations.
template <int W>
__global__ void __launch_bounds__(kPcgBlock, 6) k_cg_iter_tl(EllDev A, FieldBufs f, InnerBufs<float> ib, int it) {
  constexpr int PB = kPcgBlock;
  constexpr int NW = PB / 64;
  __shared__ double lds[4 * NW];
  __shared__ double tot[3];
  __shared__ __align__(16) float rl[kMaxCoarse];  // P^T w_i, zero past nc
  __shared__ float zl[kMaxCoarse];  // zr_{i+1} of the aggregates this block touches, else 0
  __shared__ int cl[kMaxTouched];   // the touched aggregates, own first
  __shared__ float zown;            // zr_i of the block's own aggregate
  const int G = gridDim.x;
  const int nc = A.nc, ncp = A.ncp;
  const int cur = it & 1, nxt = cur ^ 1;
  const int64_t n = A.n;
  const int bt = (int)blockIdx.x;
  const int64_t i = (int64_t)bt * PB + threadIdx.x;
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int pj = i < n ? (A.pat ? (int)A.pat[i] : (int)A.pat32[i]) : 0;
  const int q0 = A.cnb_off[bt], K = A.cnb_off[bt + 1] - q0;
  // the solve's scalars, read together up front (block 0 writes the other
  // parity's history at the end; the flags are set by earlier launches)
  PcgScalars* sc = f.sc;
  const bool idle_now = idle(sc);  // converged, or a stalled solve waits for the host
  const double g_prev = sc->hist[1 - cur], a_prev = sc->hist[2 + (1 - cur)];
  const double tol = sc->inner_tol_eff;
  const bool stopped = sc->inner_done || sc->total_it + it >= sc->max_it;
  // a launch past the end of this pass (or of the solve) returns before its
  // loads: it would only sum the partials to find that it has nothing to do
  if (idle_now) return;
  if (stopped) {
    if (blockIdx.x == 0 && threadIdx.x == 0) sc->inner_done = 1;
    return;
  }
  const int myc = (int)threadIdx.x < K ? A.cnb[q0 + threadIdx.x] : 0;  // in flight with the partials
  {
    // partials two per thread (16-byte loads; slots are 256-byte aligned, so
    // the pair past G stays inside the slot), summed in index order; R = P^T
    // w_i to LDS in FP32, zero up to a multiple of 512 columns (what the
    // coarse rows' loop reads)
    const double* pg = slot(f, cur ? S_G1 : S_G0, G);
    const double* pd = slot(f, cur ? S_D1 : S_D0, G);
    const double* pr = slot(f, cur ? S_RR1 : S_RR0, G);
    const double* pw = slot(f, cur ? S_WC1 : S_WC, G);
    const int r512 = (nc + 511) & ~511;
    double a = 0.0, b = 0.0, c = 0.0;
    for (int j = 2 * (int)threadIdx.x; j < r512; j += 2 * PB) {
      if (j < G) {
        const double2 xg = *reinterpret_cast<const double2*>(pg + j);
        const double2 xd = *reinterpret_cast<const double2*>(pd + j);
        const double2 xr = *reinterpret_cast<const double2*>(pr + j);
        const double2 xw = *reinterpret_cast<const double2*>(pw + j);
        const bool two = j + 1 < G;
        a += xg.x;
        b += xd.x;
        c += xr.x;
        if (two) {
          a += xg.y;
          b += xd.y;
          c += xr.y;
        }
        rl[j] = (float)xw.x;
        rl[j + 1] = two ? (float)xw.y : 0.0f;
        zl[j] = 0.0f;
        zl[j + 1] = 0.0f;
      } else {
        rl[j] = 0.0f;
        rl[j + 1] = 0.0f;
      }
    }
    if ((int)threadIdx.x < K) cl[threadIdx.x] = myc;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      a += __shfl_xor(a, off, 64);
      b += __shfl_xor(b, off, 64);
      c += __shfl_xor(c, off, 64);
    }
    if (lane == 0) {
      lds[wv] = a;
      lds[NW + wv] = b;
      lds[2 * NW + wv] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      double x = 0.0, y = 0.0, z = 0.0;
#pragma unroll
      for (int k = 0; k < NW; ++k) {
        x += lds[k];
        y += lds[NW + k];
        z += lds[2 * NW + k];
      }
      tot[0] = x;
      tot[1] = y;
      tot[2] = z;
    }
    __syncthreads();
  }
  const double g = tot[0], d = tot[1], rr = tot[2];
  double beta = 0.0, alpha;
  if (it > 0) {
    beta = g / g_prev;
    alpha = g / (d - beta * g / a_prev);
  } else {
    alpha = g / d;
  }
  if (idle_now) return;
  const bool done = stopped || (it > 0 && rr <= tol * tol) || !(alpha > 0.0) || !(alpha < 1e300);
  if (done) {
    if (blockIdx.x == 0 && threadIdx.x == 0) sc->inner_done = 1;
    return;
  }
  const float a = (float)alpha, b = (float)beta;
  // zw_i = Ac^-1 R on the touched aggregates, then zs_i and zr_{i+1}: wave wv
  // evaluates list entries wv + 2 NW m and wv + NW + 2 NW m, 16-byte loads
  // of both rows (zero-padded to ncp columns; R is zero past nc)
  for (int base = 0; base < K; base += 2 * NW) {
    const int qa = base + wv, qb = qa + NW;
    if (qa >= K) break;  // uniform in the wave
    const bool hb = qb < K;
    const int ca = cl[qa], cb = cl[hb ? qb : qa];
    const float4* ra = reinterpret_cast<const float4*>(A.acip32 + (size_t)ca * ncp);
    const float4* rb = reinterpret_cast<const float4*>(A.acip32 + (size_t)cb * ncp);
    const float zra = ib.zr[cur][ca], zsa = ib.zs[nxt][ca];
    const float zrb = ib.zr[cur][cb], zsb = ib.zs[nxt][cb];
    float ta = 0.0f, tb = 0.0f;
    for (int k0 = 0; k0 < ncp; k0 += 2 * 256) {
      float4 xa[2], xb[2], xr[2];
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const int k = k0 + 256 * u + 4 * lane;  // first column of this lane's four
        const int ks = k < ncp ? k / 4 : 0;     // past the row: any entry, times R = 0
        xa[u] = ra[ks];
        xb[u] = rb[ks];
        xr[u] = *reinterpret_cast<const float4*>(&rl[k]);  // k < nc rounded up to 512
      }
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        ta = fmaf(xa[u].x, xr[u].x, ta);
        ta = fmaf(xa[u].y, xr[u].y, ta);
        ta = fmaf(xa[u].z, xr[u].z, ta);
        ta = fmaf(xa[u].w, xr[u].w, ta);
        tb = fmaf(xb[u].x, xr[u].x, tb);
        tb = fmaf(xb[u].y, xr[u].y, tb);
        tb = fmaf(xb[u].z, xr[u].z, tb);
        tb = fmaf(xb[u].w, xr[u].w, tb);
      }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      ta += __shfl_xor(ta, off, 64);
      tb += __shfl_xor(tb, off, 64);
    }
    if (lane == 0) {
      const float sa = fmaf(b, zsa, ta), za = fmaf(-a, sa, zra);
      zl[ca] = za;
      if (ca == bt) {
        ib.zs[cur][ca] = sa;
        ib.zr[nxt][ca] = za;
        zown = zra;
      }
      if (hb) {
        const float sb = fmaf(b, zsb, tb), zb = fmaf(-a, sb, zrb);
        zl[cb] = zb;
        if (cb == bt) {
          ib.zs[cur][cb] = sb;
          ib.zr[nxt][cb] = zb;
          zown = zrb;
        }
      }
    }
  }
  const float* __restrict__ rc = ib.r[cur];
  const float* __restrict__ wc = ib.w[cur];
  const float* __restrict__ sp = ib.s[nxt];
  const int32_t* po = A.poff + (size_t)pj * W;
  // (A D^-1) r_{i+1}, r_{i+1} of the neighbours formed on the fly (32-bit
  // byte offsets from the uniform bases: one address register per neighbour)
  float acc = 0.0f, ri = 0.0f, wi = 0.0f, so = 0.0f, pold = 0.0f, xold = 0.0f, di = 0.0f;
  if (i < n) {
    const float* pv = A.pval32 + (size_t)pj * W;
#pragma unroll
    for (int k = 0; k < W; ++k) {
      const uint32_t jb = (uint32_t)((int)i + po[k]) * 4u;
      acc = fmaf(pv[k], r_next(at_byte(rc, jb), at_byte(wc, jb), at_byte(sp, jb), a, b), acc);
    }
    // the row's own entries, in flight across the barrier
    ri = rc[i];
    wi = wc[i];
    so = sp[i];
    pold = ib.p[i];
    xold = ib.x[i];
    di = A.pdinv32[pj];
  }
  __syncthreads();  // zl
  double g1 = 0.0, d1 = 0.0, rr1 = 0.0, wrest = 0.0;
  if (i < n) {
    // + (A P) zr_{i+1}: the table entries again (hot in L1)
    const float* pa = A.pap32 + (size_t)pj * W;
    float acz = 0.0f;
#pragma unroll
    for (int k = 0; k < W; ++k) acz = fmaf(pa[k], zl[((int)i + po[k]) / kCoarseBlockRows], acz);
    acc += acz;
    const bool free_row = pa[0] != 0.0f;  // slot 0 is self: A_ii on free rows
    const float si = fmaf(b, so, wi);
    const float rn = fmaf(-a, si, ri);
    const float ui = free_row ? fmaf(di, ri, zown) : di * ri;
    const float uc = free_row ? zl[bt] : 0.0f;
    const float pi = fmaf(b, pold, ui);
    ib.x[i] = fmaf(a, pi, xold);
    ib.p[i] = pi;
    ib.s[cur][i] = si;
    ib.r[nxt][i] = rn;
    ib.w[nxt][i] = acc;
    const float un = fmaf(di, rn, uc);
    g1 = (double)rn * (double)un;
    d1 = (double)acc * (double)un;
    rr1 = (double)rn * (double)rn;
    if (free_row) wrest = (double)acc;
  }
  // block partials of the dots and of P^T w_{i+1}
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    g1 += __shfl_xor(g1, off, 64);
    d1 += __shfl_xor(d1, off, 64);
    rr1 += __shfl_xor(rr1, off, 64);
    wrest += __shfl_xor(wrest, off, 64);
  }
  __syncthreads();
  if (lane == 0) {
    lds[wv] = g1;
    lds[NW + wv] = d1;
    lds[2 * NW + wv] = rr1;
    lds[3 * NW + wv] = wrest;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double x = 0.0, y = 0.0, z = 0.0, q = 0.0;
#pragma unroll
    for (int k = 0; k < NW; ++k) {
      x += lds[k];
      y += lds[NW + k];
      z += lds[2 * NW + k];
      q += lds[3 * NW + k];
    }
    slot(f, nxt ? S_G1 : S_G0, G)[bt] = x;
    slot(f, nxt ? S_D1 : S_D0, G)[bt] = y;
    slot(f, nxt ? S_RR1 : S_RR0, G)[bt] = z;
    slot(f, nxt ? S_WC1 : S_WC, G)[bt] = q;
    if (blockIdx.x == 0) {
      sc->hist[cur] = g;
      sc->hist[2 + cur] = alpha;
      sc->inner_it = it + 1;
    }
  }
}

// One iteration in the tiled order with the neighbours' r_{i+1} staged in
// LDS (Jacobi, or TL: two-level with the coarse step as in k_cg_iter_tl).
// Block b stages r_{i+1} = r - a (w + b s) once for every point of the 10^3
// box around its tile (its tile and the one-point halo from the 26
// neighbour tiles; points with no node hold 0), then each row reads its
// neighbours from the box by direction code.  Per row this replaces W
// three-vector gathers by about two staged loads; the values, their order
// and the fused operations are those of the row-pattern iteration, so the
// two give the same bits.
template <int W, bool TL>
__global__ void __launch_bounds__(kPcgBlock, 6) k_cg_iter_box(EllDev A, FieldBufs f, InnerBufs<float> ib, int it) {
  constexpr int PB = kPcgBlock;
  constexpr int NW = PB / 64;
  constexpr int NC = TL ? kMaxCoarse : 1;
  __shared__ double lds[4 * NW];
  __shared__ double tot[3];
  __shared__ __align__(16) float rl[NC];  // P^T w_i, zero past nc
  __shared__ float zl[NC];                // zr_{i+1} of the aggregates this block touches, else 0
  __shared__ int cl[kMaxTouched];
  __shared__ float zown;
  __shared__ float box[1000];             // r_{i+1} at (x, y, z) in [-1, 8]^3: (x+1) + 10 (y+1) + 100 (z+1)
  __shared__ int nbl[27];                 // neighbour blocks; missing ones: this block (their entries are 0)
  const int G = gridDim.x;
  const int cur = it & 1, nxt = cur ^ 1;
  const int64_t n = A.n;
  const int bt = (int)blockIdx.x;
  const int64_t i = (int64_t)bt * PB + threadIdx.x;
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int pj = i < n ? (int)A.bpat[i] : 0;
  const int mynb = threadIdx.x < 27 ? A.nbb[(size_t)bt * 27 + threadIdx.x] : 0;
  PcgScalars* sc = f.sc;
  const bool idle_now = idle(sc);  // converged, or a stalled solve waits for the host
  const double g_prev = sc->hist[1 - cur], a_prev = sc->hist[2 + (1 - cur)];
  const double tol = sc->inner_tol_eff;
  const bool stopped = sc->inner_done || sc->total_it + it >= sc->max_it;
  // a launch past the end of this pass (or of the solve) returns before its
  // loads: it would only sum the partials to find that it has nothing to do
  if (idle_now) return;
  if (stopped) {
    if (blockIdx.x == 0 && threadIdx.x == 0) sc->inner_done = 1;
    return;
  }
  int K = 0;
  if constexpr (TL) {
    const int nc = A.nc;
    const int q0 = A.cnb_off[bt];
    K = A.cnb_off[bt + 1] - q0;
    const int myc = (int)threadIdx.x < K ? A.cnb[q0 + threadIdx.x] : 0;
    // as k_cg_iter_tl: partials two per thread, R to LDS in FP32
    const double* pg = slot(f, cur ? S_G1 : S_G0, G);
    const double* pd = slot(f, cur ? S_D1 : S_D0, G);
    const double* pr = slot(f, cur ? S_RR1 : S_RR0, G);
    const double* pw = slot(f, cur ? S_WC1 : S_WC, G);
    const int r512 = (nc + 511) & ~511;
    double a = 0.0, b = 0.0, c = 0.0;
    for (int j = 2 * (int)threadIdx.x; j < r512; j += 2 * PB) {
      if (j < G) {
        const double2 xg = *reinterpret_cast<const double2*>(pg + j);
        const double2 xd = *reinterpret_cast<const double2*>(pd + j);
        const double2 xr = *reinterpret_cast<const double2*>(pr + j);
        const double2 xw = *reinterpret_cast<const double2*>(pw + j);
        const bool two = j + 1 < G;
        a += xg.x;
        b += xd.x;
        c += xr.x;
        if (two) {
          a += xg.y;
          b += xd.y;
          c += xr.y;
        }
        rl[j] = (float)xw.x;
        rl[j + 1] = two ? (float)xw.y : 0.0f;
        zl[j] = 0.0f;
        zl[j + 1] = 0.0f;
      } else {
        rl[j] = 0.0f;
        rl[j + 1] = 0.0f;
      }
    }
    if ((int)threadIdx.x < K) cl[threadIdx.x] = myc;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      a += __shfl_xor(a, off, 64);
      b += __shfl_xor(b, off, 64);
      c += __shfl_xor(c, off, 64);
    }
    if (lane == 0) {
      lds[wv] = a;
      lds[NW + wv] = b;
      lds[2 * NW + wv] = c;
    }
  } else {
    // as k_cg_iter: partials in block order, one per thread and pass
    const double* pg = slot(f, cur ? S_G1 : S_G0, G);
    const double* pd = slot(f, cur ? S_D1 : S_D0, G);
    const double* pr = slot(f, cur ? S_RR1 : S_RR0, G);
    double a = 0.0, b = 0.0, c = 0.0;
    for (int j = threadIdx.x; j < G; j += PB) {
      a += pg[j];
      b += pd[j];
      c += pr[j];
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      a += __shfl_xor(a, off, 64);
      b += __shfl_xor(b, off, 64);
      c += __shfl_xor(c, off, 64);
    }
    if (lane == 0) {
      lds[wv] = a;
      lds[NW + wv] = b;
      lds[2 * NW + wv] = c;
    }
  }
  static_assert(2 * PB >= 1000, "two box points per thread");
  // this iteration's operands do not depend on the sums: r, w, s at this
  // thread's two box points and the row's own entries are loaded now, so
  // that they arrive while the sums, the step lengths and the coarse step
  // are formed (the neighbour block of a point from the global table;
  // unconditional loads from clamped rows, since a load under a branch is
  // waited for where the branch joins)
  const float* __restrict__ rc = ib.r[cur];
  const float* __restrict__ wc = ib.w[cur];
  const float* __restrict__ sp = ib.s[nxt];
  int nr[2];
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    const int q = min((int)threadIdx.x + u * PB, 999);
    const int bx = q % 10 - 1, by = (q / 10) % 10 - 1, bz = q / 100 - 1;
    nr[u] = A.nbb[(size_t)bt * 27 + ((bx >> 3) + 1) + 3 * ((by >> 3) + 1) + 9 * ((bz >> 3) + 1)];
  }
  const int64_t is = i < n ? i : n - 1;
  const float ri = rc[is], wi = wc[is], so = sp[is], pold = ib.p[is], xold = ib.x[is];
  float br[2], bw[2], bs[2];
  bool bv[2];
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    const int q = (int)threadIdx.x + u * PB;
    const int bx = q % 10 - 1, by = (q / 10) % 10 - 1, bz = q / 100 - 1;
    const int code = ((bx >> 3) + 1) + 3 * ((by >> 3) + 1) + 9 * ((bz >> 3) + 1);
    const int nb = nr[u] >= 0 ? nr[u] : bt;
    const int64_t row = (int64_t)nb * PB + (bx & 7) + 8 * (by & 7) + 64 * (bz & 7);
    bv[u] = q < 1000 && (code == 13 || nb != bt) && row < n;
    const int64_t rs = bv[u] ? row : is;
    br[u] = rc[rs];
    bw[u] = wc[rs];
    bs[u] = sp[rs];
  }
  if (threadIdx.x < 27) nbl[threadIdx.x] = mynb >= 0 ? mynb : bt;
  __syncthreads();
  if (threadIdx.x == 0) {
    double x = 0.0, y = 0.0, z = 0.0;
#pragma unroll
    for (int k = 0; k < NW; ++k) {
      x += lds[k];
      y += lds[NW + k];
      z += lds[2 * NW + k];
    }
    tot[0] = x;
    tot[1] = y;
    tot[2] = z;
  }
  __syncthreads();
  const double g = tot[0], d = tot[1], rr = tot[2];
  double beta = 0.0, alpha;
  if (it > 0) {
    beta = g / g_prev;
    alpha = g / (d - beta * g / a_prev);
  } else {
    alpha = g / d;
  }
  if (idle_now) return;
  const bool done = stopped || (it > 0 && rr <= tol * tol) || !(alpha > 0.0) || !(alpha < 1e300);
  if (done) {
    if (blockIdx.x == 0 && threadIdx.x == 0) sc->inner_done = 1;
    return;
  }
  const float a = (float)alpha, b = (float)beta;
  if constexpr (TL) {
    const int ncp = A.ncp;
    for (int base = 0; base < K; base += 2 * NW) {
      const int qa = base + wv, qb = qa + NW;
      if (qa >= K) break;  // uniform in the wave
      const bool hb = qb < K;
      const int ca = cl[qa], cb = cl[hb ? qb : qa];
      const float4* ra = reinterpret_cast<const float4*>(A.acip32 + (size_t)ca * ncp);
      const float4* rb = reinterpret_cast<const float4*>(A.acip32 + (size_t)cb * ncp);
      const float zra = ib.zr[cur][ca], zsa = ib.zs[nxt][ca];
      const float zrb = ib.zr[cur][cb], zsb = ib.zs[nxt][cb];
      float ta = 0.0f, tb = 0.0f;
      for (int k0 = 0; k0 < ncp; k0 += 2 * 256) {
        float4 xa[2], xb[2], xr[2];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
          const int k = k0 + 256 * u + 4 * lane;
          const int ks = k < ncp ? k / 4 : 0;
          xa[u] = ra[ks];
          xb[u] = rb[ks];
          xr[u] = *reinterpret_cast<const float4*>(&rl[k]);
        }
#pragma unroll
        for (int u = 0; u < 2; ++u) {
          ta = fmaf(xa[u].x, xr[u].x, ta);
          ta = fmaf(xa[u].y, xr[u].y, ta);
          ta = fmaf(xa[u].z, xr[u].z, ta);
          ta = fmaf(xa[u].w, xr[u].w, ta);
          tb = fmaf(xb[u].x, xr[u].x, tb);
          tb = fmaf(xb[u].y, xr[u].y, tb);
          tb = fmaf(xb[u].z, xr[u].z, tb);
          tb = fmaf(xb[u].w, xr[u].w, tb);
        }
      }
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
        ta += __shfl_xor(ta, off, 64);
        tb += __shfl_xor(tb, off, 64);
      }
      if (lane == 0) {
        const float sa = fmaf(b, zsa, ta), za = fmaf(-a, sa, zra);
        zl[ca] = za;
        if (ca == bt) {
          ib.zs[cur][ca] = sa;
          ib.zr[nxt][ca] = za;
          zown = zra;
        }
        if (hb) {
          const float sb = fmaf(b, zsb, tb), zb = fmaf(-a, sb, zrb);
          zl[cb] = zb;
          if (cb == bt) {
            ib.zs[cur][cb] = sb;
            ib.zr[nxt][cb] = zb;
            zown = zrb;
          }
        }
      }
    }
  }
  // stage r_{i+1} over the box
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    const int q = (int)threadIdx.x + u * PB;
    if (q < 1000) box[q] = bv[u] ? r_next(br[u], bw[u], bs[u], a, b) : 0.0f;
  }
  __syncthreads();  // box, zl, zown
  double g1 = 0.0, d1 = 0.0, rr1 = 0.0, wrest = 0.0;
  if (i < n) {
    const int x = threadIdx.x & 7, y = (threadIdx.x >> 3) & 7, z = threadIdx.x >> 6;
    const int cpos = (x + 1) + 10 * (y + 1) + 100 * (z + 1);
    const uint4 dq = *reinterpret_cast<const uint4*>(A.bd + (size_t)pj * 16);
    const uint32_t dw[4] = {dq.x, dq.y, dq.z, dq.w};
    const float* pv = A.bv + (size_t)pj * W;
    const float* pa = A.bap + (size_t)pj * W;
    // (A D^-1) r_{i+1} in acc and, two-level, (A P) zr_{i+1} in acz, each in
    // slot order
    float acc = 0.0f, acz = 0.0f;
    // (in slices of 5: the table entries are L1 hits, and loading all 2 W of
    // them at once would spill)
#pragma unroll 5
    for (int k = 0; k < W; ++k) {
      const int dc = (int)((dw[k >> 2] >> (8 * (k & 3))) & 0xffu);
      const int dx = (dc & 3) - 1, dy = ((dc >> 2) & 3) - 1, dz = (dc >> 4) - 1;
      acc = fmaf(pv[k], box[cpos + dx + 10 * dy + 100 * dz], acc);
      if constexpr (TL) {
        const int tc = (((x + dx) >> 3) + 1) + 3 * (((y + dy) >> 3) + 1) + 9 * (((z + dz) >> 3) + 1);
        acz = fmaf(pa[k], zl[nbl[tc]], acz);
      }
    }
    const float di = A.bdi[pj];
    bool free_row = false;
    float uc = 0.0f, ui;
    if constexpr (TL) {
      acc += acz;
      free_row = pa[0] != 0.0f;  // slot 0 is self: A_ii on free rows
      ui = free_row ? fmaf(di, ri, zown) : di * ri;
      uc = free_row ? zl[bt] : 0.0f;
    } else {
      ui = di * ri;
    }
    const float si = fmaf(b, so, wi);
    const float rn = fmaf(-a, si, ri);
    const float pi = fmaf(b, pold, ui);
    ib.x[i] = fmaf(a, pi, xold);
    ib.p[i] = pi;
    ib.s[cur][i] = si;
    ib.r[nxt][i] = rn;
    ib.w[nxt][i] = acc;
    float un;
    if constexpr (TL) un = fmaf(di, rn, uc);
    else un = di * rn + uc;
    g1 += (double)rn * (double)un;
    d1 += (double)acc * (double)un;
    rr1 += (double)rn * (double)rn;
    if (free_row) wrest = (double)acc;
  }
  if constexpr (TL) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      g1 += __shfl_xor(g1, off, 64);
      d1 += __shfl_xor(d1, off, 64);
      rr1 += __shfl_xor(rr1, off, 64);
      wrest += __shfl_xor(wrest, off, 64);
    }
    __syncthreads();
    if (lane == 0) {
      lds[wv] = g1;
      lds[NW + wv] = d1;
      lds[2 * NW + wv] = rr1;
      lds[3 * NW + wv] = wrest;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      double x = 0.0, y = 0.0, z = 0.0, q = 0.0;
#pragma unroll
      for (int k = 0; k < NW; ++k) {
        x += lds[k];
        y += lds[NW + k];
        z += lds[2 * NW + k];
        q += lds[3 * NW + k];
      }
      slot(f, nxt ? S_G1 : S_G0, G)[bt] = x;
      slot(f, nxt ? S_D1 : S_D0, G)[bt] = y;
      slot(f, nxt ? S_RR1 : S_RR0, G)[bt] = z;
      slot(f, nxt ? S_WC1 : S_WC, G)[bt] = q;
    }
  } else {
    publish3<PB>(g1, d1, rr1, slot(f, nxt ? S_G1 : S_G0, G), slot(f, nxt ? S_D1 : S_D0, G),
                 slot(f, nxt ? S_RR1 : S_RR0, G), lds);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    sc->hist[cur] = g;
    sc->hist[2 + cur] = alpha;
    sc->inner_it = it + 1;
  }
}

// Closes an inner solve: records its iteration count and applies
// x += |r| d (skipped when the outer loop has already converged).
template <typename T, int PB>
__global__ void __launch_bounds__(PB) k_apply(EllDev A, FieldBufs f, InnerBufs<T> ib) {
  const SegRange sg = seg_of_block(f, blockIdx.x, gridDim.x);
  PcgScalars* sc = f.sc + sg.s;
  if (idle(sc)) return;
  if (!sc->inner_done) {
    // a pass launched without a host poll ran out of its batch: wait for
    // the host to resume it (every block sees the same flags)
    if (blockIdx.x == 0 && threadIdx.x == 0) sc->stall = 1;
    return;
  }
  if ((int)blockIdx.x == sg.b0 && threadIdx.x == 0) {
    // the pass is counted by the segment's first block (no block of this
    // launch reads the counts)
    if (sc->outer_it < 16) sc->pass_it[sc->outer_it] = sc->inner_it;
    sc->outer_it += 1;
    sc->total_it += sc->inner_it;
  }
  const double rn = sc->rnorm;
  const int64_t r0 = (int64_t)blockIdx.x * PB;
  for (int64_t i = r0 + threadIdx.x; i < r0 + PB && i < A.n; i += PB)
    f.x64[i] = fma(rn, (double)ib.x[i], f.x64[i]);
}

__global__ void k_clear_stall(PcgScalars* sc) { sc->stall = 0; }

__global__ void k_set_controls(PcgScalars* sc, int n, double tol, double inner_tol, int max_it, int max_outer) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  sc[k].tol = tol;
  sc[k].inner_tol = inner_tol;
  sc[k].max_it = max_it;
  sc[k].max_outer = max_outer;
}

// Host-polled summary of the segments: done when every segment is done,
// largest iteration counts, largest relative residual (as rnorm, bnorm2 = 1).
__global__ void __launch_bounds__(kPB) k_seg_summary(FieldBufs f) {
  __shared__ int flags[3];
  __shared__ int mx[3];
  __shared__ double lds[kPB / 64];
  if (threadIdx.x == 0) {
    flags[0] = 1;
    flags[1] = 1;
    flags[2] = 0;
    mx[0] = mx[1] = mx[2] = 0;
  }
  __syncthreads();
  double rel = 0.0;
  for (int s = threadIdx.x; s < f.n_seg; s += kPB) {
    const PcgScalars& q = f.sc[s];
    if (!q.outer_done) atomicAnd(&flags[0], 0);
    if (!q.inner_done) atomicAnd(&flags[1], 0);
    if (q.failed) atomicOr(&flags[2], 1);
    atomicMax(&mx[0], q.inner_it);
    atomicMax(&mx[1], q.total_it);
    atomicMax(&mx[2], q.outer_it);
    const double bn = sqrt(q.bnorm2);
    rel = fmax(rel, q.rnorm / (bn > 0.0 ? bn : 1.0));
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) rel = fmax(rel, __shfl_xor(rel, off, 64));
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = rel;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kPB / 64; ++w) rel = fmax(rel, lds[w]);
    rel = fmax(rel, lds[0]);
    PcgScalars* o = f.sc + f.n_seg;
    o->outer_done = flags[0];
    o->inner_done = flags[1];
    o->failed = flags[2];
    o->inner_it = mx[0];
    o->total_it = mx[1];
    o->outer_it = mx[2];
    o->rnorm = rel;
    o->bnorm2 = 1.0;
  }
}

// Floating potential from Gauss's law on the conductor (one block per
// segment; 1024 threads: the conductor's coupling entries are a few per
// thread, each a dependent gather).
constexpr int kFloatBlock = 1024;
__global__ void __launch_bounds__(kFloatBlock) k_float(FieldBufs f, FloatArgs fa) {
  __shared__ double lds[kFloatBlock / 64];
  const int sg = blockIdx.x;
  const int64_t j0 = f.kc_off ? f.kc_off[sg] : 0, j1 = f.kc_off ? f.kc_off[sg + 1] : f.n_kc;
  double s = 0.0;
  for (int64_t j = j0 + threadIdx.x; j < j1; j += blockDim.x) s += f.kc_val[j] * f.x64[f.kc_idx[j]];
  const double tot = block_sum(s, lds);
  if (threadIdx.x == 0) {
    PcgScalars* sc = f.sc + sg;
    const size_t lo = (size_t)sg * kMaxGroups * kMaxSpecies;
    const double ws = fa.seg_wscale ? fa.seg_wscale[sg] : 1.0;
    double q = fa.seg_qinit ? fa.seg_qinit[sg] : fa.q_init;
    for (int g = 0; g < kMaxGroups; ++g) {
      if (!((fa.cond_mask >> g) & 1u)) continue;
      for (int sp = 0; sp < kMaxSpecies; ++sp) {
        const double qw = fa.qw[sp] * ws;
        q += (double)fa.ledger_count[lo + g * kMaxSpecies + sp] * qw;
        if (fa.emit_count) q -= (double)fa.emit_count[lo + g * kMaxSpecies + sp] * qw;
      }
    }
    // a failed solve keeps the previous floating potential (its phi_0 is
    // restored from the previous solution, PICCycle.cpp:731-767)
    if (sc->failed) return;
    const double cap = fa.seg_ccap ? fa.seg_ccap[sg] : fa.ccap;
    sc->ccap = cap;
    sc->qcond = q;
    sc->vfloat = cap > 0.0 ? (q + sc->cond_load - tot) / cap : 0.0;
  }
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_phi(int64_t n, FieldBufs f, T* phiT, double* phi64) {
  const uint32_t lb = xcd_remap(blockIdx.x, gridDim.x);
  const int64_t i = (int64_t)lb * kBlock + threadIdx.x;
  if (i >= n) return;
  double p = f.x64[i];
  if (f.n_basis > 0) {
    for (int k = 0; k < f.n_basis; ++k) p = fma(f.coef[k], f.basis[(int64_t)k * n + i], p);
  } else if (f.phi1) {
    p = fma(f.sc[f.blk_seg ? f.blk_seg[i / kPcgBlock] : 0].vfloat, f.phi1[i], p);
  }
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

static inline int grid_rows(int64_t n, int pb) { return (int)((n + pb - 1) / pb); }
// blocks of the coarse kernel: each recomputes the restricted vector, so few
static inline int coarse_blocks(const EllDev& A) {
  const int per_block = (kPB / 64) * kCoarseRowsPerWave;
  return (A.nc + per_block - 1) / per_block;
}

// Row-block kernels by block size (EllDev::pb).
#define SCDAT_PB(call) \
  do {                 \
    if (A.pb == 256) {  \
      constexpr int PB = 256; \
      call;            \
    } else {           \
      constexpr int PB = 512; \
      call;            \
    }                  \
  } while (0)

void launch_rhs(const EllDev& A, const FieldBufs& f, const unsigned long long* rho_fx,
                double unit, int load_sign, hipStream_t s) {
  const int G = grid_rows(A.n, A.pb);
  SCDAT_PB(hipLaunchKernelGGL(k_rhs<PB>, dim3(G), dim3(PB), 0, s, A, f, rho_fx, unit, load_sign));
  hipLaunchKernelGGL(k_fin_rhs, dim3(f.blk_seg ? f.n_seg : 1), dim3(kPB), 0, s, f, G);
}
void launch_rhs_real(const EllDev& A, const FieldBufs& f, const double* src, double scale, int load_sign,
                     int homogeneous, hipStream_t s) {
  const int G = grid_rows(A.n, A.pb);
  SCDAT_PB(hipLaunchKernelGGL(k_rhs_real<PB>, dim3(G), dim3(PB), 0, s, A, f, src, scale, load_sign, homogeneous));
  hipLaunchKernelGGL(k_fin_rhs, dim3(f.blk_seg ? f.n_seg : 1), dim3(kPB), 0, s, f, G);
}
void launch_resid64(const EllDev& A, const FieldBufs& f, hipStream_t s) {
  const int G = grid_rows(A.n, A.pb);
  SCDAT_PB(hipLaunchKernelGGL(k_resid64<PB>, dim3(G), dim3(PB), 0, s, A, f));
  hipLaunchKernelGGL(k_fin_resid, dim3(f.blk_seg ? f.n_seg : 1), dim3(kPB), 0, s, f, G);
}
template <typename T>
void launch_inner_init(const EllDev& A, const FieldBufs& f, const InnerBufs<T>& ib,
                       hipStream_t s) {
  const int G = grid_rows(A.n, A.pb);
  SCDAT_PB(hipLaunchKernelGGL((k_inner_init<T, PB>), dim3(G), dim3(PB), 0, s, A, f, ib));
  if (A.nc) hipLaunchKernelGGL(k_coarse<T>, dim3(coarse_blocks(A)), dim3(kPB), 0, s, A, f, ib, 0, 0, G);
  bool started = false;
  if constexpr (sizeof(T) == 4) {
    if (A.box && (A.nc == 0 || A.fused)) {  // tile-staged, as the iterations
      const dim3 g(grid_rows(A.n, kPcgBlock)), b(kPcgBlock);
      started = true;
      switch (A.w * 2 + (A.nc ? 1 : 0)) {
#define SCDAT_W(w)                                                                  \
  case 2 * w:                                                                       \
    hipLaunchKernelGGL((k_cg_start_box<w, false>), g, b, 0, s, A, f, ib);           \
    break;                                                                          \
  case 2 * w + 1:                                                                   \
    hipLaunchKernelGGL((k_cg_start_box<w, true>), g, b, 0, s, A, f, ib);            \
    break;
        SCDAT_W(13) SCDAT_W(14) SCDAT_W(15) SCDAT_W(16)
#undef SCDAT_W
        default:
          started = false;  // ell() only stages these widths
      }
    }
  }
  if (!started) SCDAT_PB(hipLaunchKernelGGL((k_cg_start<T, PB>), dim3(G), dim3(PB), 0, s, A, f, ib));
  // fused: iteration 0 forms zw_0 from the restriction k_cg_start published
  if (A.nc && !A.fused)
    hipLaunchKernelGGL(k_coarse<T>, dim3(coarse_blocks(A)), dim3(kPB), 0, s, A, f, ib, 1, 0, G);
}
// One iteration at block size PB (compile-time ELL width where one exists).
template <typename T, int PB>
void launch_iter_pb(const EllDev& A, const FieldBufs& f, const InnerBufs<T>& ib, int it, hipStream_t s,
                    hipEvent_t e0, hipEvent_t e1) {
  const dim3 g(grid_rows(A.n, PB)), b(PB);
  if constexpr (sizeof(T) == 4) {
    if (A.npat > 0) switch (A.w) {
#define SCDAT_W(w) \
  case w:          \
    hipExtLaunchKernelGGL((k_cg_iter<T, w, false, true, PB>), g, b, 0, s, e0, e1, 0, A, f, ib, it); \
    return;
      SCDAT_W(7) SCDAT_W(8) SCDAT_W(9) SCDAT_W(10) SCDAT_W(11) SCDAT_W(12) SCDAT_W(13) SCDAT_W(14)
      SCDAT_W(15) SCDAT_W(16) SCDAT_W(17) SCDAT_W(18) SCDAT_W(19) SCDAT_W(20)
#undef SCDAT_W
      default:
        break;  // no compile-time width: the ELL form below
    }
  }
  switch (A.w) {
#define SCDAT_W(w) \
  case w:          \
    hipExtLaunchKernelGGL((k_cg_iter<T, w, false, false, PB>), g, b, 0, s, e0, e1, 0, A, f, ib, it); \
    break;
    SCDAT_W(7) SCDAT_W(8) SCDAT_W(9) SCDAT_W(10) SCDAT_W(11) SCDAT_W(12) SCDAT_W(13) SCDAT_W(14)
    SCDAT_W(15) SCDAT_W(16) SCDAT_W(17) SCDAT_W(18) SCDAT_W(19) SCDAT_W(20)
#undef SCDAT_W
    default:
      hipExtLaunchKernelGGL((k_cg_iter<T, 0, false, false, PB>), g, b, 0, s, e0, e1, 0, A, f, ib, it);
  }
}
template <typename T>
void launch_inner_iter(const EllDev& A, const FieldBufs& f, const InnerBufs<T>& ib, int it,
                       hipStream_t s, hipEvent_t* ev) {
  hipEvent_t e0 = ev ? ev[0] : nullptr, e1 = ev ? ev[1] : nullptr;
  if constexpr (sizeof(T) == 4) {
    if (A.box && (A.nc == 0 || A.fused)) {  // tile-staged (Jacobi or fused two-level)
      const dim3 g(grid_rows(A.n, kPcgBlock)), b(kPcgBlock);
      switch (A.w * 2 + (A.nc ? 1 : 0)) {
#define SCDAT_W(w) \
  case 2 * w:      \
    hipExtLaunchKernelGGL((k_cg_iter_box<w, false>), g, b, 0, s, e0, e1, 0, A, f, ib, it); \
    return;        \
  case 2 * w + 1:  \
    hipExtLaunchKernelGGL((k_cg_iter_box<w, true>), g, b, 0, s, e0, e1, 0, A, f, ib, it); \
    return;
        SCDAT_W(13) SCDAT_W(14) SCDAT_W(15) SCDAT_W(16)
#undef SCDAT_W
        default:
          break;  // ell() only stages these widths
      }
    }
    if (A.nc && A.fused && A.pap32) {  // two-level, coarse solve inside the iteration
      const dim3 g(grid_rows(A.n, kPcgBlock)), b(kPcgBlock);
      switch (A.w) {
#define SCDAT_W(w) \
  case w:          \
    hipExtLaunchKernelGGL(k_cg_iter_tl<w>, g, b, 0, s, e0, e1, 0, A, f, ib, it); \
    return;
        SCDAT_W(13) SCDAT_W(14) SCDAT_W(15) SCDAT_W(16) SCDAT_W(17) SCDAT_W(18)
#undef SCDAT_W
        default:
          break;  // ell() only fuses these widths
      }
    }
  }
  if (A.nc) {  // two-level: 512-row blocks (the aggregates)
    const dim3 g(grid_rows(A.n, kPcgBlock)), b(kPcgBlock);
    switch (A.w) {
#define SCDAT_W(w) \
  case w:          \
    hipExtLaunchKernelGGL((k_cg_iter<T, w, true, false, kPcgBlock>), g, b, 0, s, e0, e1, 0, A, f, ib, it); \
    break;
      SCDAT_W(13) SCDAT_W(14) SCDAT_W(15) SCDAT_W(16) SCDAT_W(17) SCDAT_W(18)
#undef SCDAT_W
      default:
        hipExtLaunchKernelGGL((k_cg_iter<T, 0, true, false, kPcgBlock>), g, b, 0, s, e0, e1, 0, A, f, ib, it);
    }
    hipLaunchKernelGGL(k_coarse<T>, dim3(coarse_blocks(A)), dim3(kPB), 0, s, A, f, ib, 2, it,
                       grid_rows(A.n, kPcgBlock));
    return;
  }
  if (A.pb == 256) launch_iter_pb<T, 256>(A, f, ib, it, s, e0, e1);
  else launch_iter_pb<T, 512>(A, f, ib, it, s, e0, e1);
}
template <typename T>
void launch_apply(const EllDev& A, const FieldBufs& f, const InnerBufs<T>& ib, hipStream_t s) {
  SCDAT_PB(hipLaunchKernelGGL((k_apply<T, PB>), dim3(grid_rows(A.n, A.pb)), dim3(PB), 0, s, A, f, ib));
}
size_t partial_doubles(int64_t n) {  // smallest block, slots 256-byte aligned
  return (size_t)S_COUNT * (size_t)((grid_rows(n, 256) + 31) & ~31) + 16;
}
void launch_set_controls(const FieldBufs& f, int n_scalars, double tol, double inner_tol, int max_it,
                         int max_outer, hipStream_t s) {
  hipLaunchKernelGGL(k_set_controls, dim3((n_scalars + 63) / 64), dim3(64), 0, s, f.sc, n_scalars, tol, inner_tol,
                     max_it, max_outer);
}
void launch_clear_stall(const FieldBufs& f, hipStream_t s) {
  hipLaunchKernelGGL(k_clear_stall, dim3(1), dim3(1), 0, s, f.sc);
}
void launch_float(const FieldBufs& f, const FloatArgs& fa, hipStream_t s) {
  hipLaunchKernelGGL(k_float, dim3(f.blk_seg ? f.n_seg : 1), dim3(kFloatBlock), 0, s, f, fa);
}
__global__ void __launch_bounds__(kBlock) k_extrapolate(FieldBufs f, double* prev, int64_t n, int first) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const double t = f.x64[i];
  if (!first && f.kind[i] == 0) f.x64[i] = 2.0 * t - prev[i];
  prev[i] = t;
}
void launch_extrapolate(const FieldBufs& f, double* prev, int64_t n, int first, hipStream_t s) {
  hipLaunchKernelGGL(k_extrapolate, dim3(grid_for(n)), dim3(kBlock), 0, s, f, prev, n, first);
}
// Rows of segments whose solve failed take back the previous solution, so
// that phi and E rebuilt from x64 equal the previous step's bit for bit.
__global__ void __launch_bounds__(kBlock) k_keep_failed(FieldBufs f, const double* prev, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  if (f.sc[f.blk_seg ? f.blk_seg[i / kPcgBlock] : 0].failed) f.x64[i] = prev[i];
}
void launch_keep_failed(const FieldBufs& f, const double* prev, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_keep_failed, dim3(grid_for(n)), dim3(kBlock), 0, s, f, prev, n);
}
void launch_seg_summary(const FieldBufs& f, hipStream_t s) {
  if (!f.blk_seg) return;
  hipLaunchKernelGGL(k_seg_summary, dim3(1), dim3(kPB), 0, s, f);
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
