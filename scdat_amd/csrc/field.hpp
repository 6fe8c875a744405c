// Launch interface of the Poisson / nodal-field kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "device_common.hpp"

namespace scdat {

// Threads and rows of a PCG row block (one row per thread).  512 rather than
// 256: every block sums all blocks' partials in its prologue, so half as
// many blocks cut that read 4x (12.0 -> 11.3 us per iteration at 64^3).
// Scenario node ranges and the two-level aggregates follow the row blocks.
constexpr int kPcgBlock = 512;
constexpr int kMaxCoarse = 2048;               // row blocks of the two-level preconditioner
constexpr int kCoarseBlockRows = kPcgBlock;    // nodes per aggregate
constexpr int kMaxTouched = 32;                // aggregates one row block may touch (fused two-level)

// Scalars of one solve, resident on the device; the host polls `inner_done`
// and `outer_done` between batches of iterations.
struct PcgScalars {
  double bnorm2, cond_load, rnorm, inv_rnorm;
  double rz, pq, alpha, beta, rr;
  double vfloat, qcond, ccap;
  double tol, inner_tol, inner_tol_eff;
  double hist[4];  // (g_i, alpha_i) of the last inner iteration, by parity
  int32_t outer_done, inner_done, inner_it, outer_it, total_it, max_it, max_outer, failed;
  unsigned int ticket[4];
  double init_rnorm;      // residual norm of the solve's initial guess
  int32_t pass_it[16];    // inner iterations of each outer pass (the first 16)
  // A pass launched without a host poll whose batch ended before its inner
  // solve stopped: the apply sets stall instead of applying, and every
  // later kernel of the solve does nothing until the host resumes the pass.
  int32_t stall;
};

// ELL operator over all nodes, slot-major (slot s of row i at s*n + i), self
// in slot 0.  Constrained rows (Dirichlet, conductor) are identity rows and
// their columns are eliminated from free rows (symmetric elimination).
struct EllDev {
  int64_t n;
  int w;
  const int32_t* col;
  const float* v32;
  const double* v64;
  const float* vh32;   // A D^-1 (column-scaled by the Jacobi inverse), inner solve
  const double* vh64;
  const float* dinv32;
  const double* dinv64;
  // nodal-gradient weights (LSQ), same slots, 3 components
  const float* gx32;
  const float* gy32;
  const float* gz32;
  const double* gx64;
  const double* gy64;
  const double* gz64;
  // Two-level preconditioner M^-1 = D^-1 + P Ac^-1 P^T (nc = 0: Jacobi only).
  // The aggregates are the row blocks: P maps block b to its free rows (the
  // context numbers nodes in Morton order, so a block is a compact box),
  // Ac = P^T A P over the blocks (empty blocks: zero rows).
  int nc;                  // row blocks (coarse vector length)
  const float* dco32;      // 1 / D^-1 on free rows, 0 on constrained rows
  const double* dco64;
  const float* aci32;      // Ac^-1, nc x nc row-major
  const double* aci64;
  // Fused coarse solve (FP32 row patterns): block b applies the rows of
  // Ac^-1 of the aggregates its rows touch, cnb[cnb_off[b] .. cnb_off[b+1])
  // (its own first), inside the iteration kernel; pap32 holds per pattern
  // slot the FP32 entry of A P (A_ij on free rows and free columns, else 0).
  int fused;
  const int32_t* cnb_off;
  const int32_t* cnb;
  const float* pap32;
  int ncp;                 // row stride of acip32: nc rounded up to 64, zero-padded
  const float* acip32;     // Ac^-1 in FP32 with that stride
  // Tile-staged iteration (tiled order, FP32; box = 0: not used).  Block b
  // stages r_{i+1} of its tile and the one-point halo (a 10^3 box) in LDS
  // once; row i of b then reads its neighbours there by direction code
  // (delta + 1 in two bits per axis, x lowest) from its box pattern.
  int box;
  const uint16_t* bpat;    // box pattern of each row
  const uint8_t* bd;       // 16 direction codes per pattern (21: the row itself)
  const float* bv;         // A D^-1 entries per pattern (w per pattern)
  const float* bap;        // A P entries per pattern (two-level; 0 otherwise)
  const float* bdi;        // D^-1 per pattern
  const int32_t* nbb;      // 27 neighbour blocks per block (13: itself; -1: none)
  // Row patterns of the FP32 inner operator (npat = 0: not used).  Rows with
  // the same column offsets j - i, the same FP32 values of A D^-1 and the
  // same FP32 D^-1 share one table entry, so an inner iteration reads a
  // 2-byte pattern id per row instead of the row's ELL slots.  The values
  // are the ELL's own, so both forms give the same bits.
  int npat;
  const uint16_t* pat;     // pattern of each row (npat <= 65536), else
  const uint32_t* pat32;   // 4-byte ids
  const int32_t* poff;     // npat x w column offsets (pattern-major)
  const float* pval32;     // npat x w values of A D^-1
  const float* pdinv32;    // npat values of D^-1
  // Threads (= rows) of a PCG row block: kPcgBlock, or 256 for a single
  // scenario under about 131k rows with the Jacobi preconditioner (half the
  // rows per block keeps the grid wide enough to fill the GPU).
  int pb;
};

struct FieldBufs {
  const int32_t* kind;      // node kind
  const double* lift;       // Dirichlet lifting of the load (free rows)
  const double* dval;       // Dirichlet values
  const double* dscale;     // diagonal of constrained rows
  const double* phi1;       // unit-potential conductor solution
  const int32_t* kc_idx;    // sparse vector sum of conductor rows of K
  const double* kc_val;
  int64_t n_kc;
  double* b64;
  double* x64;
  double* r64;
  double* partial;          // >= 2 * grid doubles
  PcgScalars* sc;
  // conductor patches held at circuit potentials: phi = x + sum_k coef[k] basis_k
  // (n_basis > 0 replaces the single floating conductor phi1 / vfloat)
  int n_basis;
  const double* basis;      // n_basis x n, patch k at 1 V, everything else 0
  const double* coef;       // n_basis potentials (device)
  // Independent segments (batched scenarios).  The rows of segment s are the
  // row blocks [seg_b0[s], seg_b0[s+1]) and blk_seg[b] names the segment of
  // block b; sc[s], the charge unit and the conductor sums are per segment,
  // and sc[n_seg] is a summary the host polls (all done, largest counts).
  // blk_seg null: one segment over all rows, sc[0] only.
  int n_seg;
  const int32_t* blk_seg;
  const int32_t* seg_b0;
  const double* seg_unit;
  const int64_t* kc_off;     // n_seg + 1 offsets into kc_idx / kc_val
};

// FP32 (mixed) or FP64 inner-solve vectors: r, w = A D^-1 r and s = A p by
// iteration parity, the search direction p and the correction x.
template <typename T>
struct InnerBufs {
  T* r[2];
  T* w[2];
  T* s[2];
  T* p;
  T* x;
  // two-level: coarse vectors Ac^-1 P^T r, Ac^-1 P^T s, Ac^-1 P^T w by
  // iteration parity (one entry per row block)
  T* zr[2];
  T* zs[2];
  T* zw[2];
};

void launch_rhs(const EllDev& A, const FieldBufs& f, const unsigned long long* rho_fx,
                double unit, int load_sign, hipStream_t s);
// b from an FP64 nodal vector: deposited charge (x scale, like launch_rhs),
// or homogeneous: src on free rows, 0 on constrained rows
void launch_rhs_real(const EllDev& A, const FieldBufs& f, const double* src, double scale, int load_sign,
                     int homogeneous, hipStream_t s);
void launch_resid64(const EllDev& A, const FieldBufs& f, hipStream_t s);
template <typename T>
void launch_inner_init(const EllDev& A, const FieldBufs& f, const InnerBufs<T>& ib, hipStream_t s);
// iteration `it` of the current inner solve (0, 1, ...), one kernel; with ev
// non-null its dispatch timestamps land in ev[0], ev[1] (kernel execution
// only, no launch gaps)
template <typename T>
void launch_inner_iter(const EllDev& A, const FieldBufs& f, const InnerBufs<T>& ib, int it,
                       hipStream_t s, hipEvent_t* ev);
// doubles needed by the partial-sum buffer for n rows
size_t partial_doubles(int64_t n);
template <typename T>
void launch_apply(const EllDev& A, const FieldBufs& f, const InnerBufs<T>& ib, hipStream_t s);
// Conductor charge = q_init + sum over conductor groups and species of
// absorbed count * q_s w_s; floating potential from Gauss's law.
struct FloatArgs {
  double q_init;
  const unsigned long long* ledger_count;  // [segment][kMaxGroups][kMaxSpecies]
  const unsigned long long* emit_count;    // [segment][kMaxGroups][kMaxSpecies] or null
  double qw[kMaxSpecies];
  uint32_t cond_mask;
  // per segment (null without segments): initial charge, weight multiplier
  // of every species' qw, capacitance sum_C K phi_1
  const double* seg_qinit;
  const double* seg_wscale;
  const double* seg_ccap;
  double ccap;  // without segments: the capacitance sum_C K phi_1
};
void launch_float(const FieldBufs& f, const FloatArgs& fa, hipStream_t s);
// Lets a stalled solve (one scenario) go on: the host resumes the stalled pass.
void launch_clear_stall(const FieldBufs& f, hipStream_t s);
// The stopping controls of the solve(s) on the device, without a host round trip.
void launch_set_controls(const FieldBufs& f, int n_scalars, double tol, double inner_tol, int max_it,
                         int max_outer, hipStream_t s);
// Initial guess of a time step's solve from the last two solutions: on free
// rows x = 2 x_n - x_{n-1} (first = 1: x stays x_n); prev takes x_n.
void launch_extrapolate(const FieldBufs& f, double* prev, int64_t n, int first, hipStream_t s);
// After a solve: rows of every segment whose PCG failed get x64 = prev (the
// previous step's solution), the reference's keep-the-previous-potential
// policy (PICCycle.cpp:731-767).
void launch_keep_failed(const FieldBufs& f, const double* prev, int64_t n, hipStream_t s);
// Summary of the per-segment scalars into sc[n_seg] (no-op for one segment).
void launch_seg_summary(const FieldBufs& f, hipStream_t s);
template <typename T>
void launch_phi_efield(const EllDev& A, const FieldBufs& f, T* phiT, double* phi64, Vec4<T>* E,
                       unsigned long long* rho_clear, hipStream_t s);
template <typename T>
void launch_efield_only(const EllDev& A, const T* phiT, Vec4<T>* E, hipStream_t s);

}  // namespace scdat
