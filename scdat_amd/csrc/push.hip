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
#include <type_traits>

#include "device_common.hpp"
#include "push.hpp"
#include "rng.hpp"
#include "walk.hpp"

namespace scdat {

// LDS-tiled fixed-point deposition.
//  1. wave level: runs of lanes with the same cell (particles are cell-sorted)
//     are reduced by a segmented inclusive scan on the 32-bit weights;
//  2. block level: each run's tail lane adds c_s * weight for its 4 nodes into
//     an LDS open-addressing table keyed by node id (64-bit LDS atomics);
//  3. after all chunks of the block, the table is flushed with one 64-bit
//     global atomic per distinct node.
// A node that finds no free slot within kProbe probes goes straight to the
// global accumulator, so any particle order is correct, just slower.
constexpr int kHash = 512;
constexpr int kProbe = 8;
// Particles per thread of the push kernel (the deposit tile is chunks x
// kBlock consecutive particles): 1, 2, 4 or 8, chosen per problem by
// measurement (scdat_autotune); 4 by default.

struct DepositTile {
  int32_t key[kHash];
  unsigned long long val[kHash];
};

__device__ __forceinline__ void tile_clear(DepositTile& L) {
  for (int j = threadIdx.x; j < kHash; j += kBlock) {
    L.key[j] = -1;
    L.val[j] = 0ull;
  }
}

__device__ __forceinline__ void tile_flush(DepositTile& L, unsigned long long* rho) {
  for (int j = threadIdx.x; j < kHash; j += kBlock) {
    const int32_t k = L.key[j];
    if (k >= 0) {
      const unsigned long long v = L.val[j];
      if (v) atomicAdd(&rho[k], v);
    }
  }
}

// The block's table into the deposit buckets (every thread of the block
// calls it, after a barrier that follows the last insertion).
__device__ void tile_flush_buckets(DepositTile& L, unsigned long long* rho, const DepBuckets& B) {
  static_assert(kHash == 2 * kBlock, "two table slots per thread");
  __shared__ int32_t rid[kDepRanges];
  __shared__ unsigned int rcount[kDepRanges], rbase[kDepRanges];
  if (threadIdx.x < kDepRanges) {
    rid[threadIdx.x] = -1;
    rcount[threadIdx.x] = 0u;
  }
  __syncthreads();
  int slot[2] = {-1, -1};
  unsigned int rank[2] = {0u, 0u};
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const int j = threadIdx.x + q * kBlock;
    const int32_t k = L.key[j];
    if (k < 0 || L.val[j] == 0ull) continue;
    const int32_t r = k / kDepRangeRows;
    for (int e = 0; e < kDepRanges; ++e) {
      const int32_t o = atomicCAS(&rid[e], -1, r);
      if (o == -1 || o == r) {
        slot[q] = e;
        break;
      }
    }
    if (slot[q] >= 0) rank[q] = atomicAdd(&rcount[slot[q]], 1u);
    else atomicAdd(&rho[k], L.val[j]);  // more ranges than the block tracks
  }
  __syncthreads();
  if (threadIdx.x < kDepRanges && rid[threadIdx.x] >= 0)
    rbase[threadIdx.x] = atomicAdd(&B.count[rid[threadIdx.x]], rcount[threadIdx.x]);
  __syncthreads();
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    if (slot[q] < 0) continue;
    const int j = threadIdx.x + q * kBlock;
    const unsigned int at = rbase[slot[q]] + rank[q];
    if (at < (unsigned int)B.cap) {
      DepEntry d;
      d.node = L.key[j];
      d.pad = 0;
      d.val = L.val[j];
      B.ent[(size_t)rid[slot[q]] * (size_t)B.cap + at] = d;
    } else {
      atomicAdd(&rho[L.key[j]], L.val[j]);  // bucket full
    }
  }
}

template <typename T>
__device__ __forceinline__ void tile_runs(int key, int32_t f[4], int64_t cs,
                                          const TetMap<T>* __restrict__ tets,
                                          unsigned long long* rho, DepositTile& L) {
  const int lane = threadIdx.x & 63;
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
    const int4 nd = *reinterpret_cast<const int4*>(tets[key].node);
    const int32_t nodes[4] = {nd.x, nd.y, nd.z, nd.w};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int32_t node = nodes[q];
      const unsigned long long val = (unsigned long long)(cs * (int64_t)f[q]);
      uint32_t h = ((uint32_t)node * 2654435761u) >> (32 - 9);
      bool done = false;
      for (int p = 0; p < kProbe && !done; ++p) {
        const int32_t k = atomicCAS(&L.key[h], -1, node);
        if (k == -1 || k == node) {
          atomicAdd(&L.val[h], val);
          done = true;
        }
        h = (h + 1) & (kHash - 1);
      }
      if (!done) atomicAdd(&rho[node], val);
    }
  }
}

// One block pushes CH x kBlock consecutive particles.  LAT (FP32, lattice
// meshes): sub-steps that stay clear of boundary faces locate their end tet
// on the grid instead of walking.
template <typename T, int CH, bool LAT>
__global__ void __launch_bounds__(kBlock) push_kernel(PushArgs<T> a) {
  const int64_t n = a.n_dev ? *a.n_dev : a.n;
  // a captured step's grid covers the capacity: the blocks the count needs
  // are remapped among themselves (so every XCD gets its share), the others
  // only publish empty partials
  const int64_t need = (n + (int64_t)CH * kBlock - 1) / ((int64_t)CH * kBlock);
  const uint32_t nact = need < (int64_t)gridDim.x ? (uint32_t)need : gridDim.x;
  if (a.n_pushed && blockIdx.x == 0 && threadIdx.x == 0) *a.n_pushed = n;
  if (blockIdx.x >= nact) {
    if (threadIdx.x == 0) {
      a.ke_partial[blockIdx.x] = 0.0;
      a.alive_partial[blockIdx.x] = 0;
    }
    return;
  }
  const uint32_t lb = xcd_remap(blockIdx.x, nact);
  const uint32_t mcc_step = a.step_dev ? *a.step_dev : a.mcc_step;
  __shared__ DepositTile tile;
  __shared__ unsigned int led_n[kMaxGroups];
  __shared__ double led_e[kMaxGroups];
  tile_clear(tile);
  if (threadIdx.x < kMaxGroups) {
    led_n[threadIdx.x] = 0u;
    led_e[threadIdx.x] = 0.0;
  }
  __syncthreads();
  double ke = 0.0, v2max = 0.0;
  int alive = 0, started = 0;
#pragma unroll 1
  for (int ch = 0; ch < CH; ++ch) {
    const int64_t i = ((int64_t)lb * CH + ch) * kBlock + threadIdx.x;
    int key = -1;
    bool is_cand = false;  // collision candidate: entry = slot | kMccForced when forced
    uint32_t cand = 0;
    int32_t f[4] = {0, 0, 0, 0};
    int mseg = -1;  // scenario of this lane's |v|^2 for the collision majorant
    double mv2 = 0.0;
    int c = (i < n) ? a.cell[i] : -1;
    if (c >= 0) {
      ++started;
      T x[3] = {a.x[i], a.y[i], a.z[i]};
      T v[3] = {a.vx[i], a.vy[i], a.vz[i]};
      const T v0[3] = {v[0], v[1], v[2]};
      // FP32: walk records (one dependent load per face crossing); the nodes
      // of the sub-step's start cell come from the tet map.  LAT: the tet
      // map of the start cell (affine map and nodes in one record) and its
      // hexahedron; the walk record only when a sub-step walks.
      using Map = typename std::conditional<sizeof(T) == 4, WalkMap<float>, TetMap<T>>::type;
      Map m;
      TetMap<T> tm;
      uint32_t th = 0;
      if constexpr (LAT) {
        tm = a.tets[c];
        th = a.lat.tet_hex[c];
      } else {
        load_map(a, c, m);
      }
      int sg = 0;
      T remaining = a.dt;
      double wt = a.weight;
      if (a.tet_seg) {
        sg = a.tet_seg[c];
        remaining = a.seg_dt[sg];
        wt = a.weight * a.seg_wscale[sg];
      }
      int sub = 0;
      int dead = 0;
      int group = -1, face = 0, last = c;
      T hit[3] = {x[0], x[1], x[2]};
      // Sub-step loop, PICParticleAlgorithms.cpp:593-701.
      while (remaining > T(1e-20) && sub < a.max_sub) {
        ++sub;
        T lam[4];
        T E[3];
        if constexpr (LAT) {
          bary(tm, x, lam);
          gather_E(a, tm.node, lam, E);
        } else if constexpr (sizeof(T) == 4) {
          bary(m, x, lam);
          const int4 nd4 = *reinterpret_cast<const int4*>(a.tets[c].node);
          const int32_t nd[4] = {nd4.x, nd4.y, nd4.z, nd4.w};
          gather_E(a, nd, lam, E);
        } else {
          bary(m, x, lam);
          gather_E(a, m, lam, E);
        }
        const T h =
            a.adaptive ? adaptive_dt(v0, E, a.bmag, a.qm, remaining, a.floor_ratio) : remaining;
        boris(v, E, a.B, a.has_b != 0, a.qm, h);
        T xs[3] = {x[0], x[1], x[2]};
        T xn[3] = {fma_t(v[0], h, x[0]), fma_t(v[1], h, x[1]), fma_t(v[2], h, x[2])};
        int r;
        if constexpr (LAT) {
          // the walk's first test: still inside the start tet
          T l1[4];
          bary(tm, xn, l1);
          if (l1[0] >= -a.tol && l1[1] >= -a.tol && l1[2] >= -a.tol && l1[3] >= -a.tol) {
            r = c;
          } else {
            uint32_t hn = 0;
            r = lat_locate(a.lat, th, xn, hn);
            if (r == -3) {
              // the walk record only here, so it shares registers with tm
              Map wm;
              load_map(a, c, wm);
              r = walk(a, c, xs, xn, v, group, hit, wm, face, last);
              if (r >= 0) hn = a.lat.tet_hex[r];
            }
            if (r >= 0) {
              tm = a.tets[r];
              th = hn;
            }
          }
        } else {
          r = walk(a, c, xs, xn, v, group, hit, m, face, last);
        }
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
          // PICCycle::recordSurfaceBoundaryInteraction (PICCycle.cpp:1295-1329):
          // count and incident energy KE * max(1, w), aggregated per block.
          const int g = (group >= 0 && group < kMaxGroups) ? group : 0;
          const double k = 0.5 * (double)a.mass *
                           ((double)v[0] * v[0] + (double)v[1] * v[1] + (double)v[2] * v[2]);
          if (a.tet_seg) {
            // scenarios may share a block: straight to the scenario's slots
            const size_t o = ((size_t)sg * kMaxGroups + g) * kMaxSpecies + a.species;
            atomicAdd(&a.ledger_count[o], 1ull);
            unsafeAtomicAdd(&a.ledger_energy[o], k * fmax(1.0, wt));
          } else {
            atomicAdd(&led_n[g], 1u);
            atomicAdd(&led_e[g], k * a.energy_weight);
          }
          // impact recorded for secondary emission after the push
          if ((a.see_mask >> g) & 1u) {
            const unsigned long long e = atomicAdd(a.n_events, 1ull);
            if ((int64_t)e < a.event_cap) {
              SeeEvent ev;
              ev.x[0] = hit[0];
              ev.x[1] = hit[1];
              ev.x[2] = hit[2];
              ev.v[0] = v[0];
              ev.v[1] = v[1];
              ev.v[2] = v[2];
              ev.id = a.id[i];
              ev.cell = last;
              ev.group_face = g | (face << 8);
              ev.species = a.species;
              a.events[e] = ev;
            }
          }
        }
        a.cell[i] = -1;
        a.x[i] = hit[0];
        a.y[i] = hit[1];
        a.z[i] = hit[2];
      } else {
        T lam[4];
        if constexpr (LAT) bary(tm, x, lam);
        else bary(m, x, lam);
        lambda_fixed(lam, f);
        key = c;
        alive += 1;
        ke += 0.5 * (double)a.mass * wt *
              ((double)v[0] * v[0] + (double)v[1] * v[1] + (double)v[2] * v[2]);
        a.x[i] = x[0];
        a.y[i] = x[1];
        a.z[i] = x[2];
        a.cell[i] = c;
        if (a.mcc_ctl) {
          // null-collision pre-selection on the post-push velocity (with
          // scenarios: the scenario's majorant and key)
          const double v2 = (double)v[0] * v[0] + (double)v[1] * v[1] + (double)v[2] * v[2];
          const double* ctl = a.mcc_ctl + (a.tet_seg ? (size_t)sg * kMccCtl : 0);
          const uint32_t k0 = a.tet_seg ? a.mcc_seg_key[2 * sg] : a.mcc_key0;
          const uint32_t k1 = a.tet_seg ? a.mcc_seg_key[2 * sg + 1] : a.mcc_key1;
          v2max = fmax(v2max, v2);
          mseg = sg;
          mv2 = v2;
          if (v2 > ctl[1]) {
            is_cand = true;
            cand = (uint32_t)i | kMccForced;
          } else {
            const U4 r = philox4x32((uint32_t)a.id[i], mcc_step, (uint32_t)a.species, kTagMccA, k0, k1);
            is_cand = u01(r.x) < ctl[0];
            cand = (uint32_t)i;
          }
        }
      }
      a.vx[i] = v[0];
      a.vy[i] = v[1];
      a.vz[i] = v[2];
    }
    if (a.mcc_ctl && a.tet_seg) wave_seg_max(mseg, mv2, a.mcc_v2max);
    if (a.mcc_ctl) {
      // wave-aggregated append of this chunk's candidates
      const uint64_t m = __ballot(is_cand);
      if (m) {
        const int lane = threadIdx.x & 63;
        const int leader = __ffsll((unsigned long long)m) - 1;
        unsigned int base = 0;
        if (lane == leader) base = atomicAdd(a.mcc_ncand, (unsigned int)__popcll(m));
        base = __shfl(base, leader, 64);
        if (is_cand) a.mcc_cand[base + (unsigned int)__popcll(m & ((1ull << lane) - 1ull))] = cand;
      }
    }
    tile_runs(key, f, a.cs, a.tets, a.rho, tile);
  }
  if (a.mcc_ctl && !a.tet_seg) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v2max = fmax(v2max, __shfl_xor(v2max, off, 64));
    if ((threadIdx.x & 63) == 0 && v2max > 0.0)
      atomicMax(a.mcc_v2max, (unsigned long long)__double_as_longlong(v2max));
  }
  __syncthreads();
  if (a.dep.count) tile_flush_buckets(tile, a.rho, a.dep);
  else tile_flush(tile, a.rho);
  if (threadIdx.x < kMaxGroups) {
    const unsigned int n = led_n[threadIdx.x];
    if (n) {
      atomicAdd(&a.ledger_count[threadIdx.x * kMaxSpecies + a.species], (unsigned long long)n);
      unsafeAtomicAdd(&a.ledger_energy[threadIdx.x * kMaxSpecies + a.species], led_e[threadIdx.x]);
    }
  }
  // per-block kinetic energy, alive and started counts (deterministic partials)
  __shared__ double lds[kBlock / 64];
  const double ks = block_sum(ke, lds);
  const double as = block_sum((double)alive, lds);
  const double ss = block_sum((double)started, lds);
  if (threadIdx.x == 0) {
    a.ke_partial[lb] = ks;
    a.alive_partial[lb] = (int32_t)as;
    if (ss > 0) atomicAdd(a.pushes, (unsigned long long)ss);
  }
}

// Deposit only (no motion): integer weights of the current positions.
template <typename T>
__global__ void __launch_bounds__(kBlock) deposit_kernel(PushArgs<T> a) {
  const uint32_t lb = xcd_remap(blockIdx.x, gridDim.x);
  __shared__ DepositTile tile;
  tile_clear(tile);
  __syncthreads();
#pragma unroll 1
  for (int ch = 0; ch < 4; ++ch) {
    const int64_t i = ((int64_t)lb * 4 + ch) * kBlock + threadIdx.x;
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
    tile_runs(key, f, a.cs, a.tets, a.rho, tile);
  }
  __syncthreads();
  tile_flush(tile, a.rho);
}

// One block per 512-row range: its bucket summed in LDS, added to its rows.
__global__ void __launch_bounds__(kDepRangeRows) k_dep_reduce(DepBuckets b, unsigned long long* rho, int64_t n_rows) {
  __shared__ unsigned long long acc[kDepRangeRows];
  acc[threadIdx.x] = 0ull;
  const unsigned int n = min(b.count[blockIdx.x], (unsigned int)b.cap);
  __syncthreads();
  const DepEntry* e = b.ent + (size_t)blockIdx.x * (size_t)b.cap;
  for (unsigned int q = threadIdx.x; q < n; q += kDepRangeRows) {
    const DepEntry d = e[q];
    atomicAdd(&acc[d.node - (int32_t)blockIdx.x * kDepRangeRows], d.val);
  }
  __syncthreads();
  const int64_t row = (int64_t)blockIdx.x * kDepRangeRows + threadIdx.x;
  if (row < n_rows && acc[threadIdx.x]) rho[row] += acc[threadIdx.x];
  if (threadIdx.x == 0) b.count[blockIdx.x] = 0u;
}

void launch_dep_reduce(const DepBuckets& b, unsigned long long* rho, int64_t n_rows, hipStream_t s) {
  if (!b.count || n_rows <= 0) return;
  const int64_t nr = (n_rows + kDepRangeRows - 1) / kDepRangeRows;
  hipLaunchKernelGGL(k_dep_reduce, dim3((uint32_t)nr), dim3(kDepRangeRows), 0, s, b, rho, n_rows);
}

int push_blocks(int64_t n, int chunks) {
  return (int)((n + (int64_t)kBlock * chunks - 1) / ((int64_t)kBlock * chunks));
}

template <typename T, bool LAT>
void launch_push_shape(const PushArgs<T>& a, dim3 g, dim3 b, hipStream_t s, hipEvent_t start, hipEvent_t stop) {
  if constexpr (LAT && sizeof(T) != 4) {
    return;  // the lattice path is FP32 only
  } else {
    if (!start && !stop) {
      // plain launches (stream capture records them as kernel nodes)
      switch (a.chunks) {
        case 1: hipLaunchKernelGGL((push_kernel<T, 1, LAT>), g, b, 0, s, a); break;
        case 2: hipLaunchKernelGGL((push_kernel<T, 2, LAT>), g, b, 0, s, a); break;
        case 8: hipLaunchKernelGGL((push_kernel<T, 8, LAT>), g, b, 0, s, a); break;
        default: hipLaunchKernelGGL((push_kernel<T, 4, LAT>), g, b, 0, s, a); break;
      }
      return;
    }
    switch (a.chunks) {
      case 1: hipExtLaunchKernelGGL((push_kernel<T, 1, LAT>), g, b, 0, s, start, stop, 0, a); break;
      case 2: hipExtLaunchKernelGGL((push_kernel<T, 2, LAT>), g, b, 0, s, start, stop, 0, a); break;
      case 8: hipExtLaunchKernelGGL((push_kernel<T, 8, LAT>), g, b, 0, s, start, stop, 0, a); break;
      default: hipExtLaunchKernelGGL((push_kernel<T, 4, LAT>), g, b, 0, s, start, stop, 0, a); break;
    }
  }
}

template <typename T>
void launch_push(const PushArgs<T>& a, hipStream_t s, hipEvent_t start, hipEvent_t stop) {
  if (a.n <= 0) return;
  const dim3 g(push_blocks(a.n, a.chunks)), b(kBlock);
  if (sizeof(T) == 4 && a.lat.tet_hex) {
    launch_push_shape<T, true>(a, g, b, s, start, stop);
  } else {
    launch_push_shape<T, false>(a, g, b, s, start, stop);
  }
}
template <typename T>
void launch_deposit(const PushArgs<T>& a, hipStream_t s) {
  if (a.n <= 0) return;
  hipLaunchKernelGGL(deposit_kernel<T>, dim3(push_blocks(a.n, 4)), dim3(kBlock), 0, s, a);
}
template void launch_push<float>(const PushArgs<float>&, hipStream_t, hipEvent_t, hipEvent_t);
template void launch_push<double>(const PushArgs<double>&, hipStream_t, hipEvent_t, hipEvent_t);
template void launch_deposit<float>(const PushArgs<float>&, hipStream_t);
template void launch_deposit<double>(const PushArgs<double>&, hipStream_t);

}  // namespace scdat
