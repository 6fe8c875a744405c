// Counting sort of one species by cell, removing absorbed particles.
// Replaces ParticleContainer::removeInactiveParticles
// (Tools/Particle/src/ParticleManager.cpp:157-177) and keeps the SoA arrays in
// cell order, which is what lets the fused push kernel share tet records and
// reduce deposition within a wavefront.  The cells are ranked by a key table
// (key[cell], a permutation of the cells; null: the tet index), so the order
// of the cells along the sorted arrays can follow a space-filling curve.
//   1. histogram of alive particles per cell rank (32-bit atomics)
//   2. exclusive scan of the histogram (block scan + block totals + add)
//   3. scatter of every SoA field to its slot (offset[rank] + atomic cursor)
// The order inside one cell depends on atomic arrival; no result depends on
// it (per-particle updates are independent, deposition is integer), and the
// particle id travels with the particle.
#include "device_common.hpp"
#include "sort.hpp"

namespace scdat {

namespace {
constexpr int kScanPer = 4;  // elements per thread in the block scan
constexpr int kScanTile = kBlock * kScanPer;

__global__ void __launch_bounds__(kBlock) k_zero_u32(uint32_t* p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) p[i] = 0u;
}

// Runs of equal cell inside a wavefront (particles are nearly cell-sorted):
// first lane of this lane's run and the run's length.  Every lane of the
// wavefront must call it.
__device__ __forceinline__ void wave_run(int key, int& start, int& len) {
  const int lane = threadIdx.x & 63;
  const int prev = __shfl_up(key, 1, 64);
  const bool head = lane == 0 || prev != key;
  const uint64_t heads = __ballot(head);
  const uint64_t le = lane == 63 ? ~0ull : ((1ull << (lane + 1)) - 1ull);
  start = 63 - __builtin_clzll(heads & le);
  const uint64_t above = heads & ~le;
  len = (above ? __builtin_ctzll(above) : 64) - start;
}

__device__ __forceinline__ int cell_rank(const int32_t* key, int c) { return key ? key[c] : c; }

// Particles per cell: one atomic per run of equal cell.
__global__ void __launch_bounds__(kBlock) k_hist(const int32_t* cell, const int32_t* key, int64_t n,
                                                  const int64_t* n_dev, uint32_t* cnt) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (n_dev) n = *n_dev;
  const int c = i < n ? cell[i] : -1;
  int start, len;
  wave_run(c, start, len);
  if (c >= 0 && (int)(threadIdx.x & 63) == start) atomicAdd(&cnt[cell_rank(key, c)], (uint32_t)len);
}

// Exclusive scan of one tile of 1024 counts; tile total to tot[blockIdx.x].
__global__ void __launch_bounds__(kBlock) k_scan_tile(const uint32_t* in, uint32_t* out, int64_t n,
                                                       uint32_t* tot) {
  __shared__ uint32_t s[kBlock];
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanPer;
  uint32_t v[kScanPer];
  uint32_t sum = 0;
#pragma unroll
  for (int k = 0; k < kScanPer; ++k) {
    v[k] = (base + k < n) ? in[base + k] : 0u;
    sum += v[k];
  }
  s[threadIdx.x] = sum;
  __syncthreads();
  for (int off = 1; off < kBlock; off <<= 1) {
    const uint32_t t = threadIdx.x >= (unsigned)off ? s[threadIdx.x - off] : 0u;
    __syncthreads();
    s[threadIdx.x] += t;
    __syncthreads();
  }
  uint32_t run = s[threadIdx.x] - sum;  // exclusive prefix of this thread
#pragma unroll
  for (int k = 0; k < kScanPer; ++k) {
    if (base + k < n) out[base + k] = run;
    run += v[k];
  }
  if (threadIdx.x == kBlock - 1) tot[blockIdx.x] = s[kBlock - 1];
}

// Single-block exclusive scan of the tile totals (any count, serial chunks).
__global__ void __launch_bounds__(kBlock) k_scan_totals(uint32_t* tot, int64_t nt,
                                                         uint32_t* grand) {
  __shared__ uint32_t s[kBlock];
  __shared__ uint32_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int64_t base = 0; base < nt; base += kBlock) {
    const int64_t i = base + threadIdx.x;
    const uint32_t v = i < nt ? tot[i] : 0u;
    s[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < kBlock; off <<= 1) {
      const uint32_t t = threadIdx.x >= (unsigned)off ? s[threadIdx.x - off] : 0u;
      __syncthreads();
      s[threadIdx.x] += t;
      __syncthreads();
    }
    if (i < nt) tot[i] = carry + s[threadIdx.x] - v;
    __syncthreads();
    if (threadIdx.x == kBlock - 1) carry += s[kBlock - 1];
    __syncthreads();
  }
  if (threadIdx.x == 0) *grand = carry;
}

__global__ void __launch_bounds__(kBlock) k_scan_add(uint32_t* out, int64_t n, const uint32_t* tot,
                                                      uint32_t* cursor) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const uint32_t o = out[i] + tot[i / kScanTile];
  out[i] = o;
  cursor[i] = 0u;
}

// Destination slot of every alive particle: the cell's offset plus a rank
// taken one atomic per run; only the 4-byte source index is scattered.  The
// particle is also packed into a record at its source slot (streamed).
template <typename T>
__global__ void __launch_bounds__(kBlock) k_rank(SortArgs<T> a) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const int64_t n = a.n_dev ? *a.n_dev : a.n;
  const int c = i < n ? a.src.cell[i] : -1;
  int start, len;
  wave_run(c, start, len);
  const int lane = threadIdx.x & 63;
  const int k = c >= 0 ? cell_rank(a.key, c) : -1;
  uint32_t base = 0u;
  if (c >= 0 && lane == start) base = atomicAdd(&a.cursor[k], (uint32_t)len);
  base = __shfl(base, start, 64);
  if (c >= 0) {
    a.perm[a.offset[k] + base + (uint32_t)(lane - start)] = (uint32_t)i;
    SortRec<T> r;
    r.x = a.src.x[i];
    r.y = a.src.y[i];
    r.z = a.src.z[i];
    r.vx = a.src.vx[i];
    r.vy = a.src.vy[i];
    r.vz = a.src.vz[i];
    r.cell = c;
    r.id = a.src.id[i];
    a.stage[i] = r;
  }
}

// The particles in destination order: one record read per particle (the
// random side), eight streamed writes.  Scattering the eight arrays directly
// wrote a partial line per field and particle.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_gather(SortArgs<T> a) {
  const int64_t d = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (d >= (int64_t)*a.total) return;
  const SortRec<T> r = a.stage[a.perm[d]];
  a.dst.x[d] = r.x;
  a.dst.y[d] = r.y;
  a.dst.z[d] = r.z;
  a.dst.vx[d] = r.vx;
  a.dst.vy[d] = r.vy;
  a.dst.vz[d] = r.vz;
  a.dst.cell[d] = r.cell;
  a.dst.id[d] = r.id;
}

__global__ void k_store_count(const uint32_t* total, int64_t* n_dev) { *n_dev = (int64_t)*total; }

inline uint32_t blocks(int64_t n, int64_t per) { return (uint32_t)((n + per - 1) / per); }
}  // namespace

template <typename T>
void launch_sort(const SortArgs<T>& a, hipStream_t s) {
  // zero-fill by kernel (no memset nodes, so the sort can be graph-captured)
  hipLaunchKernelGGL(k_zero_u32, dim3(blocks(a.ncell, kBlock)), dim3(kBlock), 0, s, a.count, a.ncell);
  if (a.n > 0) hipLaunchKernelGGL(k_hist, dim3(blocks(a.n, kBlock)), dim3(kBlock), 0, s, a.src.cell, a.key, a.n,
                                (const int64_t*)a.n_dev, a.count);
  const uint32_t nt = blocks(a.ncell, kScanTile);
  hipLaunchKernelGGL(k_scan_tile, dim3(nt), dim3(kBlock), 0, s, a.count, a.offset, a.ncell, a.tiles);
  hipLaunchKernelGGL(k_scan_totals, dim3(1), dim3(kBlock), 0, s, a.tiles, (int64_t)nt, a.total);
  hipLaunchKernelGGL(k_scan_add, dim3(blocks(a.ncell, kBlock)), dim3(kBlock), 0, s, a.offset, a.ncell,
                     a.tiles, a.cursor);
  if (a.n > 0) {
    hipLaunchKernelGGL(k_rank<T>, dim3(blocks(a.n, kBlock)), dim3(kBlock), 0, s, a);
    hipLaunchKernelGGL(k_gather<T>, dim3(blocks(a.n, kBlock)), dim3(kBlock), 0, s, a);
  }
  if (a.n_dev) hipLaunchKernelGGL(k_store_count, dim3(1), dim3(1), 0, s, (const uint32_t*)a.total, a.n_dev);
}

template void launch_sort<float>(const SortArgs<float>&, hipStream_t);
template void launch_sort<double>(const SortArgs<double>&, hipStream_t);

size_t sort_tiles(int64_t ncell) { return (size_t)blocks(ncell, kScanTile); }

}  // namespace scdat
