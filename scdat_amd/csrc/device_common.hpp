// Device-side data layout and per-particle math of the PIC-MCC cycle (gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <cstdint>

namespace scdat {

constexpr int kMaxGroups = 16;
constexpr int kMaxSpecies = 8;
constexpr int kBlock = 256;
constexpr int kLambdaBits = 24;  // fixed-point barycentric weight scale 2^24
constexpr int32_t kLambdaOne = 1 << kLambdaBits;

enum : int { BC_ABSORB = 0, BC_PERIODIC = 1, BC_REFLECT = 2, BC_CONDUCTOR = 3 };

// Affine barycentric map of one tet plus its node ids: lambda_{1..3} =
// a (x - x0), lambda_0 = 1 - sum.  One 64-byte record per tet (FP32), read
// once per particle-step and shared by all particles of the tet.
template <typename T>
struct TetMap;
template <>
struct alignas(64) TetMap<float> {
  float a[9];
  float x0[3];
  int32_t node[4];
};
template <>
struct alignas(128) TetMap<double> {
  double a[9];
  double x0[3];
  int32_t node[4];
  int32_t pad[4];
};

// Walk record of the FP32 push: the tet's barycentric map with its four face
// neighbours instead of its nodes, one 64-byte line, so a face crossing is a
// single dependent load (the next tet's record) rather than two (neighbour
// ids, then the next tet's map).
template <typename T>
struct WalkMap;
template <>
struct alignas(64) WalkMap<float> {
  float a[9];
  float x0[3];
  int32_t nb[4];
};

template <typename T>
struct Vec4;
template <>
struct alignas(16) Vec4<float> {
  float x, y, z, w;
};
template <>
struct alignas(32) Vec4<double> {
  double x, y, z, w;
};

template <typename T>
__host__ __device__ inline T fma_t(T a, T b, T c);
template <>
__host__ __device__ inline float fma_t<float>(float a, float b, float c) {
  return fmaf(a, b, c);
}
template <>
__host__ __device__ inline double fma_t<double>(double a, double b, double c) {
  return fma(a, b, c);
}

// Barycentric coordinates of p in tet m.  The fma form is part of the
// fixed-point deposition contract: the host checker evaluates the identical
// expression, so the integer weights agree bit for bit.
template <typename M, typename T>
__device__ __forceinline__ void bary(const M& m, const T p[3], T lam[4]) {
  const T d0 = p[0] - m.x0[0], d1 = p[1] - m.x0[1], d2 = p[2] - m.x0[2];
  lam[1] = fma_t(m.a[0], d0, fma_t(m.a[1], d1, m.a[2] * d2));
  lam[2] = fma_t(m.a[3], d0, fma_t(m.a[4], d1, m.a[5] * d2));
  lam[3] = fma_t(m.a[6], d0, fma_t(m.a[7], d1, m.a[8] * d2));
  lam[0] = T(1) - lam[1] - lam[2] - lam[3];
}

// Integer barycentric weights: lambda_1..3 clamped to [0,1], truncated to
// 2^-24, lambda_0 takes the remainder so the four weights sum to exactly 2^24
// (charge is conserved exactly in the integer accumulator).
template <typename T>
__device__ __forceinline__ void lambda_fixed(const T lam[4], int32_t f[4]) {
  int32_t s = 0;
#pragma unroll
  for (int k = 1; k < 4; ++k) {
    T l = lam[k];
    l = l < T(0) ? T(0) : (l > T(1) ? T(1) : l);
    f[k] = (int32_t)(l * T(kLambdaOne));
    s += f[k];
  }
  f[0] = kLambdaOne - s;
}

// Boris velocity update (PICParticleAlgorithms.cpp:803-851): half E kick,
// rotation about B when |B| > 1e-20 and |t|^2 < 100, half E kick.
template <typename T>
__device__ __forceinline__ void boris(T v[3], const T E[3], const T B[3], bool has_b, T qm, T dt) {
  const T h = T(0.5) * dt * qm;
  T vm[3] = {v[0] + E[0] * h, v[1] + E[1] * h, v[2] + E[2] * h};
  if (has_b) {
    const T t[3] = {B[0] * h, B[1] * h, B[2] * h};
    const T t2 = t[0] * t[0] + t[1] * t[1] + t[2] * t[2];
    if (t2 < T(100)) {
      const T sf = T(2) / (T(1) + t2);
      const T s[3] = {t[0] * sf, t[1] * sf, t[2] * sf};
      const T vs[3] = {vm[0] + (vm[1] * t[2] - vm[2] * t[1]), vm[1] + (vm[2] * t[0] - vm[0] * t[2]),
                       vm[2] + (vm[0] * t[1] - vm[1] * t[0])};
      const T vp[3] = {vm[0] + (vs[1] * s[2] - vs[2] * s[1]), vm[1] + (vs[2] * s[0] - vs[0] * s[2]),
                       vm[2] + (vs[0] * s[1] - vs[1] * s[0])};
      vm[0] = vp[0];
      vm[1] = vp[1];
      vm[2] = vp[2];
    }
  }
  v[0] = vm[0] + E[0] * h;
  v[1] = vm[1] + E[1] * h;
  v[2] = vm[2] + E[2] * h;
}

// Adaptive sub-step (PICParticleAlgorithms.cpp:750-801).  v0 is the
// start-of-step velocity, as in the reference.
template <typename T>
__device__ __forceinline__ T adaptive_dt(const T v0[3], const T E[3], T bmag, T qm, T dt_max,
                                T floor_ratio) {
  T dt = dt_max;
  const T emag = sqrt(E[0] * E[0] + E[1] * E[1] + E[2] * E[2]);
  const T aq = qm < T(0) ? -qm : qm;
  if (emag > T(1e-10)) {
    const T acc = emag * aq;
    const T vel = sqrt(v0[0] * v0[0] + v0[1] * v0[1] + v0[2] * v0[2]);
    if (acc > T(1e-10) && vel > T(1)) {
      const T dv = T(0.1) * vel / acc;
      dt = dv < dt ? dv : dt;
    }
  }
  if (bmag > T(1e-10)) {
    const T wc = bmag * aq;
    if (wc > T(1e-10)) {
      const T dc = T(0.1) * T(2.0 * 3.14159265358979323846) / wc;
      dt = dc < dt ? dc : dt;
    }
  }
  const T fl = dt_max * floor_ratio;
  dt = dt > fl ? dt : fl;
  return dt < dt_max ? dt : dt_max;
}

// XCD-aware block order: the dispatcher hands workgroup b to XCD b % 8, so
// logical chunks are remapped to give each XCD one contiguous 1/8 of the
// range (neighbouring particles share tets and nodes in that XCD's L2).
__device__ inline uint32_t xcd_remap(uint32_t b, uint32_t g) {
  const uint32_t q = g >> 3, r = g & 7u;
  const uint32_t x = b & 7u, idx = b >> 3;
  return x * q + (x < r ? x : r) + idx;
}

// Atomic max of v (>= 0, FP64 bits ordered as integers) into out[s] for the
// lanes of a wavefront, one atomic per distinct s (s < 0: no value).  Every
// lane must call it.
__device__ __forceinline__ void wave_seg_max(int s, double v, unsigned long long* out) {
  const int lane = threadIdx.x & 63;
  uint64_t pending = __ballot(s >= 0);
  while (pending) {
    const int leader = __ffsll((unsigned long long)pending) - 1;
    const int s0 = __shfl(s, leader, 64);
    double m = s == s0 ? v : 0.0;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
    if (lane == leader && m > 0.0) atomicMax(&out[s0], (unsigned long long)__double_as_longlong(m));
    pending &= ~__ballot(s == s0);
  }
}

// Deterministic block sum of doubles (wave64 shuffles, then LDS across waves).
__device__ inline double block_sum(double v, double* lds) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  __syncthreads();
  if (lane == 0) lds[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + 63) >> 6;
    for (int i = 0; i < nw; ++i) s += lds[i];
  }
  return s;  // valid in thread 0
}

}  // namespace scdat
