// Background Monte Carlo collisions, secondary-electron emission and
// photoemission on gfx950.
//
// Collisions replace MonteCarloCollisionHandler::processCollisions
// (Tools/Interactions/Collisions/src/CollisionAlgorithm.cpp:500-538) inside
// executeBackgroundMccStep (CollisionPicAdapter.cpp:338-434).  Per particle:
// nu_p = n_p sigma_p(E) |v| with the cross sections of CollisonTypes.cpp:56-176,
// one collision with probability 1 - exp(-sum nu dt) and the process chosen
// with probability nu_p / sum nu.  The null-collision method selects the
// candidates: the push tests every particle against P_null = 1 -
// exp(-nu_max dt) of a constant majorant nu_max (k_mcc_majorant) with one
// Philox draw and no cross section, and only candidates reach k_mcc, which
// accepts with probability (1 - exp(-nu dt)) / P_null.  Outcomes follow CollisionAlgorithm.cpp:
// 189-460: isotropic scatter (elastic, Coulomb), energy loss (inelastic,
// excitation above threshold), ionisation (uniform split of E - I, secondary
// electron and ion at rest at the parent's position), removal (charge
// exchange, recombination).  The reference's second draw inside
// executeCollision (effective probability p^2) is not reproduced.
//
// All arithmetic of these rare events is FP64 whatever the particle storage
// precision; random numbers come from Philox keyed by (seed, particle id,
// step, species), so results do not depend on thread order.  New particles
// take their slots from the species' append counter (one atomic per
// ionisation product, per impact for its secondaries, per wavefront for
// photoelectrons) and are deposited into the step's charge at once (fixed
// point, so the order of deposits does not matter).
#include <algorithm>
#include <cstddef>

#include "emit.hpp"
#include "rng.hpp"

namespace scdat {

namespace {

constexpr double kE = 1.602176634e-19;
constexpr double kEps0 = 8.8541878128e-12;
constexpr double kPi = 3.14159265358979323846;

__device__ double sigma_of(const MccProcDev& p, double E) {
  switch (p.type) {
    case MCC_ELASTIC:
    case MCC_INELASTIC:
    case MCC_CHARGE_EXCHANGE: {
      if (E < p.thr) return 0.0;
      const double sg = p.smax * pow(fmax(1e-6, p.eref) / E, p.pidx);
      return fmax(0.0, fmin(sg, p.smax));
    }
    case MCC_IONIZATION: {
      if (E <= p.thr) return 0.0;
      const double I = fmax(1e-6, p.thr);
      const double sg = p.smax * log(E / I) * (1.0 - I / E) / E;
      return fmax(0.0, fmin(sg, p.smax));
    }
    case MCC_EXCITATION: {
      if (E <= p.thr) return 0.0;
      const double x = E / fmax(1e-6, p.thr);
      const double sg = p.smax * pow(fmax(0.0, 1.0 - 1.0 / x), p.alpha) * pow(1.0 / x, p.beta);
      return fmax(0.0, fmin(sg, p.smax));
    }
    case MCC_RECOMBINATION: {
      const double er = fmax(1e-6, p.eref);
      const double sg = p.smax * pow(er / fmax(er, E), p.pidx);
      return fmax(0.0, fmin(sg, p.smax));
    }
    default: {
      if (E <= 0.0) return 0.0;
      const double Ej = E * kE;
      const double ke2 = (1.0 / (4.0 * kPi * kEps0)) * kE * kE;
      const double sg = kPi * p.clog * ke2 * ke2 / fmax(1e-30, 4.0 * Ej * Ej);
      return fmax(0.0, fmin(sg, p.smax));
    }
  }
}

__device__ __forceinline__ void iso_dir(double u2, double u3, double d[3]) {
  const double ct = 2.0 * u2 - 1.0;
  const double st = sqrt(fmax(0.0, 1.0 - ct * ct));
  const double ph = 2.0 * kPi * u3;
  d[0] = st * cos(ph);
  d[1] = st * sin(ph);
  d[2] = ct;
}

// Fixed-point deposit of one particle (the push kernel's weights).
template <typename T>
__device__ void deposit_one(const TetMap<T>* tets, unsigned long long* rho, int cell, const T p[3],
                            int64_t cs) {
  const TetMap<T> m = tets[cell];
  T lam[4];
  bary(m, p, lam);
  int32_t f[4];
  lambda_fixed(lam, f);
#pragma unroll
  for (int k = 0; k < 4; ++k)
    atomicAdd(&rho[m.node[k]], (unsigned long long)(cs * (int64_t)f[k]));
}

// Appends one particle; returns false (and counts an overflow) when full.
template <typename T>
__device__ bool append_at(const SpeciesOut<T>& o, int64_t r, const T x[3], const double v[3], int cell, int32_t id,
                          const TetMap<T>* tets, unsigned long long* rho, unsigned long long* counts) {
  const int64_t slot = (o.base_dev ? *o.base_dev : o.base) + r;
  if (slot >= o.cap) {
    atomicAdd(&counts[C_OVERFLOW], 1ull);
    return false;
  }
  o.x[slot] = x[0];
  o.y[slot] = x[1];
  o.z[slot] = x[2];
  o.vx[slot] = (T)v[0];
  o.vy[slot] = (T)v[1];
  o.vz[slot] = (T)v[2];
  o.cell[slot] = cell;
  o.id[slot] = id;
  deposit_one(tets, rho, cell, x, o.cs);
  return true;
}
// Appends one particle at the next free slot (r: index r among this step's
// appends to the species, reserved from o.append).
template <typename T>
__device__ bool append(const SpeciesOut<T>& o, const T x[3], const double v[3], int cell, int32_t id,
                       const TetMap<T>* tets, unsigned long long* rho, unsigned long long* counts) {
  return append_at(o, (int64_t)atomicAdd(o.append, 1ull), x, v, cell, id, tets, rho, counts);
}

// Total collision frequency sum_p n_p sigma_p(E) sqrt(2E/m) of a species;
// dens non-null: n_p = dens[slot of p] (one scenario's densities).
__device__ double nu_total(const MccProcDev* proc, int np, double E, double mass, const double* dens) {
  double t = 0.0;
  for (int j = 0; j < np; ++j) t += (dens ? dens[proc[j].slot] : proc[j].n) * sigma_of(proc[j], E);
  return t * sqrt(2.0 * E * kE / mass);
}

// The tables of the kernel arguments that a lane indexes with its own value
// (process, species, group) are copied to LDS first: indexing the argument
// block per lane, or passing it by reference to a call, copies the whole
// block to scratch memory in every thread.  The block reads bytes [off, off +
// bytes) of its argument segment one dword per lane (the kernels take one
// argument, which starts the segment with the host struct's layout).
template <int NT, size_t OFF, size_t BYTES>
__device__ __forceinline__ void stage_args(void* dst) {
  static_assert(OFF % 4 == 0 && BYTES % 4 == 0, "whole dwords");
  const uint32_t* src =
      reinterpret_cast<const uint32_t*>((const char*)__builtin_amdgcn_kernarg_segment_ptr() + OFF);
  uint32_t* d = static_cast<uint32_t*>(dst);
  constexpr int kWords = (int)(BYTES / 4);
  uint32_t v[(kWords + NT - 1) / NT];
#pragma unroll
  for (int k = 0; k * NT < kWords; ++k) {
    const int t = k * NT + (int)threadIdx.x;
    v[k] = src[t < kWords ? t : 0];
  }
#pragma unroll
  for (int k = 0; k * NT < kWords; ++k) {
    const int t = k * NT + (int)threadIdx.x;
    if (t < kWords) d[t] = v[k];
  }
}
// field of the kernel's argument struct Args to LDS, by a block of NT threads
#define SCDAT_STAGE(NT, Args, field, dst) stage_args<NT, offsetof(Args, field), sizeof(Args::field)>(dst)

// out[s] for an index s uniform over the wavefront, with compile-time indices.
template <typename T>
__device__ __forceinline__ SpeciesOut<T> out_at(const SpeciesOut<T> (&out)[kMaxSpecies], int s) {
  SpeciesOut<T> o = out[0];
#pragma unroll
  for (int q = 1; q < kMaxSpecies; ++q)
    if (q == s) o = out[q];
  return o;
}

// Slots for the lanes of a wavefront that want one (the same counter):
// one atomic per wavefront; every lane taking part must call it.
__device__ __forceinline__ int64_t wave_reserve(unsigned long long* counter, bool want) {
  const uint64_t m = __ballot(want);
  if (!m) return -1;
  const int lane = threadIdx.x & 63;
  const int leader = __ffsll((unsigned long long)m) - 1;
  unsigned long long base = 0ull;
  if (lane == leader) base = atomicAdd(counter, (unsigned long long)__popcll(m));
  base = __shfl(base, leader, 64);
  return want ? (int64_t)(base + (unsigned long long)__popcll(m & ((1ull << lane) - 1ull))) : -1;
}

// Adds the wavefront's sum of v to *p with one atomic; every lane of the
// wavefront must call it.
__device__ __forceinline__ void wave_add(unsigned long long* p, unsigned long long v) {
  unsigned long long t = v;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off, 64);
  const uint64_t m = __ballot(true);
  if ((int)(threadIdx.x & 63) == __ffsll((unsigned long long)m) - 1 && t) atomicAdd(p, t);
}

// One block per species: nu_max over log-spaced energies up to the cap (and
// just above every threshold, where a cross section may switch on), plus
// the bound sum n sigma_max v over the energies below the lowest sample;
// times kMccMajorantMargin.  Then P_null, and the push's accumulators reset.
// One energy per thread: the FP64 cross sections are a long dependent chain
// per energy, and the block is alone on its CU.
constexpr int kMajBlock = kMccMajorantPoints;
static_assert(kMajBlock <= 1024 && kMajBlock % 64 == 0, "one energy per thread of one block");
__global__ void __launch_bounds__(kMajBlock) k_mcc_majorant(MccMajorantArgs a) {
  __shared__ double lds[kMajBlock / 64];
  __shared__ MccProcDev sp[kMaxProc];
  SCDAT_STAGE(kMajBlock, MccMajorantArgs, proc, sp);
  __syncthreads();
  // one block per scenario (one block without scenarios)
  const int sg = blockIdx.x;
  const double* dens = a.seg_n ? a.seg_n + (size_t)sg * kMaxProc : nullptr;
  const double dt = a.seg_n ? a.seg_dt[sg] : a.dt;
  unsigned long long* v2p = a.v2max + sg;
  double* ctl = a.ctl + (size_t)sg * kMccCtl;
  const double v2max = __longlong_as_double((long long)*v2p);
  const double v2cap = kMccCapFactor * v2max;
  const double ecap = 0.5 * a.mass * v2cap / kE;
  double f = 0.0;
  if (ecap > 0.0) {
    const double elo = 1e-8 * ecap, la = log(elo), lb = log(ecap);
    for (int k = threadIdx.x; k < kMccMajorantPoints; k += kMajBlock) {
      const double E = exp(la + (lb - la) * (double)k / (double)(kMccMajorantPoints - 1));
      f = fmax(f, nu_total(sp, a.np, E, a.mass, dens));
    }
    if ((int)threadIdx.x < a.np) {
      const double thr = sp[threadIdx.x].thr;
      if (thr > elo && thr < ecap) f = fmax(f, nu_total(sp, a.np, thr * (1.0 + 1e-9), a.mass, dens));
    }
    if (threadIdx.x == 0) {
      double b = 0.0;
      for (int j = 0; j < a.np; ++j) b += (dens ? dens[sp[j].slot] : sp[j].n) * sp[j].smax;
      f = fmax(f, b * sqrt(2.0 * elo * kE / a.mass));
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) f = fmax(f, __shfl_xor(f, off, 64));
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = f;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kMajBlock / 64; ++w) f = fmax(f, lds[w]);
    f = fmax(f, lds[0]);
    const double numax = kMccMajorantMargin * f;
    ctl[0] = -expm1(-numax * dt);
    ctl[1] = v2cap;
    ctl[2] = numax;
    *v2p = 0ull;
    if (sg == 0) *a.n_cand = 0u;
  }
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_v2max(const T* vx, const T* vy, const T* vz, const int32_t* cell,
                                                  int64_t n, unsigned long long* out, const int32_t* tet_seg) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  double v2 = 0.0;
  const int c = i < n ? cell[i] : -1;
  if (c >= 0) v2 = (double)vx[i] * vx[i] + (double)vy[i] * vy[i] + (double)vz[i] * vz[i];
  wave_seg_max(c >= 0 ? (tet_seg ? tet_seg[c] : 0) : -1, v2, out);
}

// One candidate: the exact collision test given the pre-selection, process
// choice, outcome.  sp, so: the process and output tables in LDS.
template <typename T>
__device__ __forceinline__ void mcc_candidate(const MccArgs<T>& a, const MccProcDev* sp, const SpeciesOut<T>* so,
                                              uint32_t entry) {
  const bool forced = (entry & kMccForced) != 0u;
  const int64_t i = (int64_t)(entry & ~kMccForced);
  const int cell = a.cell[i];
  if (cell < 0) return;
  atomicAdd(&a.counts[C_CAND], 1ull);
  // one scenario's densities, time step, key and majorant
  const int sg = a.tet_seg ? a.tet_seg[cell] : 0;
  const double* dens = a.tet_seg ? a.seg_n + (size_t)sg * kMaxProc : nullptr;
  const double dt = a.tet_seg ? a.seg_dt[sg] : a.dt;
  const uint32_t key0 = a.tet_seg ? a.seg_key[2 * sg] : a.key0, key1 = a.tet_seg ? a.seg_key[2 * sg + 1] : a.key1;
  double v[3] = {(double)a.vx[i], (double)a.vy[i], (double)a.vz[i]};
  const double v2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
  const double speed = sqrt(v2);
  const double E = 0.5 * a.mass * v2 / kE;
  double nu[kMaxProc];
  double tot = 0.0;
#pragma unroll
  for (int j = 0; j < kMaxProc; ++j) {
    nu[j] = j < a.np ? (dens ? dens[sp[j].slot] : sp[j].n) * sigma_of(sp[j], E) * speed : 0.0;
    tot += nu[j];
  }
  if (!(tot > 0.0)) return;
  const uint32_t id = (uint32_t)a.id[i];
  const uint32_t step = a.step_dev ? *a.step_dev : a.step;
  const U4 r = philox4x32(id, step, (uint32_t)a.species, kTagMccA, key0, key1);
  // accept with probability pex / P_c: overall probability pex = 1 - exp(-nu dt)
  const double pc = forced ? 1.0 : a.ctl[(size_t)sg * kMccCtl];
  const double pex = -expm1(-tot * dt);
  if (pex > pc) atomicAdd(&a.counts[C_MAJOR], 1ull);
  const double u = u01(r.y) * pc;
  if (!(u < pex)) return;
  // the first process whose cumulative frequency exceeds R (zero frequencies
  // are never chosen); u / pex is uniform on (0, 1) given the collision
  const double R = (u / pex) * tot;
  int j = -1;
  double cum = 0.0;
  bool stop = false;
#pragma unroll
  for (int q = 0; q < kMaxProc; ++q) {
    if (!stop && q < a.np) {
      cum += nu[q];
      if (nu[q] > 0.0) j = q;
      if (R < cum) stop = true;
    }
  }
  const MccProcDev& p = sp[j];
  atomicAdd(&a.counts[p.slot], 1ull);
  double d[3];
  switch (p.type) {
    case MCC_ELASTIC:
    case MCC_COULOMB:
      iso_dir(u01(r.z), u01(r.w), d);
      a.vx[i] = (T)(speed * d[0]);
      a.vy[i] = (T)(speed * d[1]);
      a.vz[i] = (T)(speed * d[2]);
      break;
    case MCC_INELASTIC:
    case MCC_EXCITATION: {
      if (p.type == MCC_EXCITATION && E <= p.thr) break;
      const double loss = p.type == MCC_INELASTIC ? fmax(0.0, p.loss) : (p.loss > 0.0 ? p.loss : p.thr);
      const double sp = sqrt(2.0 * fmax(1e-6, E - loss) * kE / a.mass);
      iso_dir(u01(r.z), u01(r.w), d);
      a.vx[i] = (T)(sp * d[0]);
      a.vy[i] = (T)(sp * d[1]);
      a.vz[i] = (T)(sp * d[2]);
      break;
    }
    case MCC_IONIZATION: {
      if (E <= p.thr) break;
      const U4 b = philox4x32(id, step, (uint32_t)a.species, kTagMccB, key0, key1);
      const U4 c = philox4x32(id, step, (uint32_t)a.species, kTagMccId, key0, key1);
      const double avail = E - p.thr;
      const double sec = u01(b.x) * avail;
      const double prim = avail - sec;
      const double sp = sqrt(2.0 * prim * kE / a.mass);
      iso_dir(u01(r.z), u01(r.w), d);
      a.vx[i] = (T)(sp * d[0]);
      a.vy[i] = (T)(sp * d[1]);
      a.vz[i] = (T)(sp * d[2]);
      const T xp[3] = {a.x[i], a.y[i], a.z[i]};
      const SpeciesOut<T>& oe = so[p.e_species];
      const double se = sqrt(2.0 * sec * kE / oe.m);
      iso_dir(u01(b.y), u01(b.z), d);
      const double ve[3] = {se * d[0], se * d[1], se * d[2]};
      int made = append(oe, xp, ve, cell, gen_id(c.x), a.tets, a.rho, a.counts) ? 1 : 0;
      const double vi[3] = {0.0, 0.0, 0.0};
      made += append(so[p.i_species], xp, vi, cell, gen_id(c.y), a.tets, a.rho, a.counts) ? 1 : 0;
      atomicAdd(&a.counts[C_CREATED], (unsigned long long)made);
      break;
    }
    default:  // charge exchange, recombination
      a.cell[i] = -1;
      atomicAdd(&a.counts[C_REMOVED], 1ull);
      break;
  }
}

// Candidates of one species (grid-stride over the count the push left).
template <typename T>
__global__ void __launch_bounds__(kBlock) k_mcc(MccArgs<T> a) {
  __shared__ MccProcDev sp[kMaxProc];
  __shared__ SpeciesOut<T> so[kMaxSpecies];
  SCDAT_STAGE(kBlock, MccArgs<T>, proc, sp);
  SCDAT_STAGE(kBlock, MccArgs<T>, out, so);
  __syncthreads();
  const unsigned int nc = *a.n_cand;
  for (unsigned int q = blockIdx.x * kBlock + threadIdx.x; q < nc; q += gridDim.x * kBlock)
    mcc_candidate(a, sp, so, a.cand[q]);
}

// ---- surface emission ------------------------------------------------------
__device__ double see_yield(int model, double dmax, double emax, double E, double cos_inc) {
  if (model == 1) {  // SurfaceMaterialModel.cpp:163-170
    const double x = fmax(1e-6, E / fmax(1.0, emax));
    return fmax(0.0, dmax * x * exp(1.0 - x) / sqrt(fmax(0.1, cos_inc)));
  }
  if (model == 2) {  // Sternglass
    if (!(E > 0.0)) return 0.0;
    const double r = E / emax;
    return dmax * r * exp(2.0 - 2.0 * sqrt(r));
  }
  return 0.0;
}

template <typename T>
__device__ void inward_normal(const TetMap<T>& m, int k, double n[3]) {
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    const double g1 = m.a[r], g2 = m.a[3 + r], g3 = m.a[6 + r];
    n[r] = k == 0 ? -(g1 + g2 + g3) : (k == 1 ? g1 : (k == 2 ? g2 : g3));
  }
  const double l = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
#pragma unroll
  for (int r = 0; r < 3; ++r) n[r] /= l;
}

// Surface point moved 1e-6 of the way to the tet centroid (inside the tet).
__device__ void nudge(const double* xyz, const int32_t* tet4, int t, const double p[3], double x[3]) {
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    double c = 0.0;
#pragma unroll
    for (int k = 0; k < 4; ++k) c += xyz[3 * (int64_t)tet4[4 * (int64_t)t + k] + r];
    x[r] = p[r] + 1e-6 * (0.25 * c - p[r]);
  }
}

// Flux-weighted Maxwellian of temperature T (eV) about the inward normal:
// energy -T ln(ua ub), cos(theta) = sqrt(uc), azimuth 2 pi ud.
__device__ void emit_velocity(const double n[3], double T, double m, double ua, double ub, double uc,
                              double ud, double v[3]) {
  const double E = -T * (log(ua) + log(ub));
  const double sp = sqrt(2.0 * E * kE / m);
  const bool xa = fabs(n[0]) < 0.9;
  const double ax[3] = {xa ? 1.0 : 0.0, xa ? 0.0 : 1.0, 0.0};
  double t1[3] = {n[1] * ax[2] - n[2] * ax[1], n[2] * ax[0] - n[0] * ax[2], n[0] * ax[1] - n[1] * ax[0]};
  const double l = sqrt(t1[0] * t1[0] + t1[1] * t1[1] + t1[2] * t1[2]);
#pragma unroll
  for (int r = 0; r < 3; ++r) t1[r] /= l;
  const double t2[3] = {n[1] * t1[2] - n[2] * t1[1], n[2] * t1[0] - n[0] * t1[2], n[0] * t1[1] - n[1] * t1[0]};
  const double ct = sqrt(uc), st = sqrt(1.0 - uc), ph = 2.0 * kPi * ud;
  const double cp = cos(ph), sn = sin(ph);
#pragma unroll
  for (int r = 0; r < 3; ++r) v[r] = sp * (st * cp * t1[r] + st * sn * t2[r] + ct * n[r]);
}

// Secondary emission of one recorded impact.  sg, so, sm: the group,
// output and mass tables in LDS.
template <typename T>
__device__ __forceinline__ void see_event(const EmitArgs<T>& a, const SurfEmitDev* sg, const SpeciesOut<T>* so,
                                          const double* sm, int64_t e) {
  const SeeEvent ev = a.events[e];
  if (!((a.electron_mask >> ev.species) & 1u)) return;  // ion impacts: field emission, on the host
  const int g = ev.group_face & 0xFF, face = ev.group_face >> 8;
  int sgm = 0;
  uint32_t k0 = a.key0, k1 = a.key1;
  if (a.tet_seg) {
    sgm = a.tet_seg[ev.cell];
    k0 = a.seg_key[2 * sgm];
    k1 = a.seg_key[2 * sgm + 1];
  }
  const SurfEmitDev& cf = a.tet_seg ? a.gseg[sgm * kMaxGroups + g] : sg[g];
  const TetMap<T> m = a.tets[ev.cell];
  double n[3];
  inward_normal(m, face, n);
  const double v2 = ev.v[0] * ev.v[0] + ev.v[1] * ev.v[1] + ev.v[2] * ev.v[2];
  const double sp = sqrt(v2);
  const double E = 0.5 * sm[ev.species] * v2 / kE;
  const double cosi = sp > 0.0 ? fabs(ev.v[0] * n[0] + ev.v[1] * n[1] + ev.v[2] * n[2]) / sp : 1.0;
  const double dl = see_yield(cf.see_model, cf.dmax, cf.emax, E, cosi);
  const uint32_t step = a.step_dev ? *a.step_dev : a.step;
  const U4 r = philox4x32((uint32_t)ev.id, step, (uint32_t)ev.species, kTagSee, k0, k1);
  const double fl = floor(dl);
  int nem = (int)fmin(fl, (double)kMaxEmit) + (u01(r.x) < dl - fl ? 1 : 0);
  nem = nem < kMaxEmit ? nem : kMaxEmit;
  if (nem <= 0) return;
  double xe[3];
  nudge(a.xyz, a.tet4, ev.cell, ev.x, xe);
  const T xt[3] = {(T)xe[0], (T)xe[1], (T)xe[2]};
  const SpeciesOut<T>& o = so[cf.e_species];
  int made = 0;
  // the impact's secondaries take consecutive slots (one atomic per impact)
  const int64_t r0 = (int64_t)atomicAdd(o.append, (unsigned long long)nem);
  for (int k = 0; k < nem; ++k) {
    const uint32_t sk = ((uint32_t)ev.species << 8) | (uint32_t)k;
    const U4 b = philox4x32((uint32_t)ev.id, step, sk, kTagSeeK, k0, k1);
    const U4 c = philox4x32((uint32_t)ev.id, step, sk, kTagSeeId, k0, k1);
    double v[3];
    emit_velocity(n, cf.t_see, o.m, u01(b.x), u01(b.y), u01(b.z), u01(b.w), v);
    made += append_at(o, r0 + k, xt, v, ev.cell, gen_id(c.x), a.tets, a.rho, a.counts) ? 1 : 0;
  }
  atomicAdd(&a.emit_count[((size_t)sgm * kMaxGroups + g) * kMaxSpecies + cf.e_species], (unsigned long long)made);
  atomicAdd(&a.counts[C_SEE], (unsigned long long)made);
}

// Impacts of the step, grid-stride over however many the push recorded.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_see(EmitArgs<T> a) {
  __shared__ SurfEmitDev sg[kMaxGroups];
  __shared__ SpeciesOut<T> so[kMaxSpecies];
  __shared__ double sm[kMaxSpecies];
  SCDAT_STAGE(kBlock, EmitArgs<T>, g, sg);
  SCDAT_STAGE(kBlock, EmitArgs<T>, out, so);
  SCDAT_STAGE(kBlock, EmitArgs<T>, mass, sm);
  __syncthreads();
  const unsigned long long ne = *a.n_events;
  const int64_t nev = (int64_t)ne < a.event_cap ? (int64_t)ne : a.event_cap;
  for (int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x; e < nev; e += (int64_t)gridDim.x * kBlock)
    see_event(a, sg, so, sm, e);
}

template <typename T>
__device__ bool photo_one(const EmitArgs<T>& a, const SpeciesOut<T>& o, int64_t r, int64_t k, int sg, int64_t f0,
                          int64_t f1, double area, const SurfEmitDev& cf, uint32_t step, uint32_t key0,
                          uint32_t key1);

// Photoemission of one group, SurfaceMaterialModel.cpp:199-208 current
// density at the group's potential; J A dt / (|q| w) macro-electrons per step.
// With scenarios, grid row blockIdx.y is scenario s: its faces of the group,
// its settings, time step, weight, key and floating potential, so a case
// emits what it emits alone.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_photo(EmitArgs<T> a) {
  const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const int sg = blockIdx.y;
  const bool seg = a.seg_face_off != nullptr;
  const int64_t f0 = seg ? a.seg_face_off[sg] : a.f0, f1 = seg ? a.seg_face_off[sg + 1] : a.f1;
  if (f1 <= f0) return;
  const SurfEmitDev& cf = seg ? a.gseg[sg * kMaxGroups + a.group] : a.g[a.group];
  if (!(cf.photo_j > 0.0)) return;
  const double dt = seg ? a.seg_dt[sg] : a.dt;
  const double qaw = seg ? a.q_abs_w[cf.e_species] * a.seg_wscale[sg] : a.q_abs_w[cf.e_species];
  const uint32_t key0 = seg ? a.seg_key[2 * sg] : a.key0, key1 = seg ? a.seg_key[2 * sg + 1] : a.key1;
  const double area = a.face_cum[f1 - 1];
  const double V = a.is_conductor ? a.vfloat[seg ? sg * a.vfloat_stride : 0] : 0.0;
  const double j = fmax(0.0, cf.photo_j) * exp(-fmax(0.0, V - cf.v_ref) / fmax(1e-3, cf.t_ph));
  const double expect = j * area * dt / qaw;
  const uint32_t step = a.step_dev ? *a.step_dev : a.step;
  const U4 r0 = philox4x32((uint32_t)a.group, step, 0u, kTagPhoto, key0, key1);
  const double fl = floor(expect);
  const int64_t nem = (int64_t)fl + (u01(r0.x) < expect - fl ? 1 : 0);
  if ((int64_t)blockIdx.x * kBlock >= nem) return;  // a block past the count (uniform)
  const bool emit = k < nem;
  // one slot per emitting lane, one atomic per wavefront
  const SpeciesOut<T> o = out_at(a.out, cf.e_species);
  const int64_t r = wave_reserve(o.append, emit);
  bool made = false;
  if (emit) made = photo_one(a, o, r, k, sg, f0, f1, area, cf, step, key0, key1);
  wave_add(&a.emit_count[((size_t)sg * kMaxGroups + a.group) * kMaxSpecies + cf.e_species], made ? 1ull : 0ull);
  wave_add(&a.counts[C_PHOTO], made ? 1ull : 0ull);
}

// One photoelectron of group a.group (scenario sg): its face, the point on
// it, the flux-weighted velocity, appended at slot index r.
template <typename T>
__device__ bool photo_one(const EmitArgs<T>& a, const SpeciesOut<T>& o, int64_t r, int64_t k, int sg, int64_t f0,
                          int64_t f1, double area, const SurfEmitDev& cf, uint32_t step, uint32_t key0,
                          uint32_t key1) {
  const U4 b = philox4x32((uint32_t)k, step, (uint32_t)a.group, kTagPhotoK, key0, key1);
  const U4 c = philox4x32((uint32_t)k, step, (uint32_t)a.group, kTagPhotoK2, key0, key1);
  // face: first cumulative area above u * area
  const double rr = u01(b.x) * area;
  int64_t lo = f0, hi = f1;  // search [lo, hi)
  while (lo < hi) {
    const int64_t mid = lo + (hi - lo) / 2;
    if (a.face_cum[mid] > rr) hi = mid;
    else lo = mid + 1;
  }
  const int64_t fi = lo < f1 ? lo : f1 - 1;
  const int t = a.face_tet[fi], kf = a.face_k[fi];
  const double* p[3];
  int mm = 0;
#pragma unroll
  for (int q = 0; q < 4; ++q)
    if (q != kf) p[mm++] = &a.xyz[3 * (int64_t)a.tet4[4 * (int64_t)t + q]];
  const double sq = sqrt(u01(b.y)), u2 = u01(b.z);
  double pt[3], xe[3], n[3], v[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) pt[q] = (1.0 - sq) * p[0][q] + sq * (1.0 - u2) * p[1][q] + sq * u2 * p[2][q];
  nudge(a.xyz, a.tet4, t, pt, xe);
  inward_normal(a.tets[t], kf, n);
  emit_velocity(n, cf.t_ph, o.m, u01(b.w), u01(c.x), u01(c.y), u01(c.z), v);
  const T xt[3] = {(T)xe[0], (T)xe[1], (T)xe[2]};
  return append_at(o, r, xt, v, t, gen_id(c.w), a.tets, a.rho, a.counts);
}

// Appends host-made particles (field emission) to a species and deposits
// them into this step's charge.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_inject(SpeciesOut<T> o, int64_t n, const double* x, const double* v,
                                                   const int32_t* cell, const int32_t* id, const TetMap<T>* tets,
                                                   unsigned long long* rho, unsigned long long* counts) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const T xt[3] = {(T)x[3 * i], (T)x[3 * i + 1], (T)x[3 * i + 2]};
  const double vv[3] = {v[3 * i], v[3 * i + 1], v[3 * i + 2]};
  append(o, xt, vv, cell[i], id[i], tets, rho, counts);
}

__global__ void k_fill_u64(unsigned long long* p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

__global__ void k_commit_counts(int64_t* n_dev, unsigned long long* append, CountArgs a) {
  const int k = threadIdx.x;
  if (k >= a.n_species) return;
  const int64_t n = n_dev[k] + (int64_t)append[k];
  n_dev[k] = n < a.cap[k] ? n : a.cap[k];
  append[k] = 0ull;
}

__global__ void k_step_advance(uint32_t* step) { *step += 1u; }

inline uint32_t blocks_for(int64_t n) { return (uint32_t)((n + kBlock - 1) / kBlock); }

}  // namespace

void launch_commit_counts(int64_t* n_dev, unsigned long long* append, const CountArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_commit_counts, dim3(1), dim3(kMaxSpecies), 0, s, n_dev, append, a);
}
void launch_step_advance(uint32_t* step_dev, hipStream_t s) {
  hipLaunchKernelGGL(k_step_advance, dim3(1), dim3(1), 0, s, step_dev);
}

template <typename T>
void launch_mcc(const MccArgs<T>& a, hipStream_t s) {
  if (a.n <= 0 || a.np <= 0) return;
  // candidates are few (P_null n plus the particles above the cap): a fixed
  // grid that loops over however many the push found
  const int blocks = (int)std::min<int64_t>(256, std::max<int64_t>(1, blocks_for(a.n)));
  hipLaunchKernelGGL(k_mcc<T>, dim3(blocks), dim3(kBlock), 0, s, a);
}
template <typename T>
void launch_see(const EmitArgs<T>& a, int64_t max_events, hipStream_t s) {
  if (max_events <= 0) return;
  // impacts are few: a small grid that loops over however many there are
  hipLaunchKernelGGL(k_see<T>, dim3((uint32_t)std::min<int64_t>(64, blocks_for(max_events))), dim3(kBlock), 0, s, a);
}
template <typename T>
void launch_photo(const EmitArgs<T>& a, int64_t max_count, int n_seg, hipStream_t s) {
  if (max_count <= 0 || (!a.seg_face_off && a.f1 <= a.f0) || n_seg < 1) return;
  hipLaunchKernelGGL(k_photo<T>, dim3(blocks_for(max_count), a.seg_face_off ? n_seg : 1), dim3(kBlock), 0, s, a);
}
template <typename T>
void launch_inject(const SpeciesOut<T>& o, int64_t n, const double* x, const double* v, const int32_t* cell,
                   const int32_t* id, const TetMap<T>* tets, unsigned long long* rho, unsigned long long* counts,
                   hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_inject<T>, dim3(blocks_for(n)), dim3(kBlock), 0, s, o, n, x, v, cell, id, tets, rho, counts);
}
void launch_fill_u64(unsigned long long* p, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill_u64, dim3(blocks_for(n)), dim3(kBlock), 0, s, p, n);
}

void launch_mcc_majorant(const MccMajorantArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_mcc_majorant, dim3(a.seg_n ? a.n_seg : 1), dim3(kMajBlock), 0, s, a);
}
template <typename T>
void launch_v2max(const T* vx, const T* vy, const T* vz, const int32_t* cell, int64_t n, unsigned long long* v2max,
                  const int32_t* tet_seg, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_v2max<T>, dim3(blocks_for(n)), dim3(kBlock), 0, s, vx, vy, vz, cell, n, v2max, tet_seg);
}

#define SCDAT_EMIT_INST(T)                                                        \
  template void launch_v2max<T>(const T*, const T*, const T*, const int32_t*, int64_t, unsigned long long*, \
                                const int32_t*, hipStream_t);                     \
  template void launch_inject<T>(const SpeciesOut<T>&, int64_t, const double*, const double*, const int32_t*, \
                                 const int32_t*, const TetMap<T>*, unsigned long long*, unsigned long long*, \
                                 hipStream_t);                                    \
  template void launch_mcc<T>(const MccArgs<T>&, hipStream_t);                    \
  template void launch_see<T>(const EmitArgs<T>&, int64_t, hipStream_t);          \
  template void launch_photo<T>(const EmitArgs<T>&, int64_t, int, hipStream_t);
SCDAT_EMIT_INST(float)
SCDAT_EMIT_INST(double)
#undef SCDAT_EMIT_INST

}  // namespace scdat
