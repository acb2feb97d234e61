// Device primitives for the DMC propagation kernels (gfx950, wave64).
//
//  * Philox4x32-10 counter-based generator (Salmon, Moraes, Dror, Shaw, SC'11),
//    keyed by the run seed and addressed by (walker slot, step, purpose, block).
//  * Box-Muller normals, FP32 and FP64.
//  * wave64 inclusive scans / reductions built on cross-lane shuffles, and block
//    versions through LDS.  Reductions use a fixed shuffle tree, so a sum is
//    bit-reproducible from run to run.
//  * agent-scope acquire/release helpers for the single-pass look-back scan: the eight
//    XCD L2 caches are not coherent with each other, so every flag that one workgroup
//    publishes for another goes through agent-scope atomics.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define PVD_WAVE 64
#define PVD_DEV __device__ __forceinline__

namespace pvd {

// ------------------------------------------------------------------------------ RNG
struct Philox4 {
  uint32_t v[4];
};

PVD_DEV uint32_t mulhi32(uint32_t a, uint32_t b) { return __umulhi(a, b); }

// Philox4x32 with 10 rounds.  ctr = (c0..c3), key = (k0, k1).
PVD_DEV Philox4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                              uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    // one 32x32 -> 64-bit product gives both halves (v_mad_u64_u32)
    const uint64_t p0 = (uint64_t)M0 * (uint64_t)c0;
    const uint64_t p1 = (uint64_t)M1 * (uint64_t)c2;
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    const uint32_t n0 = hi1 ^ c1 ^ k0;
    const uint32_t n2 = hi0 ^ c3 ^ k1;
    c0 = n0;
    c1 = lo1;
    c2 = n2;
    c3 = lo0;
    k0 += W0;
    k1 += W1;
  }
  Philox4 o;
  o.v[0] = c0;
  o.v[1] = c1;
  o.v[2] = c2;
  o.v[3] = c3;
  return o;
}

// purposes of a draw (third counter word, low byte)
enum : uint32_t { RNG_DISP = 1u, RNG_BRANCH = 2u, RNG_METRO = 3u, RNG_STEP = 4u };

struct Rng {
  uint32_t k0, k1;
  PVD_DEV Philox4 block(uint32_t slot, uint64_t step, uint32_t purpose, uint32_t blk) const {
    return philox4x32_10(slot, (uint32_t)step, purpose | (blk << 8), (uint32_t)(step >> 32), k0, k1);
  }
};

// uniform in [0,1) with 24 bits, and in (0,1] for logarithms
PVD_DEV float u01f(uint32_t x) { return (float)(x >> 8) * 5.9604644775390625e-08f; }
PVD_DEV float u01f_open0(uint32_t x) { return (float)((x >> 8) + 1u) * 5.9604644775390625e-08f; }
// uniform in [0,1) with 53 bits from two words
PVD_DEV double u01d(uint32_t a, uint32_t b) {
  const uint64_t m = ((uint64_t)(a >> 5) << 26) | (uint64_t)(b >> 6);
  return (double)m * 1.1102230246251565e-16;
}
PVD_DEV double u01d_open0(uint32_t a, uint32_t b) {
  const uint64_t m = ((uint64_t)(a >> 5) << 26) | (uint64_t)(b >> 6);
  return (double)(m + 1u) * 1.1102230246251565e-16;
}

// Box-Muller: two standard normals from two uniforms.
PVD_DEV void box_muller_f(uint32_t a, uint32_t b, float& z0, float& z1) {
  const float u1 = u01f_open0(a);
  const float u2 = u01f(b);
  const float r = __fsqrt_rn(-2.0f * __logf(u1));
  float s, c;
  __sincosf(6.283185307179586f * u2, &s, &c);
  z0 = r * c;
  z1 = r * s;
}
// Box-Muller from a 24-bit radius uniform r24 (u1 = (r24 + 1) 2^-24 in (0, 1]) and a
// 16-bit angle uniform a16 (u2 = a16 2^-16 in [0, 1)): z0 = rho cos(2 pi u2),
// z1 = rho sin(2 pi u2), rho = sqrt(-2 ln u1) = sqrt(-2 ln2 log2 u1).  The hardware sine /
// cosine take the angle in revolutions.
PVD_DEV void box_muller_24_16(uint32_t r24, uint32_t a16, float& z0, float& z1) {
  const float u1 = (float)(r24 + 1u) * 5.9604644775390625e-08f;
  const float u2 = (float)a16 * 1.52587890625e-05f;
  const float rho = __builtin_amdgcn_sqrtf(-1.3862943611198906f * __builtin_amdgcn_logf(u1));
  z0 = rho * __builtin_amdgcn_cosf(u2);
  z1 = rho * __builtin_amdgcn_sinf(u2);
}
PVD_DEV void box_muller_d(uint32_t a0, uint32_t a1, uint32_t b0, uint32_t b1, double& z0, double& z1) {
  const double u1 = u01d_open0(a0, a1);
  const double u2 = u01d(b0, b1);
  const double r = sqrt(-2.0 * log(u1));
  double s, c;
  sincospi(2.0 * u2, &s, &c);
  z0 = r * c;
  z1 = r * s;
}

// NP standard normals for walker `slot` at `step`.
template <typename Real, int NP>
struct Normals;

template <int NP>
struct Normals<float, NP> {
  PVD_DEV static void draw(const Rng& g, uint32_t slot, uint64_t step, float* z) {
#pragma unroll
    for (int b = 0; b < (NP + 3) / 4; ++b) {
      const Philox4 r = g.block(slot, step, RNG_DISP, (uint32_t)b);
      float a0, a1, a2, a3;
      box_muller_f(r.v[0], r.v[1], a0, a1);
      box_muller_f(r.v[2], r.v[3], a2, a3);
      if (4 * b + 0 < NP) z[4 * b + 0] = a0;
      if (4 * b + 1 < NP) z[4 * b + 1] = a1;
      if (4 * b + 2 < NP) z[4 * b + 2] = a2;
      if (4 * b + 3 < NP) z[4 * b + 3] = a3;
    }
  }
};

template <int NP>
struct Normals<double, NP> {
  PVD_DEV static void draw(const Rng& g, uint32_t slot, uint64_t step, double* z) {
#pragma unroll
    for (int b = 0; b < (NP + 1) / 2; ++b) {
      const Philox4 r = g.block(slot, step, RNG_DISP, (uint32_t)b);
      double a0, a1;
      box_muller_d(r.v[0], r.v[1], r.v[2], r.v[3], a0, a1);
      if (2 * b + 0 < NP) z[2 * b + 0] = a0;
      if (2 * b + 1 < NP) z[2 * b + 1] = a1;
    }
  }
};

// NW consecutive 32-bit words of one walker's stream at one step (blocks 0, 1, ...)
template <int NW>
PVD_DEV void philox_words(const Rng& g, uint32_t slot, uint64_t step, uint32_t* w) {
#pragma unroll
  for (int b = 0; b < (NW + 3) / 4; ++b) {
    const Philox4 r = g.block(slot, step, RNG_STEP, (uint32_t)b);
#pragma unroll
    for (int j = 0; j < 4; ++j)
      if (4 * b + j < NW) w[4 * b + j] = r.v[j];
  }
}

PVD_DEV double uniform_d(const Rng& g, uint32_t slot, uint64_t step, uint32_t purpose) {
  const Philox4 r = g.block(slot, step, purpose, 0u);
  return u01d(r.v[0], r.v[1]);
}

// ------------------------------------------------------------------------------ lanes
PVD_DEV int lane_id() { return (int)__lane_id(); }

template <typename T>
PVD_DEV T wave_incl_scan(T v) {
  const int l = lane_id();
#pragma unroll
  for (int o = 1; o < PVD_WAVE; o <<= 1) {
    const T t = __shfl_up(v, o, PVD_WAVE);
    if (l >= o) v += t;
  }
  return v;
}

template <typename T>
PVD_DEV T wave_sum(T v) {
#pragma unroll
  for (int o = PVD_WAVE / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, PVD_WAVE);
  return v;
}

template <typename T>
PVD_DEV T wave_max(T v) {
#pragma unroll
  for (int o = PVD_WAVE / 2; o > 0; o >>= 1) {
    const T t = __shfl_xor(v, o, PVD_WAVE);
    v = t > v ? t : v;
  }
  return v;
}

// Block exclusive scan of ints; returns this thread's exclusive prefix and writes the
// block total.  `sh` needs NW ints.
template <int NT>
PVD_DEV int block_excl_scan(int v, int* sh, int& total) {
  constexpr int NW = NT / PVD_WAVE;
  const int w = threadIdx.x / PVD_WAVE, l = lane_id();
  const int inc = wave_incl_scan(v);
  if (l == PVD_WAVE - 1) sh[w] = inc;
  __syncthreads();
  int base = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < NW; ++i) {
    const int s = sh[i];
    base += (i < w) ? s : 0;
    tot += s;
  }
  __syncthreads();
  total = tot;
  return base + inc - v;
}

// Block sum with a fixed reduction tree (deterministic).  `sh` needs NW entries.
template <int NT, typename T>
PVD_DEV T block_sum(T v, T* sh) {
  constexpr int NW = NT / PVD_WAVE;
  const int w = threadIdx.x / PVD_WAVE, l = lane_id();
  v = wave_sum(v);
  if (l == 0) sh[w] = v;
  __syncthreads();
  T r = sh[0];
#pragma unroll
  for (int i = 1; i < NW; ++i) r += sh[i];
  __syncthreads();
  return r;
}

template <int NT, typename T>
PVD_DEV T block_max(T v, T* sh) {
  constexpr int NW = NT / PVD_WAVE;
  const int w = threadIdx.x / PVD_WAVE, l = lane_id();
  v = wave_max(v);
  if (l == 0) sh[w] = v;
  __syncthreads();
  T r = sh[0];
#pragma unroll
  for (int i = 1; i < NW; ++i) r = sh[i] > r ? sh[i] : r;
  __syncthreads();
  return r;
}

// ------------------------------------------------------------------------------ atomics
PVD_DEV uint64_t ld_acquire(const uint64_t* p) {
  return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
PVD_DEV void st_release(uint64_t* p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}
PVD_DEV int ld_acquire_i(const int* p) {
  return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
PVD_DEV void st_release_i(int* p, int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}
PVD_DEV unsigned fetch_add_u(unsigned* p, unsigned v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
}
PVD_DEV void fence_acquire() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent"); }
// Relaxed agent-scope atomics: coherent across the XCDs (they bypass the per-XCD L2) and,
// unlike acquire / release, they need no L2 write-back or invalidate.  Used for values that
// are self-contained in one 64-bit word (look-back flags carry their payload).
PVD_DEV uint64_t ld_relaxed(const uint64_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
PVD_DEV void st_relaxed(uint64_t* p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
PVD_DEV unsigned fetch_add_relaxed(unsigned* p, unsigned v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
PVD_DEV void fence_release() { __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent"); }
PVD_DEV unsigned ld_relaxed_u(const unsigned* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
PVD_DEV void st_relaxed_u(unsigned* p, unsigned v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Wait until this wave's earlier memory operations have completed (relaxed agent-scope
// stores are then performed at the device's coherence point), with no cache write-back:
// a later store by this wave cannot be observed before them.  The signal fences keep the
// compiler from moving memory operations across the wait.
PVD_DEV void store_complete() {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0);
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

}  // namespace pvd
