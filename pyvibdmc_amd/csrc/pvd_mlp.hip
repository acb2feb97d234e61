// Surrogate potential: permutation-sorted inverse-distance descriptor + MLP (D -> H -> H -> H -> 1,
// SiLU) on BF16 MFMA with FP32 accumulation, for inference inside the DMC step and for
// training (forward with stored pre-activations, backward, weight gradients, Adam).
//
// MFMA formulation (v_mfma_f32_16x16x32_bf16, wave64).  Weights are the A operand
// (16 output units x 32 inputs) and activations the B operand (32 inputs x 16 walkers);
// lane l holds A[l%16][8(l/16)+j], B[8(l/16)+j][l%16] (j < 8) and D[4(l/16)+i][l%16]
// (i < 4).  A layer's accumulators therefore already hold, for walker l%16, the units
// 16m + 4(l/16) + i.  Reading the next layer's 32-wide input chunk c as
//   k = 8g + j  ->  unit 16(2c + j/4) + 4g + j%4
// makes those registers the next B operand directly, so activations never leave
// registers; the weights are stored pre-permuted to match (pvd_mlp_pack).
#include <hip/hip_runtime.h>
#include <mutex>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>

#include "../../include/pyvibdmc_amd.h"
#include "pvd_models.hpp"

namespace pvd {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(4))) float f32x4;

constexpr int MLP_MAXD = 32;
constexpr int MLP_MAXA = 6;
constexpr int MLP_MAXP = 15;
constexpr int MLP_NQ = 1 + MLP_MAXP + MLP_MAXP * (MLP_MAXP + 1) / 2;  // 136

struct MlpNet {
  int D, H;
  const bf16x8* Wf[3];
  const bf16x8* Wt[2];  // transposed packs of layers 1 and 2
  const float* b[3];
  const float* w3;
  const float* b3;
  int na, use_range;
  float zz[MLP_MAXP];             // z_i z_j of atom pair q in order (0,1), (0,2), ..., (1,2), ...
  float cd2[MLP_MAXA];            // squared Coulomb diagonal (0.5 z_a^2.4)^2
  int cmask[MLP_MAXA];            // bit b: atom b is sorted together with atom a (0: a stays)
  int slot[MLP_MAXA][MLP_MAXA];   // position of the r-th largest-norm member of a's sublist
  int ng, gs;                     // ng > 1: ng position groups of gs positions are sorted
  int gpos[MLP_MAXA][MLP_MAXA];   // position m of group g
  float xmean[MLP_MAXD], xinv[MLP_MAXD], lo[MLP_MAXD], hi[MLP_MAXD];
  float ymean, ystd;
  float quad[MLP_NQ];  // quadratic baseline on the normalised features
};

// quadratic baseline q(x) of the NPR normalised features (FP32): constant, linear terms,
// then x_i x_j for i <= j in row order
template <int NPR>
PVD_DEV float quad_baseline(const MlpNet& M, const float* xn) {
  float q = M.quad[0];
#pragma unroll
  for (int i = 0; i < NPR; ++i) q += M.quad[1 + i] * xn[i];
  int o = 1 + NPR;
#pragma unroll
  for (int i = 0; i < NPR; ++i)
#pragma unroll
    for (int j = i; j < NPR; ++j) q += M.quad[o++] * (xn[i] * xn[j]);
  return q;
}

// SiLU and its derivative with the hardware exp2 and reciprocal (about 1 ulp each); the
// same functions serve training, backward and inference, so the network is consistent
PVD_DEV float sigmoid_f(float z) {
  return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(-1.4426950408889634f * z));
}
PVD_DEV float silu_f(float z) { return z * sigmoid_f(z); }
PVD_DEV float dsilu_f(float z) {
  const float s = sigmoid_f(z);
  return s * (1.0f + z * (1.0f - s));
}

// unit index of element j of input chunk c for lane group g (the permuted K order)
PVD_DEV int phys_unit(int c, int g, int j) { return 16 * (2 * c + (j >> 2)) + 4 * g + (j & 3); }

// descriptor of one walker (the DistIt permutation sorting of
// simulation_utilities/tensorflow_descriptors/distance_descriptors.py:115-152): atoms are
// permuted, never pair distances, so the features stay the upper triangle of one inverse
// distance matrix.
//  1. sort key: the row norm of the Coulomb matrix (z_i z_j / r_ij, diagonal 0.5 z_i^2.4);
//  2. sorted atoms: within each sublist, position sublist[r] receives the member with the
//     r-th largest norm (ties: lower atom index first);
//  3. sorted groups: equal-size position groups are reordered by the sum of their atoms'
//     norms, largest first (ties: lower group index first);
//  4. features: 1/r of the permuted atoms for position pairs (0,1), (0,2), ..., (1,2), ...
//     (the Coulomb values divided by z_i z_j, a per-feature constant the normalisation
//     removes).
// The permutation is packed as 4-bit fields (position p holds atom (perm >> 4p) & 15), so
// no lane indexes a register array at run time; every table lookup has a compile-time index.
template <typename R, int NP>
PVD_DEV void mlp_descriptor(const MlpNet& M, const R* x, float* f) {
  constexpr int NA = NP / 3, NPR = NA * (NA - 1) / 2;
  static_assert(NA <= MLP_MAXA && NPR <= MLP_MAXP, "too many atoms");
  float dc[NPR];
  float rn2[NA];
#pragma unroll
  for (int a = 0; a < NA; ++a) rn2[a] = M.cd2[a];
  {
    int q = 0;
#pragma unroll
    for (int i = 0; i < NA; ++i)
#pragma unroll
      for (int j = i + 1; j < NA; ++j) {
        const float d0 = (float)(x[3 * i] - x[3 * j]);
        const float d1 = (float)(x[3 * i + 1] - x[3 * j + 1]);
        const float d2 = (float)(x[3 * i + 2] - x[3 * j + 2]);
        dc[q] = 1.0f / __fsqrt_rn(d0 * d0 + d1 * d1 + d2 * d2);
        const float c = M.zz[q] * dc[q];
        rn2[i] += c * c;
        rn2[j] += c * c;
        ++q;
      }
  }
  unsigned perm = 0;
#pragma unroll
  for (int a = 0; a < NA; ++a) {
    int r = 0;
#pragma unroll
    for (int b = 0; b < NA; ++b)
      if (b != a && ((M.cmask[a] >> b) & 1)) r += (rn2[b] > rn2[a] || (rn2[b] == rn2[a] && b < a)) ? 1 : 0;
    int p = a;
    if (M.cmask[a]) {
#pragma unroll
      for (int t = 0; t < NA; ++t)
        if (r == t) p = M.slot[a][t];
    }
    perm |= (unsigned)a << (4 * p);
  }
  if (M.ng > 1) {
    float tot[NA];
#pragma unroll
    for (int g = 0; g < NA; ++g) {
      tot[g] = 0.0f;
      if (g < M.ng) {
#pragma unroll
        for (int m = 0; m < NA; ++m) {
          if (m < M.gs) {
            const unsigned at = (perm >> (4 * M.gpos[g][m])) & 15u;
            float nrm = 0.0f;
#pragma unroll
            for (int a = 0; a < NA; ++a)
              if (at == (unsigned)a) nrm = __fsqrt_rn(rn2[a]);
            tot[g] += nrm;
          }
        }
      }
    }
    unsigned fp = perm;
#pragma unroll
    for (int g = 0; g < NA; ++g) {
      if (g < M.ng) {
        int k = 0;
#pragma unroll
        for (int h = 0; h < NA; ++h)
          if (h < M.ng && h != g) k += (tot[h] > tot[g] || (tot[h] == tot[g] && h < g)) ? 1 : 0;
#pragma unroll
        for (int m = 0; m < NA; ++m) {
          if (m < M.gs) {
            int dst = 0;
#pragma unroll
            for (int kk = 0; kk < NA; ++kk)
              if (kk == k) dst = M.gpos[kk][m];
            const unsigned src = (perm >> (4 * M.gpos[g][m])) & 15u;
            fp = (fp & ~(15u << (4 * dst))) | (src << (4 * dst));
          }
        }
      }
    }
    perm = fp;
  }
  {
    int q = 0;
#pragma unroll
    for (int p1 = 0; p1 < NA; ++p1)
#pragma unroll
      for (int p2 = p1 + 1; p2 < NA; ++p2) {
        const int a1 = (int)((perm >> (4 * p1)) & 15u), a2 = (int)((perm >> (4 * p2)) & 15u);
        const int lo = a1 < a2 ? a1 : a2, hi = a1 < a2 ? a2 : a1;
        const int cq = lo * NA - lo * (lo + 1) / 2 + (hi - lo - 1);
        float v = 0.0f;
#pragma unroll
        for (int s = 0; s < NPR; ++s)
          if (cq == s) v = dc[s];
        f[q++] = v;
      }
  }
#pragma unroll
  for (int k = NPR; k < MLP_MAXP; ++k) f[k] = 0.0f;
}

// normalised features 8g..8g+7 of one walker as a B operand chunk (D = NPR features)
template <int NPR>
PVD_DEV bf16x8 input_chunk(const MlpNet& M, const float* f, int g, bool valid) {
  float fn[NPR];
#pragma unroll
  for (int k = 0; k < NPR; ++k) fn[k] = (f[k] - M.xmean[k]) * M.xinv[k];
  bf16x8 v;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    float s = 0.0f;
#pragma unroll
    for (int k = j; k < NPR; k += 8)
      if (k == 8 * g + j) s = fn[k];
    v[j] = (__bf16)(valid ? s : 0.0f);
  }
  return v;
}

// D(m) += sum_c A(m, c) B(c) for NT walker tiles; A read from LDS
template <int NT, int MT, int KC>
PVD_DEV void mfma_layer(const bf16x8* __restrict__ Wl, const bf16x8 (&bin)[NT][KC], f32x4 (&acc)[NT][MT], int lane) {
#pragma unroll
  for (int m = 0; m < MT; ++m)
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[t][m] = f32x4{0.f, 0.f, 0.f, 0.f};
  // input chunk outer, output tile inner: consecutive MFMAs write different accumulators
  // (each accumulator still sums its chunks in order c = 0, 1, ...)
#pragma unroll
  for (int c = 0; c < KC; ++c) {
#pragma unroll
    for (int m = 0; m < MT; ++m) {
      const bf16x8 a = Wl[(m * KC + c) * 64 + lane];
#pragma unroll
      for (int t = 0; t < NT; ++t) acc[t][m] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bin[t][c], acc[t][m], 0, 0, 0);
    }
  }
}

template <int N>
PVD_DEV void stage(bf16x8* lds, const bf16x8* __restrict__ src, int words) {
  __syncthreads();
  for (int q = threadIdx.x; q < words; q += N) lds[q] = src[q];
  __syncthreads();
}

// ------------------------------------------------------------------------------------
// Network forward.  MODE 0: inference on walkers (SoA planes of R) -> V (double), with the
// analytic fallback outside the training range.  MODE 1: training forward on normalised
// descriptors X^T (D x B) -> pre-activations Z^T (3 x H x B) and y (B).
template <typename R, int NP, int H, int NT, int MODE>
__global__ __launch_bounds__(256) void k_mlp_forward(MlpNet M, const R* planes, long long stride,
                                                     const int* n_dev, long long n_fix, double* V, PesParams fb,
                                                     const float* XT, float* ZT, float* Y) {
  extern __shared__ bf16x8 lds[];
  constexpr int MT = H / 16, KC = H / 32;
  const int n = n_dev ? *n_dev : (int)n_fix;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int g = lane >> 4, col = lane & 15;
  const long long base = ((long long)blockIdx.x * 4 + wave) * (16 * NT);
  if ((long long)blockIdx.x * 4 * 16 * NT >= n) return;  // whole block past the end
  bf16x8 bin0[NT][1];
  bf16x8 bin[NT][KC];
  bool fbk[NT];
  double vfb[NT];
  float qb[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const long long w = base + t * 16 + col;
    const bool ok = w < n;
    fbk[t] = false;
    vfb[t] = 0.0;
    qb[t] = 0.0f;
    if constexpr (MODE == 0) {
      float f[MLP_MAXP];
      R x[NP];
#pragma unroll
      for (int p = 0; p < NP; ++p) x[p] = ok ? planes[p * stride + w] : R(0);
      mlp_descriptor<R, NP>(M, x, f);
      if (ok && M.use_range) {
        bool out = false;
#pragma unroll
        for (int k = 0; k < MLP_MAXP; ++k)
          if (k < M.D) out |= !(f[k] >= M.lo[k] && f[k] <= M.hi[k]);
        if (out) {
          fbk[t] = true;
          vfb[t] = (double)pes_eval<R, NP>(fb, x);
        }
      }
      constexpr int NPR = NP / 3 * (NP / 3 - 1) / 2;
      bin0[t][0] = input_chunk<NPR>(M, f, g, ok);
      if (g == 0) {
        float xn[NPR];
#pragma unroll
        for (int k = 0; k < NPR; ++k) xn[k] = (f[k] - M.xmean[k]) * M.xinv[k];
        qb[t] = quad_baseline<NPR>(M, xn);
      }
    } else {
      bf16x8 v;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int k = 8 * g + j;
        v[j] = (__bf16)((ok && k < M.D) ? XT[(long long)k * n + w] : 0.0f);
      }
      bin0[t][0] = v;
    }
  }
  f32x4 acc[NT][MT];
  float part[NT];
  // layer 0 (input chunk c = 0 in natural order)
  stage<256>(lds, M.Wf[0], MT * 64);
  mfma_layer<NT, MT, 1>(lds, bin0, acc, lane);
#pragma unroll
  for (int layer = 0; layer < 3; ++layer) {
    if (layer > 0) {
      stage<256>(lds, M.Wf[layer], MT * KC * 64);
      mfma_layer<NT, MT, KC>(lds, bin, acc, lane);
    }
    const float* bl = M.b[layer];
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const long long w = base + t * 16 + col;
      const bool ok = w < n;
      if (layer == 2) part[t] = 0.0f;
#pragma unroll
      for (int m = 0; m < MT; ++m) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int u = 16 * m + 4 * g + i;
          const float z = acc[t][m][i] + bl[u];
          if constexpr (MODE == 1) {
            if (ok) ZT[((long long)layer * H + u) * n + w] = z;
          }
          const float h = silu_f(z);
          if (layer < 2) bin[t][m >> 1][(m & 1) * 4 + i] = (__bf16)h;
          else part[t] += M.w3[u] * h;
        }
      }
    }
  }
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    float s = part[t];
    s += __shfl_xor(s, 16, 64);
    s += __shfl_xor(s, 32, 64);
    const long long w = base + t * 16 + col;
    if (g == 0 && w < n) {
      const float y = s + M.b3[0];
      if constexpr (MODE == 0) {
        V[w] = fbk[t] ? vfb[t] : (double)(y * M.ystd + M.ymean + qb[t]);
      } else {
        Y[w] = y;
      }
    }
  }
}

// Backward through the hidden layers for one batch: dZ_l^T = dL/dz_l (3 x H x B) from the
// stored pre-activations and dy = dL/dy.  dh_{l-1} = W_l^T dz_l uses the transposed packs.
template <int H, int NT>
__global__ __launch_bounds__(256) void k_mlp_backward(MlpNet M, long long n, const float* ZT, const float* dY,
                                                      float* DZT) {
  extern __shared__ bf16x8 lds[];
  constexpr int MT = H / 16, KC = H / 32;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int g = lane >> 4, col = lane & 15;
  const long long base = ((long long)blockIdx.x * 4 + wave) * (16 * NT);
  if ((long long)blockIdx.x * 4 * 16 * NT >= n) return;
  bf16x8 bin[NT][KC];
  f32x4 acc[NT][MT];
  // layer 2: dz2 = w3 * dy * silu'(z2)
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const long long w = base + t * 16 + col;
    const bool ok = w < n;
    const float dy = ok ? dY[w] : 0.0f;
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int u = 16 * m + 4 * g + i;
        const float z = ok ? ZT[((long long)2 * H + u) * n + w] : 0.0f;
        const float dz = M.w3[u] * dy * dsilu_f(z);
        if (ok) DZT[((long long)2 * H + u) * n + w] = dz;
        bin[t][m >> 1][(m & 1) * 4 + i] = (__bf16)dz;
      }
  }
#pragma unroll
  for (int layer = 2; layer >= 1; --layer) {
    stage<256>(lds, M.Wt[layer - 1], MT * KC * 64);
    mfma_layer<NT, MT, KC>(lds, bin, acc, lane);
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const long long w = base + t * 16 + col;
      const bool ok = w < n;
#pragma unroll
      for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int u = 16 * m + 4 * g + i;
          const float z = ok ? ZT[((long long)(layer - 1) * H + u) * n + w] : 0.0f;
          const float dz = acc[t][m][i] * dsilu_f(z);
          if (ok) DZT[((long long)(layer - 1) * H + u) * n + w] = dz;
          bin[t][m >> 1][(m & 1) * 4 + i] = (__bf16)dz;
        }
    }
  }
}

// Weight gradient of layer l over a batch slice: P[split][o][i] = sum_b dz[o][b] * hin[i][b]
// with hin = X (l = 0) or silu(z_{l-1}).  One wave per 32 x 32 output block.
__global__ __launch_bounds__(64) void k_mlp_wgrad(int layer, int H, int Din, long long n, int kslice,
                                                  const float* XT, const float* ZT, const float* DZT, float* P) {
  const int lane = threadIdx.x;
  const int g = lane >> 4, col = lane & 15;
  const int ob = blockIdx.x * 32, ib = blockIdx.y * 32, split = blockIdx.z;
  const long long k0 = (long long)split * kslice;
  const long long k1 = k0 + kslice < n ? k0 + kslice : n;
  const float* dz = DZT + (long long)layer * H * n;
  const float* hz = layer == 0 ? XT : ZT + (long long)(layer - 1) * H * n;
  f32x4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (long long kb = k0; kb < k1; kb += 32) {
    bf16x8 A[2], B[2];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      const int o = ob + 16 * a + col;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const long long b = kb + 8 * g + j;
        A[a][j] = (__bf16)((b < k1 && o < H) ? dz[(long long)o * n + b] : 0.0f);
      }
    }
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const int i = ib + 16 * c + col;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const long long b = kb + 8 * g + j;
        float h = 0.0f;
        if (b < k1 && i < Din) {
          const float v = hz[(long long)i * n + b];
          h = layer == 0 ? v : silu_f(v);
        }
        B[c][j] = (__bf16)h;
      }
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int c = 0; c < 2; ++c) acc[a][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[a], B[c], acc[a][c], 0, 0, 0);
  }
  const int Hin = layer == 0 ? Din : H;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int i4 = 0; i4 < 4; ++i4) {
        const int o = ob + 16 * a + 4 * g + i4;
        const int i = ib + 16 * c + col;
        if (o < H && i < Hin) P[((long long)split * H + o) * Hin + i] = acc[a][c][i4];
      }
}

// fixed-order sum of the split partials into the gradient vector
__global__ void k_mlp_wgrad_reduce(const float* P, int splits, long long size, float* G) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= size) return;
  float s = 0.0f;
  for (int k = 0; k < splits; ++k) s += P[k * size + q];
  G[q] = s;
}

// bias gradients and the output layer: one block per row, fixed-order reduction
__global__ __launch_bounds__(256) void k_mlp_rowsum(int H, long long n, const float* ZT, const float* DZT,
                                                    const float* dY, float* gb0, float* gb1, float* gb2, float* gw3,
                                                    float* gb3) {
  __shared__ float sh[4];
  const int row = blockIdx.x;  // 0..3H: b0, b1, b2 rows, then w3 rows, then b3
  float s = 0.0f;
  for (long long b = threadIdx.x; b < n; b += 256) {
    if (row < 3 * H) {
      s += DZT[(long long)row * n + b];
    } else if (row < 4 * H) {
      const int u = row - 3 * H;
      s += dY[b] * silu_f(ZT[((long long)2 * H + u) * n + b]);
    } else {
      s += dY[b];
    }
  }
  s = block_sum<256>(s, sh);
  if (threadIdx.x == 0) {
    if (row < H) gb0[row] = s;
    else if (row < 2 * H) gb1[row - H] = s;
    else if (row < 3 * H) gb2[row - 2 * H] = s;
    else if (row < 4 * H) gw3[row - 3 * H] = s;
    else gb3[0] = s;
  }
}

// residual of the MSE loss: dy = 2 (y - t) / n; per-block squared error into loss parts
__global__ __launch_bounds__(256) void k_mlp_loss(long long n, const float* Y, const float* T, float* dY,
                                                  double* parts) {
  __shared__ double sh[4];
  const long long b = (long long)blockIdx.x * 256 + threadIdx.x;
  double e = 0.0;
  if (b < n) {
    const float r = Y[b] - T[b];
    dY[b] = 2.0f * r / (float)n;
    e = (double)r * (double)r;
  }
  e = block_sum<256>(e, sh);
  if (threadIdx.x == 0) parts[blockIdx.x] = e;
}

// Adam with bias correction on the flat FP32 parameter vector
__global__ void k_adam(float* p, float* m, float* v, const float* g, long long size, float lr, float b1, float b2,
                       float eps, float c1, float c2) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= size) return;
  const float gg = g[q];
  const float mm = b1 * m[q] + (1.0f - b1) * gg;
  const float vv = b2 * v[q] + (1.0f - b2) * gg * gg;
  m[q] = mm;
  v[q] = vv;
  p[q] = p[q] - lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}

// FP32 master weights -> packed BF16 operands.  Forward pack of layer l (A = W_l):
// element (m, c, lane, j) = W[16m + lane%16][k], k = 8(lane/16)+j (layer 0) or the
// permuted unit (layers 1, 2).  Transposed pack (A = W_l^T): W[permuted k][16m + lane%16].
__global__ void k_mlp_pack(const float* W, int rows, int cols, int in_natural, int transpose, int MT, int KC,
                           bf16x8* out) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long total = (long long)MT * KC * 64;
  if (q >= total) return;
  const int lane = (int)(q % 64);
  const int c = (int)((q / 64) % KC);
  const int m = (int)(q / (64 * KC));
  const int g = lane >> 4, col = lane & 15;
  bf16x8 v;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int k = in_natural ? 32 * c + 8 * g + j : phys_unit(c, g, j);
    const int r = 16 * m + col;
    float x = 0.0f;
    if (!transpose) {
      if (r < rows && k < cols) x = W[(long long)r * cols + k];
    } else {
      // A = W^T: row r is an input unit of W, k an output unit
      if (k < rows && r < cols) x = W[(long long)k * cols + r];
    }
    v[j] = (__bf16)x;
  }
  out[q] = v;
}

// descriptors of AoS walkers (n, M, 3) -> raw features F^T (D x n), float32
template <typename R, int NP>
__global__ void k_mlp_desc(MlpNet M, const R* x, long long n, float* FT) {
  const long long w = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= n) return;
  R xx[NP];
#pragma unroll
  for (int p = 0; p < NP; ++p) xx[p] = x[w * NP + p];
  float f[MLP_MAXP];
  mlp_descriptor<R, NP>(M, xx, f);
#pragma unroll
  for (int k = 0; k < MLP_MAXP; ++k)
    if (k < M.D) FT[(long long)k * n + w] = f[k];
}

// raw features -> normalised features in place
__global__ void k_mlp_normalise(MlpNet M, long long n, float* FT) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (long long)M.D * n) return;
  const int k = (int)(q / n);
  FT[q] = (FT[q] - M.xmean[k]) * M.xinv[k];
}

}  // namespace pvd

using namespace pvd;

namespace {
thread_local std::string g_merr;
int mfail(int code, const std::string& m) {
  g_merr = m;
  return code;
}
#define MHIP(x)                                                                                  \
  do {                                                                                           \
    hipError_t e_ = (x);                                                                         \
    if (e_ != hipSuccess) return mfail(8, std::string(#x) + ": " + hipGetErrorString(e_));       \
  } while (0)

long long off_W(int l, int D, int H) {
  // flat parameter order: W0 (H x D), b0, W1 (H x H), b1, W2, b2, w3, b3
  switch (l) {
    case 0: return 0;
    case 1: return (long long)H * D + H;
    default: return (long long)H * D + H + (long long)H * H + H;
  }
}
long long off_b(int l, int D, int H) { return off_W(l, D, H) + (l == 0 ? (long long)H * D : (long long)H * H); }
long long off_w3(int D, int H) { return (long long)H * D + 3ll * H + 2ll * H * H; }

// the atom-sorting tables of the descriptor from the DistIt-style specification; n_atoms = 0
// describes a network trained on given features only (no descriptor)
bool build_sorting(const pvd_mlp* m, MlpNet& N) {
  const int na = m->n_atoms;
  N.na = na;
  if (na == 0) return true;
  if (na < 2 || na > MLP_MAXA || m->D != na * (na - 1) / 2) return false;
  int q = 0;
  for (int i = 0; i < na; ++i) {
    if (m->z[i] < 1) return false;
    const double d = 0.5 * std::pow((double)m->z[i], 2.4);
    N.cd2[i] = (float)(d * d);
    for (int j = i + 1; j < na; ++j) N.zz[q++] = (float)m->z[i] * (float)m->z[j];
  }
  // sorted atoms: sublists partition the atoms (or there are none)
  if (m->n_sorted_atoms < 0 || m->n_sorted_atoms > na) return false;
  if (m->n_sorted_atoms > 0) {
    int seen = 0, o = 0;
    for (int s = 0; s < m->n_sorted_atoms; ++s) {
      const int len = m->sorted_atoms_len[s];
      if (len < 1 || o + len > na) return false;
      int mask = 0;
      for (int t = 0; t < len; ++t) {
        const int a = m->sorted_atoms[o + t];
        if (a < 0 || a >= na || ((seen >> a) & 1)) return false;
        seen |= 1 << a;
        mask |= 1 << a;
      }
      for (int t = 0; t < len; ++t) {
        const int a = m->sorted_atoms[o + t];
        N.cmask[a] = len > 1 ? (mask & ~(1 << a)) : 0;
        for (int r = 0; r < len; ++r) N.slot[a][r] = m->sorted_atoms[o + r];
      }
      o += len;
    }
    if (seen != (1 << na) - 1) return false;
  }
  // sorted groups: n_groups groups of group_size distinct positions
  N.ng = m->n_groups;
  N.gs = m->group_size;
  if (N.ng > 1) {
    if (N.gs < 1 || N.ng * N.gs > na) return false;
    int seen = 0;
    for (int g = 0; g < N.ng; ++g)
      for (int k = 0; k < N.gs; ++k) {
        const int p = m->sorted_groups[g * N.gs + k];
        if (p < 0 || p >= na || ((seen >> p) & 1)) return false;
        seen |= 1 << p;
        N.gpos[g][k] = p;
      }
  } else {
    N.ng = 0;
  }
  return true;
}

bool build_net(const pvd_mlp* m, MlpNet& N) {
  std::memset(&N, 0, sizeof(N));
  const int D = m->D, H = m->H;
  if (D < 1 || D > 32 || (H != 128 && H != 256)) return false;
  N.D = D;
  N.H = H;
  const float* p = (const float*)m->params;
  for (int l = 0; l < 3; ++l) N.b[l] = p + off_b(l, D, H);
  N.w3 = p + off_w3(D, H);
  N.b3 = N.w3 + H;
  const bf16x8* pk = (const bf16x8*)m->packs;
  const long long s0 = (long long)(H / 16) * 64, s1 = (long long)(H / 16) * (H / 32) * 64;
  N.Wf[0] = pk;
  N.Wf[1] = pk + s0;
  N.Wf[2] = pk + s0 + s1;
  N.Wt[0] = pk + s0 + 2 * s1;
  N.Wt[1] = pk + s0 + 3 * s1;
  N.use_range = m->use_range;
  if (!build_sorting(m, N)) return false;
  for (int k = 0; k < 32; ++k) {
    N.xmean[k] = m->xmean[k];
    N.xinv[k] = m->xinv[k];
    N.lo[k] = m->lo[k];
    N.hi[k] = m->hi[k];
  }
  N.ymean = m->ymean;
  for (int k = 0; k < MLP_NQ; ++k) N.quad[k] = m->quad[k];
  N.ystd = m->ystd;
  return true;
}

// ------------------------------------------------------------------------------------
// Inference on walkers, H = 128: persistent workgroups keep all three weight packs (72 KB),
// the biases and the output weights in LDS for the whole launch and loop over walker
// chunks (E_WAVES waves x NT_E tiles of 16 walkers).  Each A fragment read from LDS feeds NT_E MFMAs; the
// biases enter as the MFMA accumulator input.  Walkers outside the training range get the
// FALLBACK_MARK bit pattern, which k_mlp_fallback replaces with the analytic potential.
constexpr int NT_E = 2;
constexpr int E_WAVES = 8;  // waves per workgroup sharing one LDS copy of the weights
constexpr unsigned long long FALLBACK_MARK = 0x7FF4D1CEFA11BAC4ull;  // a signalling-NaN payload

template <int H>
constexpr size_t energy_lds_bytes() {
  return (size_t)((H / 16) * 64 + 2 * (H / 16) * (H / 32) * 64) * 16 + (size_t)(4 * H + 4) * 4;
}

template <typename R, int NP, int H>
__global__ __launch_bounds__(64 * E_WAVES, 2) void k_mlp_energy(MlpNet M, const R* planes, long long stride, const int* n_dev,
                                                       long long n_fix, double* V) {
  extern __shared__ bf16x8 lds[];
  constexpr int MT = H / 16, KC = H / 32;
  bf16x8* L0 = lds;
  bf16x8* L1 = L0 + MT * 64;
  bf16x8* L2 = L1 + MT * KC * 64;
  float* lb = (float*)(L2 + MT * KC * 64);  // b0 | b1 | b2 | w3 | b3
  for (int q = threadIdx.x; q < MT * 64; q += 64 * E_WAVES) L0[q] = M.Wf[0][q];
  for (int q = threadIdx.x; q < MT * KC * 64; q += 64 * E_WAVES) {
    L1[q] = M.Wf[1][q];
    L2[q] = M.Wf[2][q];
  }
  for (int q = threadIdx.x; q < H; q += 64 * E_WAVES) {
    lb[q] = M.b[0][q];
    lb[H + q] = M.b[1][q];
    lb[2 * H + q] = M.b[2][q];
    lb[3 * H + q] = M.w3[q];
  }
  if (threadIdx.x == 0) lb[4 * H] = M.b3[0];
  __syncthreads();
  const int n = n_dev ? *n_dev : (int)n_fix;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int g = lane >> 4, col = lane & 15;
  constexpr long long CHUNK = (long long)E_WAVES * 16 * NT_E;
  for (long long c0 = (long long)blockIdx.x * CHUNK; c0 < n; c0 += (long long)gridDim.x * CHUNK) {
    // compiler-only barrier: the weight fragments are re-read from LDS each chunk instead of
    // being hoisted out of the loop into (spilled) registers
    asm volatile("" ::: "memory");
    const long long base = c0 + (long long)wave * 16 * NT_E;
    constexpr int NPR = NP / 3 * (NP / 3 - 1) / 2;
    bf16x8 bin0[NT_E][1];
    bf16x8 bin[NT_E][KC];
    bool fbk[NT_E];
    float qb[NT_E];
#pragma unroll
    for (int t = 0; t < NT_E; ++t) {
      const long long w = base + t * 16 + col;
      const bool ok = w < n;
      float f[MLP_MAXP];
      R x[NP];
#pragma unroll
      for (int p = 0; p < NP; ++p) x[p] = ok ? planes[p * stride + w] : R(0);
      mlp_descriptor<R, NP>(M, x, f);
      bool out = false;
      if (M.use_range) {
#pragma unroll
        for (int k = 0; k < NPR; ++k) out |= !(f[k] >= M.lo[k] && f[k] <= M.hi[k]);
      }
      fbk[t] = out;
      bin0[t][0] = input_chunk<NPR>(M, f, g, ok);
      qb[t] = 0.0f;
      if (g == 0) {
        float xn[NPR];
#pragma unroll
        for (int k = 0; k < NPR; ++k) xn[k] = (f[k] - M.xmean[k]) * M.xinv[k];
        qb[t] = quad_baseline<NPR>(M, xn);
      }
    }
    f32x4 acc[NT_E][MT];
    float part[NT_E];
#pragma unroll
    for (int layer = 0; layer < 3; ++layer) {
      asm volatile("" ::: "memory");
      const float* bl = lb + layer * H;
#pragma unroll
      for (int m = 0; m < MT; ++m) {
        const f32x4 bias = *(const f32x4*)(bl + 16 * m + 4 * g);
        if (layer == 0) {
          const bf16x8 a = L0[m * 64 + lane];
#pragma unroll
          for (int t = 0; t < NT_E; ++t) acc[t][m] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bin0[t][0], bias, 0, 0, 0);
        } else {
          const bf16x8* Wl = layer == 1 ? L1 : L2;
#pragma unroll
          for (int c = 0; c < KC; ++c) {
            const bf16x8 a = Wl[(m * KC + c) * 64 + lane];
#pragma unroll
            for (int t = 0; t < NT_E; ++t)
              acc[t][m] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bin[t][c], c == 0 ? bias : acc[t][m], 0, 0, 0);
          }
        }
      }
      if (layer < 2) {
#pragma unroll
        for (int t = 0; t < NT_E; ++t)
#pragma unroll
          for (int m = 0; m < MT; ++m)
#pragma unroll
            for (int i = 0; i < 4; ++i) bin[t][m >> 1][(m & 1) * 4 + i] = (__bf16)silu_f(acc[t][m][i]);
      } else {
#pragma unroll
        for (int t = 0; t < NT_E; ++t) {
          float s = 0.0f;
#pragma unroll
          for (int m = 0; m < MT; ++m) {
            const f32x4 w3 = *(const f32x4*)(lb + 3 * H + 16 * m + 4 * g);
#pragma unroll
            for (int i = 0; i < 4; ++i) s += w3[i] * silu_f(acc[t][m][i]);
          }
          part[t] = s;
        }
      }
    }
    const float b3 = lb[4 * H];
#pragma unroll
    for (int t = 0; t < NT_E; ++t) {
      float s = part[t];
      s += __shfl_xor(s, 16, 64);
      s += __shfl_xor(s, 32, 64);
      const long long w = base + t * 16 + col;
      if (g == 0 && w < n) {
        const float y = s + b3;
        V[w] = fbk[t] ? __longlong_as_double((long long)FALLBACK_MARK) : (double)(y * M.ystd + M.ymean + qb[t]);
      }
    }
  }
}

// Inference on walkers when the two hidden-layer packs do not fit LDS together (H = 256:
// 128 KB each).  Persistent workgroups of E_WAVES waves keep layer 0's pack, the biases and
// the output weights in LDS for the whole launch, and stage the packs of layers 1 and 2 in
// turn through one slot, once per chunk of WV x 16 x NTS walkers (the packs come from L2;
// the chunk's activations stay in registers between the layers).
template <int H>
constexpr size_t staged_lds_bytes() {
  return (size_t)((H / 16) * 64 + (H / 16) * (H / 32) * 64) * 16 + (size_t)(4 * H + 4) * 4;
}

template <typename R, int NP, int H, int NTS, int WV>
__global__ __launch_bounds__(64 * WV, 1) void k_mlp_energy_st(MlpNet M, const R* planes, long long stride,
                                                                  const int* n_dev, long long n_fix, double* V) {
  extern __shared__ bf16x8 lds[];
  constexpr int MT = H / 16, KC = H / 32, NTH = 64 * WV;
  bf16x8* L0 = lds;
  bf16x8* LW = L0 + MT * 64;
  float* lb = (float*)(LW + MT * KC * 64);  // b0 | b1 | b2 | w3 | b3
  for (int q = threadIdx.x; q < MT * 64; q += NTH) L0[q] = M.Wf[0][q];
  for (int q = threadIdx.x; q < H; q += NTH) {
    lb[q] = M.b[0][q];
    lb[H + q] = M.b[1][q];
    lb[2 * H + q] = M.b[2][q];
    lb[3 * H + q] = M.w3[q];
  }
  if (threadIdx.x == 0) lb[4 * H] = M.b3[0];
  __syncthreads();
  const int n = n_dev ? *n_dev : (int)n_fix;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int g = lane >> 4, col = lane & 15;
  constexpr long long CHUNK = (long long)WV * 16 * NTS;
  // the loop bounds depend on the workgroup only: every wave meets every barrier
  for (long long c0 = (long long)blockIdx.x * CHUNK; c0 < n; c0 += (long long)gridDim.x * CHUNK) {
    asm volatile("" ::: "memory");
    const long long base = c0 + (long long)wave * 16 * NTS;
    constexpr int NPR = NP / 3 * (NP / 3 - 1) / 2;
    bf16x8 bin0[NTS][1];
    bf16x8 bin[NTS][KC];
    bool fbk[NTS];
    float qb[NTS];
#pragma unroll
    for (int t = 0; t < NTS; ++t) {
      const long long w = base + t * 16 + col;
      const bool ok = w < n;
      float f[MLP_MAXP];
      R x[NP];
#pragma unroll
      for (int p = 0; p < NP; ++p) x[p] = ok ? planes[p * stride + w] : R(0);
      mlp_descriptor<R, NP>(M, x, f);
      bool out = false;
      if (M.use_range) {
#pragma unroll
        for (int k = 0; k < NPR; ++k) out |= !(f[k] >= M.lo[k] && f[k] <= M.hi[k]);
      }
      fbk[t] = out;
      bin0[t][0] = input_chunk<NPR>(M, f, g, ok);
      qb[t] = 0.0f;
      if (g == 0) {
        float xn[NPR];
#pragma unroll
        for (int k = 0; k < NPR; ++k) xn[k] = (f[k] - M.xmean[k]) * M.xinv[k];
        qb[t] = quad_baseline<NPR>(M, xn);
      }
    }
    f32x4 acc[NTS][MT];
    float part[NTS];
    // layer 0 from the resident pack
#pragma unroll
    for (int m = 0; m < MT; ++m) {
      const f32x4 bias = *(const f32x4*)(lb + 16 * m + 4 * g);
      const bf16x8 a = L0[m * 64 + lane];
#pragma unroll
      for (int t = 0; t < NTS; ++t) acc[t][m] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bin0[t][0], bias, 0, 0, 0);
    }
#pragma unroll
    for (int t = 0; t < NTS; ++t)
#pragma unroll
      for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int i = 0; i < 4; ++i) bin[t][m >> 1][(m & 1) * 4 + i] = (__bf16)silu_f(acc[t][m][i]);
    // layers 1 and 2 through the staging slot
#pragma unroll
    for (int layer = 1; layer < 3; ++layer) {
      __syncthreads();  // every wave is done with the slot
      const bf16x8* src = M.Wf[layer];
      for (int q = threadIdx.x; q < MT * KC * 64; q += NTH) LW[q] = src[q];
      __syncthreads();
      const float* bl = lb + layer * H;
      // input chunk outer, output tile inner: consecutive MFMAs write different
      // accumulators, so none waits on the previous one's result
#pragma unroll
      for (int c = 0; c < KC; ++c) {
#pragma unroll
        for (int m = 0; m < MT; ++m) {
          const bf16x8 a = LW[(m * KC + c) * 64 + lane];
#pragma unroll
          for (int t = 0; t < NTS; ++t)
            acc[t][m] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                a, bin[t][c], c == 0 ? *(const f32x4*)(bl + 16 * m + 4 * g) : acc[t][m], 0, 0, 0);
        }
      }
      if (layer == 1) {
#pragma unroll
        for (int t = 0; t < NTS; ++t)
#pragma unroll
          for (int m = 0; m < MT; ++m)
#pragma unroll
            for (int i = 0; i < 4; ++i) bin[t][m >> 1][(m & 1) * 4 + i] = (__bf16)silu_f(acc[t][m][i]);
      } else {
#pragma unroll
        for (int t = 0; t < NTS; ++t) {
          float s = 0.0f;
#pragma unroll
          for (int m = 0; m < MT; ++m) {
            const f32x4 w3 = *(const f32x4*)(lb + 3 * H + 16 * m + 4 * g);
#pragma unroll
            for (int i = 0; i < 4; ++i) s += w3[i] * silu_f(acc[t][m][i]);
          }
          part[t] = s;
        }
      }
    }
    const float b3 = lb[4 * H];
#pragma unroll
    for (int t = 0; t < NTS; ++t) {
      float s = part[t];
      s += __shfl_xor(s, 16, 64);
      s += __shfl_xor(s, 32, 64);
      const long long w = base + t * 16 + col;
      if (g == 0 && w < n) {
        const float y = s + b3;
        V[w] = fbk[t] ? __longlong_as_double((long long)FALLBACK_MARK) : (double)(y * M.ystd + M.ymean + qb[t]);
      }
    }
  }
}

// analytic potential for the walkers the surrogate marked as outside its training range
template <typename R, int NP>
__global__ __launch_bounds__(256) void k_mlp_fallback(const R* planes, long long stride, const int* n_dev,
                                                      long long n_fix, double* V, PesParams fb) {
  const int n = n_dev ? *n_dev : (int)n_fix;
  for (long long w = (long long)blockIdx.x * 256 + threadIdx.x; w < n; w += (long long)gridDim.x * 256) {
    if ((unsigned long long)__double_as_longlong(V[w]) != FALLBACK_MARK) continue;
    R x[NP];
#pragma unroll
    for (int p = 0; p < NP; ++p) x[p] = planes[p * stride + w];
    V[w] = (double)pes_eval<R, NP>(fb, x);
  }
}

// Error judge: analytic potential on every `every`-th walker against the surrogate's V
// (one workgroup, fixed-order reduction); record = {step, sum of squared errors, max |error|,
// samples} in hartree.
template <typename R, int NP>
__global__ __launch_bounds__(1024) void k_mlp_judge(const R* planes, long long stride, const int* n_dev,
                                                    long long n_fix, const double* V, int every, PesParams pes,
                                                    long long step, double* rec) {
  __shared__ double s_a[16], s_b[16];
  __shared__ int s_c[16];
  const int n = n_dev ? *n_dev : (int)n_fix;
  double sse = 0.0, mx = 0.0;
  int cnt = 0;
  for (long long w = (long long)threadIdx.x * every; w < n; w += 1024ll * every) {
    R x[NP];
#pragma unroll
    for (int p = 0; p < NP; ++p) x[p] = planes[p * stride + w];
    const double d = V[w] - (double)pes_eval<R, NP>(pes, x);
    sse += d * d;
    mx = fmax(mx, fabs(d));
    ++cnt;
  }
  const int wv = threadIdx.x >> 6;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    sse += __shfl_xor(sse, o, 64);
    mx = fmax(mx, __shfl_xor(mx, o, 64));
    cnt += __shfl_xor(cnt, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    s_a[wv] = sse;
    s_b[wv] = mx;
    s_c[wv] = cnt;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0.0, b = 0.0;
    int c = 0;
    for (int k = 0; k < 16; ++k) {
      a += s_a[k];
      b = fmax(b, s_b[k]);
      c += s_c[k];
    }
    rec[0] = (double)step;
    rec[1] = a;
    rec[2] = b;
    rec[3] = (double)c;
  }
}

constexpr int NT_FWD = 2;

template <typename R, int NP, int H>
int launch_energy(const MlpNet& N, const void* planes, long long stride, const int* n_dev, long long n_max,
                  double* V, const PesParams& fb, hipStream_t st) {
  if constexpr (H == 128) {
    // persistent: two workgroups per CU (LDS-bound), each looping over walker chunks
    int dev = 0, cus = 256;
    MHIP(hipGetDevice(&dev));
    MHIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const long long chunks = (n_max + (long long)E_WAVES * 16 * NT_E - 1) / ((long long)E_WAVES * 16 * NT_E);
    const unsigned grid = (unsigned)std::max(1ll, std::min(chunks, 2ll * cus));
    const size_t lds = energy_lds_bytes<H>();
    auto k = k_mlp_energy<R, NP, H>;
    MHIP(hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(k, dim3(grid), dim3(64 * E_WAVES), lds, st, N, (const R*)planes, stride, n_dev, n_max, V);
    MHIP(hipGetLastError());
    if (N.use_range) {
      const unsigned gf = (unsigned)std::max(1ll, std::min((n_max + 255) / 256, 4ll * cus));
      hipLaunchKernelGGL((k_mlp_fallback<R, NP>), dim3(gf), dim3(256), 0, st, (const R*)planes, stride, n_dev, n_max,
                         V, fb);
      MHIP(hipGetLastError());
    }
    return 0;
  } else {
  // H = 256: persistent, one workgroup per CU (LDS-bound), layers 1 and 2 staged per chunk
  int dev = 0, cus = 256;
  MHIP(hipGetDevice(&dev));
  MHIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  // one 16-walker tile per wave, 8 waves: two tiles per wave at one wave per SIMD (no spills)
  // measured 18.5 ms against 14.0 ms at 16M walkers (too little latency hiding)
  constexpr int NTS = 1, WV = E_WAVES;
  const long long chunks = (n_max + (long long)WV * 16 * NTS - 1) / ((long long)WV * 16 * NTS);
  const unsigned grid = (unsigned)std::max(1ll, std::min(chunks, (long long)cus));
  const size_t lds = staged_lds_bytes<H>();
  auto k = k_mlp_energy_st<R, NP, H, NTS, WV>;
  MHIP(hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(k, dim3(grid), dim3(64 * WV), lds, st, N, (const R*)planes, stride, n_dev, n_max, V);
  MHIP(hipGetLastError());
  if (N.use_range) {
    const unsigned gf = (unsigned)std::max(1ll, std::min((n_max + 255) / 256, 4ll * cus));
    hipLaunchKernelGGL((k_mlp_fallback<R, NP>), dim3(gf), dim3(256), 0, st, (const R*)planes, stride, n_dev, n_max,
                       V, fb);
    MHIP(hipGetLastError());
  }
  return 0;
  }
}

}  // namespace

extern "C" {

const char* pvd_mlp_last_error(void) { return g_merr.c_str(); }

int64_t pvd_mlp_param_count(int32_t D, int32_t H) { return off_w3(D, H) + H + 1; }

int64_t pvd_mlp_pack_bytes(int32_t H) {
  return ((int64_t)(H / 16) * 64 + 4ll * (H / 16) * (H / 32) * 64) * 16;
}

int pvd_mlp_pack(const pvd_mlp* m, void* stream) {
  if (!m || !m->params || !m->packs) return mfail(7, "null argument");
  const int D = m->D, H = m->H;
  if (H != 128 && H != 256) return mfail(7, "hidden width must be 128 or 256");
  hipStream_t st = (hipStream_t)stream;
  const float* p = (const float*)m->params;
  bf16x8* pk = (bf16x8*)m->packs;
  const int MT = H / 16, KC = H / 32;
  const long long s0 = (long long)MT * 64, s1 = (long long)MT * KC * 64;
  hipLaunchKernelGGL(k_mlp_pack, dim3((unsigned)((s0 + 255) / 256)), dim3(256), 0, st, p + off_W(0, D, H), H, D, 1, 0,
                     MT, 1, pk);
  hipLaunchKernelGGL(k_mlp_pack, dim3((unsigned)((s1 + 255) / 256)), dim3(256), 0, st, p + off_W(1, D, H), H, H, 0, 0,
                     MT, KC, pk + s0);
  hipLaunchKernelGGL(k_mlp_pack, dim3((unsigned)((s1 + 255) / 256)), dim3(256), 0, st, p + off_W(2, D, H), H, H, 0, 0,
                     MT, KC, pk + s0 + s1);
  hipLaunchKernelGGL(k_mlp_pack, dim3((unsigned)((s1 + 255) / 256)), dim3(256), 0, st, p + off_W(1, D, H), H, H, 0, 1,
                     MT, KC, pk + s0 + 2 * s1);
  hipLaunchKernelGGL(k_mlp_pack, dim3((unsigned)((s1 + 255) / 256)), dim3(256), 0, st, p + off_W(2, D, H), H, H, 0, 1,
                     MT, KC, pk + s0 + 3 * s1);
  MHIP(hipGetLastError());
  return 0;
}

int pvd_mlp_descriptors(const pvd_mlp* m, int32_t precision, int32_t n_atoms, const void* coords, int64_t n,
                        int32_t normalise, float* features_t, void* stream) {
  MlpNet N;
  if (!m || !coords || !features_t || !build_net(m, N)) return mfail(7, "bad network description");
  if (n_atoms != m->n_atoms || n_atoms < 2)
    return mfail(7, "the network's descriptor is for another number of atoms");
  if (n <= 0) return 0;
  hipStream_t st = (hipStream_t)stream;
  const unsigned grid = (unsigned)((n + 255) / 256);
  const int np = n_atoms * 3;
  if (precision == PVD_FP64) {
    if (np == 9) hipLaunchKernelGGL((k_mlp_desc<double, 9>), dim3(grid), dim3(256), 0, st, N, (const double*)coords, n, features_t);
    else if (np == 18) hipLaunchKernelGGL((k_mlp_desc<double, 18>), dim3(grid), dim3(256), 0, st, N, (const double*)coords, n, features_t);
    else return mfail(7, "descriptor kernels support 3 or 6 atoms");
  } else {
    if (np == 9) hipLaunchKernelGGL((k_mlp_desc<float, 9>), dim3(grid), dim3(256), 0, st, N, (const float*)coords, n, features_t);
    else if (np == 18) hipLaunchKernelGGL((k_mlp_desc<float, 18>), dim3(grid), dim3(256), 0, st, N, (const float*)coords, n, features_t);
    else return mfail(7, "descriptor kernels support 3 or 6 atoms");
  }
  if (normalise)
    hipLaunchKernelGGL(k_mlp_normalise, dim3((unsigned)((N.D * n + 255) / 256)), dim3(256), 0, st, N, (long long)n,
                       features_t);
  MHIP(hipGetLastError());
  return 0;
}

int pvd_mlp_energy(const pvd_mlp* m, int32_t precision, int32_t n_atoms, const void* planes, int64_t stride,
                   const int32_t* n_dev, int64_t n_max, double* v_out, const pvd_pes_desc* fallback, void* stream) {
  MlpNet N;
  if (!m || !planes || !v_out || !build_net(m, N)) return mfail(7, "bad network description");
  if (n_atoms != m->n_atoms || n_atoms < 2)
    return mfail(7, "the network's descriptor is for another number of atoms");
  if (n_max <= 0) return 0;
  PesParams fb;
  std::memset(&fb, 0, sizeof(fb));
  if (fallback) std::memcpy(&fb, fallback, sizeof(fb));
  if (N.use_range && !fallback) return mfail(7, "range check needs a fallback potential");
  hipStream_t st = (hipStream_t)stream;
  const int np = n_atoms * 3;
#define PVD_E(R, NP, H) return launch_energy<R, NP, H>(N, planes, stride, n_dev, n_max, v_out, fb, st)
  if (precision == PVD_FP64) {
    if (np == 9 && N.H == 128) PVD_E(double, 9, 128);
    if (np == 9 && N.H == 256) PVD_E(double, 9, 256);
    if (np == 18 && N.H == 128) PVD_E(double, 18, 128);
    if (np == 18 && N.H == 256) PVD_E(double, 18, 256);
  } else {
    if (np == 9 && N.H == 128) PVD_E(float, 9, 128);
    if (np == 9 && N.H == 256) PVD_E(float, 9, 256);
    if (np == 18 && N.H == 128) PVD_E(float, 18, 128);
    if (np == 18 && N.H == 256) PVD_E(float, 18, 256);
  }
#undef PVD_E
  return mfail(7, "surrogate supports 3 or 6 atoms and hidden width 128 or 256");
}

int pvd_mlp_judge(const pvd_pes_desc* pes, int32_t precision, int32_t n_atoms, const void* planes, int64_t stride,
                  const int32_t* n_dev, int64_t n_max, const double* v, int32_t every, int64_t step, double* record,
                  void* stream) {
  if (!pes || !planes || !v || !record || every < 1) return mfail(7, "bad arguments");
  PesParams P;
  std::memcpy(&P, pes, sizeof(P));
  hipStream_t st = (hipStream_t)stream;
  const int np = n_atoms * 3;
#define PVD_J(R, NP)                                                                                 \
  hipLaunchKernelGGL((k_mlp_judge<R, NP>), dim3(1), dim3(1024), 0, st, (const R*)planes, stride, n_dev, n_max, v, \
                     every, P, step, record)
  if (precision == PVD_FP64) {
    if (np == 9) PVD_J(double, 9);
    else if (np == 18) PVD_J(double, 18);
    else return mfail(7, "judge supports 3 or 6 atoms");
  } else {
    if (np == 9) PVD_J(float, 9);
    else if (np == 18) PVD_J(float, 18);
    else return mfail(7, "judge supports 3 or 6 atoms");
  }
#undef PVD_J
  MHIP(hipGetLastError());
  return 0;
}

// Gradients of the mean squared error of the network on (X, T) into grads (flat, the
// parameter order), plus the loss.  ws: device workspace of pvd_mlp_workspace_bytes().
int64_t pvd_mlp_workspace_bytes(int32_t D, int32_t H, int64_t batch) {
  const int64_t n = batch;
  const int64_t splits = (n + 255) / 256;
  const int64_t P = splits * (int64_t)H * (H > D ? H : D);
  return (3 * H * n * 2 + 2 * n + P + (n + 255) / 256 * 2) * 4 + 64;
}

}  // extern "C"

namespace {

// Scratch of pvd_mlp_grads_noalloc: one array per device, held until the process ends.
// Host code never uses the name `scratch` as an address: the device address comes from
// hipGetSymbolAddress, looked up once per device with that device current.
constexpr long long PVD_MLP_SCRATCH_BYTES = 32ll << 20;
static __device__ __align__(256) unsigned char scratch[PVD_MLP_SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch[MAX_DEVICES];

int scratch_for_current_device(void** out) {
  int dev = 0;
  MHIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= MAX_DEVICES) return mfail(7, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) return mfail(8, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch[dev] = ptr;
  }
  *out = g_scratch[dev];
  return 0;
}

int mlp_grads(const pvd_mlp* m, const float* x_t, const float* targets, int64_t n, float* grads, float* y_out,
              double* loss_parts, void* workspace, void* stream);

}  // namespace

extern "C" {

int pvd_mlp_grads(const pvd_mlp* m, const float* x_t, const float* targets, int64_t n, float* grads, float* y_out,
                  double* loss_parts, void* workspace, void* stream) {
  if (!workspace) return mfail(7, "bad arguments");
  return mlp_grads(m, x_t, targets, n, grads, y_out, loss_parts, workspace, stream);
}

int64_t pvd_mlp_noalloc_scratch_bytes(void) { return PVD_MLP_SCRATCH_BYTES; }

int pvd_mlp_grads_noalloc(const pvd_mlp* m, const float* x_t, const float* targets, int64_t n, float* grads,
                          float* y_out, double* loss_parts, void* stream) {
  if (!m) return mfail(7, "bad arguments");
  if (n <= 0) return mfail(7, "empty batch");
  const int64_t need = pvd_mlp_workspace_bytes(m->D, m->H, n);
  if (need > PVD_MLP_SCRATCH_BYTES)
    return mfail(7, "batch too large for the static scratch: needs " + std::to_string(need) + " bytes, the maximum is " +
                        std::to_string(PVD_MLP_SCRATCH_BYTES));
  void* ws = nullptr;
  const int rc = scratch_for_current_device(&ws);
  if (rc) return rc;
  return mlp_grads(m, x_t, targets, n, grads, y_out, loss_parts, ws, stream);
}

}  // extern "C"

namespace {

int mlp_grads(const pvd_mlp* m, const float* x_t, const float* targets, int64_t n, float* grads, float* y_out,
              double* loss_parts, void* workspace, void* stream) {
  MlpNet N;
  if (!m || !x_t || !targets || !grads || !workspace || !build_net(m, N)) return mfail(7, "bad arguments");
  if (n <= 0) return mfail(7, "empty batch");
  hipStream_t st = (hipStream_t)stream;
  const int D = N.D, H = N.H;
  float* ZT = (float*)workspace;
  float* DZT = ZT + 3ll * H * n;
  float* Y = DZT + 3ll * H * n;
  float* dY = Y + n;
  float* P = dY + n;
  const long long per_block = 4ll * 16 * NT_FWD;
  const unsigned grid = (unsigned)((n + per_block - 1) / per_block);
  const size_t lds = (size_t)(H / 16) * (H / 32) * 64 * 16;
  PesParams fb;
  std::memset(&fb, 0, sizeof(fb));
  if (H == 128) {
    auto kf = k_mlp_forward<float, 9, 128, NT_FWD, 1>;
    auto kb = k_mlp_backward<128, NT_FWD>;
    MHIP(hipFuncSetAttribute((const void*)kf, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    MHIP(hipFuncSetAttribute((const void*)kb, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(kf, dim3(grid), dim3(256), lds, st, N, (const float*)nullptr, 0ll, (const int*)nullptr,
                       (long long)n, (double*)nullptr, fb, x_t, ZT, Y);
    hipLaunchKernelGGL(k_mlp_loss, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, (long long)n, Y, targets, dY,
                       loss_parts);
    hipLaunchKernelGGL(kb, dim3(grid), dim3(256), lds, st, N, (long long)n, ZT, dY, DZT);
  } else {
    auto kf = k_mlp_forward<float, 9, 256, 1, 1>;
    auto kb = k_mlp_backward<256, 1>;
    const unsigned grid1 = (unsigned)((n + 63) / 64);
    MHIP(hipFuncSetAttribute((const void*)kf, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    MHIP(hipFuncSetAttribute((const void*)kb, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(kf, dim3(grid1), dim3(256), lds, st, N, (const float*)nullptr, 0ll, (const int*)nullptr,
                       (long long)n, (double*)nullptr, fb, x_t, ZT, Y);
    hipLaunchKernelGGL(k_mlp_loss, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, (long long)n, Y, targets, dY,
                       loss_parts);
    hipLaunchKernelGGL(kb, dim3(grid1), dim3(256), lds, st, N, (long long)n, ZT, dY, DZT);
  }
  const int kslice = 256;
  const int splits = (int)((n + kslice - 1) / kslice);
  for (int l = 0; l < 3; ++l) {
    const int Din = l == 0 ? D : H;
    dim3 g3((unsigned)(H / 32), (unsigned)((Din + 31) / 32), (unsigned)splits);
    hipLaunchKernelGGL(k_mlp_wgrad, g3, dim3(64), 0, st, l, H, Din, (long long)n, kslice, x_t, ZT, DZT, P);
    const long long size = (long long)H * Din;
    hipLaunchKernelGGL(k_mlp_wgrad_reduce, dim3((unsigned)((size + 255) / 256)), dim3(256), 0, st, P, splits, size,
                       grads + off_W(l, D, H));
  }
  hipLaunchKernelGGL(k_mlp_rowsum, dim3((unsigned)(4 * H + 1)), dim3(256), 0, st, H, (long long)n, ZT, DZT, dY,
                     grads + off_b(0, D, H), grads + off_b(1, D, H), grads + off_b(2, D, H), grads + off_w3(D, H),
                     grads + off_w3(D, H) + H);
  if (y_out) MHIP(hipMemcpyAsync(y_out, Y, sizeof(float) * n, hipMemcpyDeviceToDevice, st));
  MHIP(hipGetLastError());
  return 0;
}

}  // namespace

extern "C" {

int pvd_mlp_adam(float* params, float* m1, float* m2, const float* grads, int64_t size, float lr, float beta1,
                 float beta2, float eps, int64_t t, void* stream) {
  if (!params || !m1 || !m2 || !grads || size <= 0 || t <= 0) return mfail(7, "bad arguments");
  const float c1 = 1.0f - powf(beta1, (float)t), c2 = 1.0f - powf(beta2, (float)t);
  hipLaunchKernelGGL(k_adam, dim3((unsigned)((size + 255) / 256)), dim3(256), 0, (hipStream_t)stream, params, m1, m2,
                     grads, (long long)size, lr, beta1, beta2, eps, c1, c2);
  MHIP(hipGetLastError());
  return 0;
}

}  // extern "C"
