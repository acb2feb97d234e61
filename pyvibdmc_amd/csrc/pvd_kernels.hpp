// DMC step kernels.
//
// Walkers live in HBM as structure-of-arrays planes: component p of walker slot i is at
// coords[p * cap + i] (p = atom*d + c), double-buffered for the discrete branch.  One
// step of the discrete algorithm (reference pyvibdmc.py:540-547, 786-793, 391-431,
// 651-661) is ONE launch of k_discrete: each tile (one workgroup, 4 walkers per thread)
// moves its walkers (Philox + Box-Muller, or injected draws), evaluates the PES in
// registers, forms the branching weight and copy count, and writes its survivors to their
// final slots in np.repeat order.  The output offset of a tile is the sum of the published
// counts of its predecessors, read in two levels (tiles within a 64-tile group, then group
// totals).  Counts are 64-bit words (epoch | value) written and read with relaxed
// agent-scope atomics, so publishing needs no fence: on CDNA an agent-scope release writes
// back the whole L2 of the XCD, which per tile would cost more than the step itself.  Each
// tile also publishes a record of its partial sums, which the last tile reduces in a fixed
// order into vref and the population of the next step.  All control values are
// double-buffered by step parity `par`, which is a launch argument, so a launch never reads
// a value that the same launch writes.
#pragma once
#include "pvd_models.hpp"

namespace pvd {

constexpr int TILE = 256;
constexpr int MAXNP = 96;  // components per walker on the runtime-count paths

enum Status : int {
  ST_OK = 0,
  ST_POP = 1,          // population outside [0.5, 1.5] N0       (pyvibdmc.py:409-413, 714-717)
  ST_WEIGHT = 3,       // non-finite or huge branching weight    (pyvibdmc.py:397-400)
  ST_CAPACITY = 4,     // more walkers than slots
  ST_BRANCH = 5,       // continuous split outside supported sizes
  ST_SPIN = 6,         // a bounded device wait expired
};

struct Ctrl {
  int n[2];
  double vref[2];
  unsigned long long step[2];
  unsigned epoch[2];
  unsigned ticket[2];
  unsigned done[2];
  int kills[2];
  double sumw[2];
  double sumwe[2];
  int status;
  int pad0;
  long long status_step;
  double eff_t;
  unsigned list_n;
  unsigned pad1;
  int ntd[2];
  // continuous split replay: the weight kernel appends the walkers of weight >= cand_t to
  // list A and those in [cand_t2, cand_t) to list B (0: no lists; cand_t2 = cand_t: no
  // list B), each in 8 sub-lists (by workgroup index) of up to SUB_CAP entries: entry o of
  // sub-list sub of list `list` is at cand[(list * 8 + sub) * SUB_CAP + o];
  // cand_t = cut * cand_g,
  // where cut is the last step's smallest halved weight and cand_g is fed back from the
  // size of list A; the entry counts of the sub-lists are in StepArgs::ncnt (see list_count)
  double cand_t;
  double cand_t2;
  int ncopy;  // split copies listed for the copy kernel (candidate replay), else 0
  float cand_g;
  int cand_b;  // steps list B stays on (it is kept after list A alone fell short)
  // the global-memory candidate replay leaves its kill slots' sources and weights to the copy
  // kernel: gc_n sorted candidates in the replay scratch (0: the copies are (slot, source)
  // pairs in cpy), and the family where the kills ran out (sorted index, slot, weight, split
  // depth, first kill index in its last octave)
  int gc_n;
  int gc_p;
  int gc_slot;
  int gc_dp;
  int gc_start;
  int pad2;
  double gc_w;
};

struct TileRec {
  double s0;  // discrete: sum c*E | continuous: sum w*E
  double s1;  // continuous: sum w | discrete: unused
  double s2;  // continuous: sum w*E over the walkers not killed
  double s3;  // continuous: sum w over the walkers not killed
  double mx;  // continuous: max weight in the tile
  int count;  // discrete: sum c   | continuous: kills
  int births;
  int deaths;
  int bad;
};

// A walker at or above the candidate threshold of the continuous split replay.
struct Cand {
  double w;
  int slot;
  int pad;
};
constexpr int CAND_CAP = 4096;   // candidates the replay sorts in LDS
constexpr int CAND_G = 131072;   // candidates the global-memory replay sorts (2 KILL_CAP)
// scratch of the global-memory replay: keys and bucketed keys (8 B each), order and bucketed
// order (4 B each), family starts (4 B), split depths (1 B) per candidate
constexpr size_t RSG_BYTES = (size_t)CAND_G * (8 + 8 + 4 + 4 + 4 + 1);
constexpr int LIST_SUB = 8;       // sub-lists per list (spreads the appends' counters)
constexpr int SUB_CAP = 32768;    // entries per sub-list and step
constexpr int LIST_CAP = LIST_SUB * SUB_CAP;
// sub-list entries the LDS replay reads: it keeps a candidate's position in 16 bits, the
// sub-list (4 bits) and the offset in it (12 bits)
constexpr int SUB_CAP16 = 4096;
constexpr int KILL_CAP = 65536;  // kills per step

struct Piece {
  double v;
  int slot;
  int src;
  int birth;  // kill index at which the piece was created (-1: an original walker)
  int pad;
};

template <typename R>
struct StepArgs {
  const R* cin;
  R* cout;
  long long cap;
  int np;              // components per walker (runtime copy of NP)
  const int* who_in;   // descendant tracker (nullptr when no window is open)
  int* who_out;
  double* w;           // continuous weights
  double* E;           // continuous: per-walker energy used for the weight
  double* ke;          // split IS path: local kinetic energy of each walker
  const double* V_ext; // split path: energies computed outside the kernel
  unsigned long long* kmask;  // continuous: per-wave kill masks
  Ctrl* ctrl;
  uint64_t* flags;
  uint64_t* gflags;  // discrete: per 64-tile group totals
  TileRec* rec;
  uint64_t* drec;      // discrete: 3 coherent words per tile (sum c*E | births,count | epoch,bad,deaths)
  unsigned* accs;      // importance sampling: accepted-move counters (see acc_slot)
  double* tr_vref;
  double* tr_pop;
  int* tr_births;
  int* tr_deaths;
  double* tr_dteff;
  long long tmax;
  const double* inj_disp;  // injected displacements (N, M, d) C order
  const double* inj_umet;  // injected Metropolis uniforms
  const double* inj_ubr;   // injected branching uniforms
  Piece* pend;             // continuous split scratch
  Piece* fin;
  int* klist;              // kill list of the step (ascending slots)
  unsigned long long* killtag;  // per slot: (step << 32) | kill index, valid for this step only
  double* rstage;          // staged source walkers of the split copies
  Cand* cand;              // candidate walkers of the split replay (unordered)
  unsigned char* rsg;      // scratch of the global-memory candidate replay (RSG_BYTES)
  unsigned* ncnt;          // continuous: entry counts of the candidate sub-lists (see list_count)
  int* asg;                // per kill index: the candidate (sorted position) it copies
  int* cpy;                // per kill index: (kill slot, source slot) for the copy kernel
  PesParams pes;
  TrialParams trial;
  double sigma[MAXNP];
  double inv_m[MAXNP];
  float h2s[MAXNP];  // 1 / (2 sigma^2), FP32 Metropolis ratio
  double dt, alpha, n0, pop_lo, pop_hi, thresh_lo;
  Rng rng;
  int par;
  int slots;   // discrete: resident workgroups of the step kernel (tiling unit)
  int branch;
  int ms, ds;  // atoms, dims
  int is_;     // importance sampling on
  // trial wave function evaluated by the caller: caches of the importance-sampled move as
  // FP64 planes (see k_his_propose), current and output buffers, plane count (0: none)
  const double* xc_in;
  double* xc_out;
  int nx;
  int htrial;
};

// ---------------------------------------------------------------------------- helpers
// Accepted-move counters of the importance-sampled step: ACC_SLOTS per parity, one per
// 128-byte line, so the per-block atomic adds of the acceptance pass spread over many
// addresses.  Integer sums: the total does not depend on the order of the adds.
constexpr int ACC_SLOTS = 64;
constexpr int ACC_STRIDE = 32;  // unsigned words between slots

template <typename R>
PVD_DEV unsigned* acc_slot(const StepArgs<R>& A, int par, int b) {
  return A.accs + (size_t)(par * ACC_SLOTS + (b & (ACC_SLOTS - 1))) * ACC_STRIDE;
}

// Entry counts of the continuous candidate sub-lists (list 0 = A, 1 = B), one per 256-byte
// line: every workgroup of the weight kernel that holds candidates adds its count to one of
// them, and counters sharing a line serialise those adds (at 1M walkers, 3 900 workgroups
// on one line cost the kernel 150 us).  Zero between steps: the replay reads and clears them.
constexpr int CNT_STRIDE = 64;  // unsigned words between counters
constexpr size_t CNT_BYTES = sizeof(unsigned) * 2 * LIST_SUB * CNT_STRIDE;

template <typename R>
PVD_DEV unsigned* list_count(const StepArgs<R>& A, int list, int sub) {
  return A.ncnt + (size_t)(list * LIST_SUB + sub) * CNT_STRIDE;
}

// Total accepted moves of parity `par`, on every lane of the calling wave.
template <typename R>
PVD_DEV unsigned acc_total(const StepArgs<R>& A, int par) {
  return (unsigned)wave_sum((int)ld_relaxed_u(acc_slot(A, par, lane_id())));
}

// FP32 coordinate buffers are addressed through one buffer resource each: a 32-bit lane
// offset (4 i) plus a uniform plane offset (4 p cap), so walker accesses carry no 64-bit
// address arithmetic.  The host keeps np * cap * 4 below 2^32 bytes on this path; the
// resource's record count bounds every access to the buffer.
PVD_DEV __amdgpu_buffer_rsrc_t coord_rsrc(const float* base, long long cap, int np) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(base), 0, (int)(uint32_t)(cap * np * 4),
                                           0x00020000);
}
PVD_DEV float coord_ld(__amdgpu_buffer_rsrc_t r, long long cap, int p, long long i) {
  return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, (int)(uint32_t)(i * 4),
                                                              (int)(uint32_t)(p * cap * 4), 0));
}
PVD_DEV void coord_st(__amdgpu_buffer_rsrc_t r, long long cap, int p, long long i, float v) {
  __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, (int)(uint32_t)(i * 4),
                                        (int)(uint32_t)(p * cap * 4), 0);
}

// walker i written back into the current (input) buffer
template <typename R, int NP>
PVD_DEV void store_walker_in(const StepArgs<R>& A, long long i, const R* x) {
  if constexpr (sizeof(R) == 4) {
    const __amdgpu_buffer_rsrc_t r = coord_rsrc((const float*)A.cin, A.cap, NP);
#pragma unroll
    for (int p = 0; p < NP; ++p) coord_st(r, A.cap, p, i, x[p]);
  } else {
#pragma unroll
    for (int p = 0; p < NP; ++p) const_cast<R*>(A.cin)[p * A.cap + i] = x[p];
  }
}

template <typename R, int NP>
PVD_DEV void load_walker(const StepArgs<R>& A, long long i, R* x) {
  if constexpr (sizeof(R) == 4) {
    const __amdgpu_buffer_rsrc_t r = coord_rsrc((const float*)A.cin, A.cap, NP);
#pragma unroll
    for (int p = 0; p < NP; ++p) x[p] = coord_ld(r, A.cap, p, i);
  } else {
#pragma unroll
    for (int p = 0; p < NP; ++p) x[p] = A.cin[p * A.cap + i];
  }
}

// Word layout of one walker's FP32 draws at one step (Philox stream RNG_STEP): NPAIR
// Box-Muller pairs, each from a 24-bit radius uniform (the high 24 bits of its own word)
// and a 16-bit angle uniform (16-bit halves of the following words); the 24-bit branching
// and Metropolis uniforms are assembled from the spare low bytes of the radius words and
// the unused angle half, with extra words only where those run short.
template <int NP, bool IS>
struct DrawLayout {
  static constexpr int NPAIR = (NP + 1) / 2;
  static constexpr int NANG = (NPAIR + 1) / 2;              // angle words
  static constexpr int SPARE = NPAIR + ((NPAIR & 1) ? 2 : 0);  // spare bytes without extras
  static constexpr int NEED = IS ? 6 : 3;
  static constexpr int NEXTRA = SPARE >= NEED ? 0 : (NEED - SPARE + 3) / 4;
  static constexpr int NW = NPAIR + NANG + NEXTRA;
};

// spare byte b of the layout: radius-word low bytes, then the free angle half, then extras
template <int NP, bool IS>
PVD_DEV uint32_t spare_byte(const uint32_t* w, int b) {
  using L = DrawLayout<NP, IS>;
  if (b < L::NPAIR) return w[b] & 0xFFu;
  b -= L::NPAIR;
  if (L::NPAIR & 1) {
    if (b < 2) return (w[L::NPAIR + L::NANG - 1] >> (16 + 8 * b)) & 0xFFu;
    b -= 2;
  }
  return (w[L::NPAIR + L::NANG + b / 4] >> (8 * (b % 4))) & 0xFFu;
}

// All draws of walker slot i at one step: the displacement disp = sigma * z, the
// Metropolis uniform and the branching uniform.  Injected values are read from the
// caller's arrays (the reference's draw order).  Philox, FP32: DrawLayout words from
// ceil(NW / 4) blocks of one stream per (slot, step) - 2 blocks for 9 components.
// Philox, FP64: 53-bit uniforms from separate blocks per purpose.
template <typename R, int NP, bool INJ, bool IS>
PVD_DEV void draw_step(const StepArgs<R>& A, long long i, unsigned long long step, R* disp, double& um,
                       double& ub, bool want_branch) {
  if constexpr (INJ) {
#pragma unroll
    for (int p = 0; p < NP; ++p) disp[p] = (R)A.inj_disp[i * NP + p];
    um = IS ? A.inj_umet[i] : 0.0;
    ub = want_branch ? A.inj_ubr[i] : 0.0;
  } else if constexpr (sizeof(R) == 4) {
    using L = DrawLayout<NP, IS>;
    uint32_t w[L::NW];
    philox_words<L::NW>(A.rng, (uint32_t)i, step, w);
#pragma unroll
    for (int k = 0; k < L::NPAIR; ++k) {
      const uint32_t ang = (w[L::NPAIR + k / 2] >> (16 * (k & 1))) & 0xFFFFu;
      float z0, z1;
      box_muller_24_16(w[k] >> 8, ang, z0, z1);
      disp[2 * k] = (R)((float)A.sigma[2 * k] * z0);
      if (2 * k + 1 < NP) disp[2 * k + 1] = (R)((float)A.sigma[2 * k + 1] * z1);
    }
    const uint32_t b24 = (spare_byte<NP, IS>(w, 0) << 16) | (spare_byte<NP, IS>(w, 1) << 8) | spare_byte<NP, IS>(w, 2);
    ub = (double)((float)b24 * 5.9604644775390625e-08f);
    if constexpr (IS) {
      const uint32_t m24 =
          (spare_byte<NP, IS>(w, 3) << 16) | (spare_byte<NP, IS>(w, 4) << 8) | spare_byte<NP, IS>(w, 5);
      um = (double)((float)m24 * 5.9604644775390625e-08f);
    } else {
      um = 0.0;
    }
  } else {
    R z[NP];
    Normals<R, NP>::draw(A.rng, (uint32_t)i, step, z);
#pragma unroll
    for (int p = 0; p < NP; ++p) disp[p] = (R)A.sigma[p] * z[p];
    um = IS ? uniform_d(A.rng, (uint32_t)i, step, RNG_METRO) : 0.0;
    ub = want_branch ? uniform_d(A.rng, (uint32_t)i, step, RNG_BRANCH) : 0.0;
  }
}

template <typename R, bool INJ>
PVD_DEV double draw_u(const StepArgs<R>& A, long long i, unsigned long long step, uint32_t purpose) {
  if constexpr (INJ) {
    return purpose == RNG_BRANCH ? A.inj_ubr[i] : A.inj_umet[i];
  } else {
    return uniform_d(A.rng, (uint32_t)i, step, purpose);
  }
}

// Local kinetic energy -0.5 sum_a sum_c inv_m * d2psi/psi, summed over atoms first then
// over components (imp_samp.py:49-53).
template <typename R, int NP, int MS = MS_ANY>
PVD_DEV R local_kin(const StepArgs<R>& A, const R* sec) {
  if constexpr (MS != MS_ANY && NP == 9) {
    // 3 atoms x 3 components, same summation order
    R tot = R(0);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const R s = ((R)A.inv_m[c] * sec[c] + (R)A.inv_m[3 + c] * sec[3 + c]) + (R)A.inv_m[6 + c] * sec[6 + c];
      tot = (c == 0) ? s : tot + s;
    }
    return R(-0.5) * tot;
  }
  R tot = R(0);
  for (int c = 0; c < A.ds; ++c) {
    R s = R(0);
    for (int a = 0; a < A.ms; ++a) {
      const int p = a * A.ds + c;
      R v = R(0);
#pragma unroll
      for (int q = 0; q < NP; ++q)
        if (q == p) v = (R)A.inv_m[q] * sec[q];
      s = (a == 0) ? v : s + v;
    }
    tot = (c == 0) ? s : tot + s;
  }
  return R(-0.5) * tot;
}

// Importance-sampled move of one walker (pyvibdmc.py:549-612, imp_samp.py:28-47).
// On return x holds the new position and the function returns the local energy's
// kinetic part at that position; `acc` tells whether the proposal was accepted.
// FP32 importance-sampled move on a Gaussian-product trial: gradients and the closed-form
// mass-weighted Laplacian only (trial_water_fast / trial_gauss_fast).  Same mathematics as the
// literal path below (imp_samp.py:28-53): Metropolis in log form, local kinetic energy
// -1/2 sum_p (1/m_p) (g_p^2 + d2 ln psi / dx_p^2) summed over components.
template <typename R, int NP, int MS>
PVD_DEV R is_move_fast(const StepArgs<R>& A, const R* disp, double u, R* x, bool& acc) {
  const R dt = (R)A.dt;
  R lx, ly, lapx, lapy, gx[NP], gy[NP], y[NP];
  if constexpr ((MS & MS_TRIAL_WATER) != 0 && NP == 9) {
    trial_water_fast<R>(A.trial, x, (R)A.inv_m[0], (R)A.inv_m[6], lx, gx, lapx);
  } else if constexpr ((MS & MS_TRIAL_CLUSTER) != 0 && NP == 18) {
    trial_cluster_fast<R>(A.trial, x, (R)A.inv_m[3], (R)A.inv_m[0], lx, gx, lapx);
  } else {
    trial_gauss_fast<R, NP>(A.trial, A.inv_m, x, lx, gx, lapx);
  }
#pragma unroll
  for (int p = 0; p < NP; ++p) y[p] = (x[p] + disp[p]) + ((R)A.inv_m[p] * gx[p]) * dt;
  if constexpr ((MS & MS_TRIAL_WATER) != 0 && NP == 9) {
    trial_water_fast<R>(A.trial, y, (R)A.inv_m[0], (R)A.inv_m[6], ly, gy, lapy);
  } else if constexpr ((MS & MS_TRIAL_CLUSTER) != 0 && NP == 18) {
    trial_cluster_fast<R>(A.trial, y, (R)A.inv_m[3], (R)A.inv_m[0], ly, gy, lapy);
  } else {
    trial_gauss_fast<R, NP>(A.trial, A.inv_m, y, ly, gy, lapy);
  }
  R la = R(2) * (ly - lx);
#pragma unroll
  for (int p = 0; p < NP; ++p) {
    const R a1 = (x[p] - y[p]) - ((R)A.inv_m[p] * gy[p]) * dt;
    const R a2 = (y[p] - x[p]) - ((R)A.inv_m[p] * gx[p]) * dt;
    la += (a2 * a2 - a1 * a1) * (R)A.h2s[p];
  }
  acc = (double)pvd_exp<R>(la) > u;
  R gs = R(0);
  if (acc) {
#pragma unroll
    for (int p = 0; p < NP; ++p) {
      x[p] = y[p];
      gs += (R)A.inv_m[p] * (gy[p] * gy[p]);
    }
    return R(-0.5) * (gs + lapy);
  }
#pragma unroll
  for (int p = 0; p < NP; ++p) gs += (R)A.inv_m[p] * (gx[p] * gx[p]);
  return R(-0.5) * (gs + lapx);
}

template <typename R, int NP, bool INJ, int MS = MS_ANY>
PVD_DEV R is_move(const StepArgs<R>& A, const R* disp, double u, R* x, bool& acc) {
  if constexpr (sizeof(R) == 4 && NP % 3 == 0 && NP >= 6) {
    if ((MS & MS_TRIAL_ANY_SPEC) != 0 || A.trial.kind == TRIAL_GAUSS) return is_move_fast<R, NP, MS>(A, disp, u, x, acc);
  }
  TrialOut<R, NP> tx, ty;
  trial_eval_s<R, NP, MS>(A.trial, x, tx);
  R y[NP], Dx[NP], Dy[NP];
  const R dt = (R)A.dt;
#pragma unroll
  for (int p = 0; p < NP; ++p) {
    Dx[p] = (R)A.inv_m[p] * tx.d1[p];
    y[p] = (x[p] + disp[p]) + Dx[p] * dt;
  }
  trial_eval_s<R, NP, MS>(A.trial, y, ty);
#pragma unroll
  for (int p = 0; p < NP; ++p) Dy[p] = (R)A.inv_m[p] * ty.d1[p];
  R met;
  if constexpr (sizeof(R) == 8) {
    // literal form of imp_samp.py:37-46
    const R r = ty.psi / tx.psi;
    const R ratio = r * r;
    R tc[NP];
#pragma unroll
    for (int p = 0; p < NP; ++p) {
      const R s2 = (R)A.sigma[p] * (R)A.sigma[p];
      const R a1 = (x[p] - y[p]) - Dy[p] * dt;
      const R a2 = (y[p] - x[p]) - Dx[p] * dt;
      const R t1 = pvd_exp<R>((R(-1) * (a1 * a1)) / (R(2) * s2));
      const R t2 = pvd_exp<R>((R(-1) * (a2 * a2)) / (R(2) * s2));
      tc[p] = t1 / t2;
    }
    if constexpr (MS != MS_ANY && NP == 9) {
      R prod = R(1);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const R pc = (tc[c] * tc[3 + c]) * tc[6 + c];
        prod = (c == 0) ? pc : prod * pc;
      }
      met = prod * ratio;
    } else if (A.ds == 1 && NP == 1) {
      met = tc[0] * ratio;
    } else {
      R prod = R(1);
      for (int c = 0; c < A.ds; ++c) {
        R pc = R(1);
        for (int a = 0; a < A.ms; ++a) {
          const int p = a * A.ds + c;
          R v = R(1);
#pragma unroll
          for (int q = 0; q < NP; ++q)
            if (q == p) v = tc[q];
          pc = (a == 0) ? v : pc * v;
        }
        prod = (c == 0) ? pc : prod * pc;
      }
      met = prod * ratio;
    }
    if (tx.psi * ty.psi <= R(0)) met = R(0);
  } else {
    // log form (same mathematics, no overflow in single precision)
    R la = R(2) * (ty.lnpsi - tx.lnpsi);
#pragma unroll
    for (int p = 0; p < NP; ++p) {
      const R inv2s2 = (R)A.h2s[p];
      const R a1 = (x[p] - y[p]) - Dy[p] * dt;
      const R a2 = (y[p] - x[p]) - Dx[p] * dt;
      la += (a2 * a2 - a1 * a1) * inv2s2;
    }
    met = pvd_exp<R>(la);
  }
  acc = (double)met > u;
  if (acc) {
#pragma unroll
    for (int p = 0; p < NP; ++p) x[p] = y[p];
  }
  return local_kin<R, NP, MS>(A, acc ? ty.d2 : tx.d2);
}

// Exclusive prefix of tile `tile` from published counts, by all lanes of wave 0:
//   flags[t]  = epoch | count of tile t                       (published by every tile)
//   gflags[g] = epoch | total of the LB_GROUP tiles of group g (published by its last tile)
// prefix(t) = sum of gflags over the groups before t's + sum of flags over all of t's
// predecessors in its group, read directly (4 loads per lane in flight per pass).  Every
// value is final once published, so a tile waits only for its predecessors to finish their
// own work: one coherent round trip for up to LB_GROUP tiles (1M walkers), two beyond.
// Predecessors are resident or finished (tiles are dispatched in index order), so every wait
// ends; the spin is still bounded and reports ST_SPIN if it expires.
constexpr int LB_GROUP = 1024;

PVD_DEV bool spin_epoch(const uint64_t* p, unsigned epoch, uint64_t& f) {
  long long spins = 0;
  while ((unsigned)(f >> 32) != epoch) {
    if (++spins > (1ll << 26)) return false;
    __builtin_amdgcn_s_sleep(1);
    f = ld_relaxed(p);
  }
  return true;
}

// sum of the epoch-tagged values base[0 .. cnt), by all lanes of a wave
PVD_DEV unsigned sum_published(const uint64_t* base, int cnt, unsigned epoch, bool& ok) {
  const int l = lane_id();
  unsigned v = 0;
  for (int t0 = 0; t0 < cnt; t0 += 4 * PVD_WAVE) {
    uint64_t f[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int t = t0 + q * PVD_WAVE + l;
      f[q] = t < cnt ? ld_relaxed(base + t) : ((uint64_t)epoch << 32);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int t = t0 + q * PVD_WAVE + l;
      if (t < cnt) {
        if (ok) ok = spin_epoch(base + t, epoch, f[q]);
        v += (unsigned)f[q];
      }
    }
  }
  return (unsigned)wave_sum((int)v);
}

PVD_DEV int grouped_prefix(uint64_t* flags, uint64_t* gflags, int tile, unsigned epoch, unsigned total,
                           bool& spun_out) {
  const int l = lane_id();
  const int g = tile / LB_GROUP, j = tile % LB_GROUP;
  bool ok = true;
  const unsigned ip = sum_published(flags + (size_t)g * LB_GROUP, j, epoch, ok);
  if (j == LB_GROUP - 1 && l == 0) st_relaxed(gflags + g, ((uint64_t)epoch << 32) | (uint64_t)(ip + total));
  const unsigned gp = g > 0 ? sum_published(gflags, g, epoch, ok) : 0u;
  spun_out = __any(!ok);
  return (int)(gp + ip);
}

// ----------------------------------------------------------------------------- discrete
// Walkers per thread of the discrete step: a tile is TILE * W walkers, walker
// (k, thread) of tile b is slot b*TILE*W + k*TILE + thread (coalesced per k).  Larger tiles
// mean fewer published counts and cheaper per-tile scans.
// Tiling of n walkers: the number of tiles is a multiple of the resident workgroup count
// `slots` (or n / TILE when that is smaller), and all tiles are equally large (at most TW).
// Every CU then holds the same number of equally long tiles, so no tile's output offset
// waits on a predecessor from a more crowded CU.
PVD_DEV void discrete_tiles(int n, int slots, int TW, int& ntiles, int& tw) {
  const long long per_round = (long long)(slots > 0 ? slots : 1) * TW;
  long long nt = ((n + per_round - 1) / per_round) * (long long)(slots > 0 ? slots : 1);
  const long long nmin = ((long long)n + TILE - 1) / TILE;
  if (nt > nmin) nt = nmin;
  if (nt < 1) nt = 1;
  ntiles = (int)nt;
  tw = (int)(((long long)n + nt - 1) / nt);
}

template <typename R, int NP, bool EXT, bool IS = false>
constexpr int disc_wpt() {
  return EXT ? 4 : ((int)sizeof(R) * NP <= 36 ? 4 : ((int)sizeof(R) * NP <= 72 ? 2 : 1));
}

// Finalise a discrete step inside its last tile: the records of all tiles, each read once
// its epoch word shows the current step (a tile writes that word after its other words
// have completed, and before its count), a fixed-order reduction, the population guards
// (pyvibdmc.py:397-400, 409-413), calc_vref (pyvibdmc.py:655-661) and the per-step record
// (pyvibdmc.py:674-682).  `expect` is the population from the prefix read, checked against
// the records.  Called by every thread of the tile.
template <typename R, bool IS>
PVD_DEV void finalize_discrete(const StepArgs<R>& A, Ctrl* C, int par, int n, int ntiles, unsigned epoch,
                               unsigned long long step, double dte, int expect) {
  __shared__ double f_d[TILE / PVD_WAVE];
  __shared__ int f_i[TILE / PVD_WAVE];
  // FQ records per thread and pass, their loads in flight together: one round trip for the
  // epoch words, one for the sums
  constexpr int FQ = 8;
  double s0 = 0.0;
  int cnt = 0, nb = 0, nd = 0, nbad = 0, late = 0;
  for (int t0 = 0; t0 < ntiles; t0 += TILE * FQ) {
    uint64_t w2[FQ];
#pragma unroll
    for (int q = 0; q < FQ; ++q) {
      const int t = t0 + q * TILE + (int)threadIdx.x;
      w2[q] = t < ntiles ? ld_relaxed(A.drec + 3 * (size_t)t + 2) : ((uint64_t)epoch << 32);
    }
#pragma unroll
    for (int q = 0; q < FQ; ++q) {
      const int t = t0 + q * TILE + (int)threadIdx.x;
      long long spins = 0;
      while ((unsigned)(w2[q] >> 32) != epoch) {
        if (++spins > (1ll << 26)) {
          late = 1;
          break;
        }
        __builtin_amdgcn_s_sleep(1);
        w2[q] = ld_relaxed(A.drec + 3 * (size_t)t + 2);
      }
    }
    __atomic_signal_fence(__ATOMIC_SEQ_CST);  // the sums are read after their epochs matched
    uint64_t w0[FQ], w1[FQ];
#pragma unroll
    for (int q = 0; q < FQ; ++q) {
      const int t = t0 + q * TILE + (int)threadIdx.x;
      w0[q] = t < ntiles ? ld_relaxed(A.drec + 3 * (size_t)t) : 0ull;
      w1[q] = t < ntiles ? ld_relaxed(A.drec + 3 * (size_t)t + 1) : 0ull;
    }
#pragma unroll
    for (int q = 0; q < FQ; ++q) {
      const int t = t0 + q * TILE + (int)threadIdx.x;
      if (t < ntiles) {
        s0 += __longlong_as_double((long long)w0[q]);
        cnt += (int)(uint32_t)w1[q];
        nb += (int)(uint32_t)(w1[q] >> 32);
        nd += (int)(w2[q] & 0xFFFFu);
        nbad += (int)((w2[q] >> 16) & 0xFFFFu);
      }
    }
  }
  s0 = block_sum<TILE>(s0, f_d);
  cnt = block_sum<TILE>(cnt, f_i);
  nb = block_sum<TILE>(nb, f_i);
  nd = block_sum<TILE>(nd, f_i);
  nbad = block_sum<TILE>(nbad, f_i);
  late = block_sum<TILE>(late, f_i);
  if (threadIdx.x != 0) return;
  const int np1 = par ^ 1;
  int st = ST_OK;
  if (late || cnt != expect) st = ST_SPIN;
  else if (nbad) st = ST_WEIGHT;
  else if ((double)cnt < A.pop_lo || (double)cnt > A.pop_hi) st = ST_POP;
  else if (cnt > A.cap) st = ST_CAPACITY;
  if (st != ST_OK) {
    C->status = st;
    C->status_step = (long long)step;
    return;
  }
  const double vbar = s0 / (double)cnt;
  const double corr = (double)(cnt - (int)A.n0) / A.n0;
  const double vnew = vbar - (A.alpha * corr);
  C->n[np1] = cnt;
  C->vref[np1] = vnew;
  C->step[np1] = step + 1;
  C->epoch[np1] = epoch + 1;
  if ((long long)step < A.tmax) {
    A.tr_vref[step] = vnew;
    A.tr_pop[step] = (double)cnt;
    A.tr_births[step] = nb;
    A.tr_deaths[step] = nd;
    if (IS) {
      // the effective time advances on branching steps only (update_effetive_timestep is
      // called from birth_or_death, pyvibdmc.py:384-385, 372-378); a step without branching
      // leaves its entry at zero, and the next branching step adds to that zero
      C->eff_t = A.branch ? (step == 0 ? 0.0 : C->eff_t) + dte : 0.0;
      A.tr_dteff[step] = C->eff_t;
    }
  }
  // every tile read the acceptance count before publishing: clear it for step + 2
  if (IS) {
    for (int b = 0; b < ACC_SLOTS; ++b) st_relaxed_u(acc_slot(A, par, b), 0u);
  }
}

// MODE: INJ = injected draws; IS = importance sampling (analytic trial); EXT = energies
// from V_ext (walkers already moved; NP = 0 reads the component count at run time).
// Tile b is block b.  Workgroups are dispatched in index order within each XCD, so the
// predecessors a tile waits for are resident or finished; the wait is still bounded.
template <typename R, int NP, bool INJ, bool IS, bool EXT, int MS = MS_ANY>
__global__ __launch_bounds__(TILE) void k_discrete(StepArgs<R> A) {
  constexpr int W = disc_wpt<R, NP, EXT, IS>();
  constexpr int TW = TILE * W;
  __shared__ int s_prefix, s_status;
  __shared__ int s_wt[W][TILE / PVD_WAVE];
  __shared__ double s_d[TILE / PVD_WAVE];
  __shared__ int s_i[TILE / PVD_WAVE], s_i2[TILE / PVD_WAVE], s_i3[TILE / PVD_WAVE];
  const int par = A.par;
  Ctrl* C = A.ctrl;
  const int tile = blockIdx.x;
  const int n = C->n[par];
  int ntiles, tw;
  discrete_tiles(n, A.slots, TW, ntiles, tw);
  if (C->status != ST_OK || tile >= ntiles) return;
  const unsigned long long step = C->step[par];
  const unsigned epoch = C->epoch[par];
  const double vref = C->vref[par];
  double dte = A.dt;
  if (IS) dte = A.dt * ((double)acc_total(A, par) / (double)n);

  // external energies: walkers of at most 36 bytes are loaded with their energies, before the
  // look-back, and written from registers (loading them after it put the read latency
  // between the prefix and the writes); larger ones are copied plane by plane in the write loop
  constexpr bool PREF = EXT && NP > 0 && (int)sizeof(R) * NP <= 36;
  R x[W][(!EXT || PREF) ? (NP > 0 ? NP : 1) : 1];
  int c[W];
  double sce = 0.0;
  int births = 0, deaths = 0, bad = 0;
#pragma unroll
  for (int k = 0; k < W; ++k) {
    const int li = k * TILE + threadIdx.x;
    const long long i = (long long)tile * tw + li;
    const bool valid = li < tw && i < n;
    double E = 0.0, um = 0.0, ub = 0.0;
    if (valid) {
      if constexpr (EXT) {
        E = A.V_ext[i];
        if (IS) E = E + A.ke[i];
        if constexpr (PREF) load_walker<R, NP>(A, i, x[k]);
      } else if constexpr (IS) {
        // moved, and its local energy computed, by k_is_move
        load_walker<R, NP>(A, i, x[k]);
        E = A.E[i];
      } else {
        load_walker<R, NP>(A, i, x[k]);
        R disp[NP];
        draw_step<R, NP, INJ, false>(A, i, step, disp, um, ub, A.branch != 0);
#pragma unroll
        for (int p = 0; p < NP; ++p) x[k][p] = x[k][p] + disp[p];
        E = (double)pes_eval_s<R, NP, MS>(A.pes, x[k]);
      }
    }
    // branching weight and copy count (pyvibdmc.py:392-407)
    int ck = valid ? 1 : 0;
    if (valid && A.branch) {
      const double u = (EXT || IS) ? draw_u<R, INJ>(A, i, step, RNG_BRANCH) : ub;
      if constexpr (sizeof(R) == 4 && !INJ) {
        // FP32 production path: single-precision weight (the FP64 parity path below
        // follows NumPy's double arithmetic)
        const float wf = pvd_exp<float>((float)((-1. * (E - vref)) * dte));
        if (!isfinite(wf) || (double)wf > A.pop_hi + 1.0) {
          bad = 1;
          ck = 0;
        } else {
          const float fl = floorf(wf);
          ck = (int)fl + ((float)u < (wf - fl) ? 1 : 0);
        }
      } else {
        const double w = exp((-1. * (E - vref)) * dte);
        if (!isfinite(w) || w > A.pop_hi + 1.0) {
          bad = 1;
          ck = 0;
        } else {
          const double fl = floor(w);
          ck = (int)fl + (u < (w - fl) ? 1 : 0);
        }
      }
    }
    c[k] = ck;
    sce += (double)ck * E;
    births += ck > 1 ? ck - 1 : 0;
    deaths += (valid && ck == 0) ? 1 : 0;
  }
  // exclusive prefix of every walker inside the tile, in slot order (k-major), and the
  // tile sums: wave scans and sums, one exchange through LDS, fixed summation order
  constexpr int NWV = TILE / PVD_WAVE;
  const int wv = threadIdx.x / PVD_WAVE, ln = lane_id();
  int incl[W];
#pragma unroll
  for (int k = 0; k < W; ++k) incl[k] = wave_incl_scan(c[k]);
  sce = wave_sum(sce);
  births = wave_sum(births);
  deaths = wave_sum(deaths);
  bad = wave_sum(bad);
  if (ln == PVD_WAVE - 1) {
#pragma unroll
    for (int k = 0; k < W; ++k) s_wt[k][wv] = incl[k];
  }
  if (ln == 0) {
    s_d[wv] = sce;
    s_i[wv] = births;
    s_i2[wv] = deaths;
    s_i3[wv] = bad;
  }
  __syncthreads();
  int excl[W];
  int total = 0;
#pragma unroll
  for (int k = 0; k < W; ++k) {
    int before = 0, tk = 0;
#pragma unroll
    for (int v = 0; v < NWV; ++v) {
      before += v < wv ? s_wt[k][v] : 0;
      tk += s_wt[k][v];
    }
    excl[k] = total + before + incl[k] - c[k];
    total += tk;
  }
  sce = s_d[0];
  births = s_i[0];
  deaths = s_i2[0];
  bad = s_i3[0];
#pragma unroll
  for (int v = 1; v < NWV; ++v) {
    sce += s_d[v];
    births += s_i[v];
    deaths += s_i2[v];
    bad += s_i3[v];
  }

  // publish the tile record (the epoch word last, after the others have completed), then
  // the tile's count, then read the exclusive prefix
  if (threadIdx.x == 0) {
    uint64_t* dr = A.drec + 3 * (size_t)tile;
    st_relaxed(dr + 0, (uint64_t)__double_as_longlong(sce));
    st_relaxed(dr + 1, (uint64_t)(uint32_t)total | ((uint64_t)(uint32_t)births << 32));
    store_complete();
    st_relaxed(dr + 2, ((uint64_t)epoch << 32) | ((uint64_t)(uint32_t)bad << 16) | (uint64_t)(uint32_t)deaths);
    store_complete();
    st_relaxed(&A.flags[tile], ((uint64_t)epoch << 32) | (uint64_t)(unsigned)total);
  }
  if (threadIdx.x < PVD_WAVE) {
    bool spun = false;
    const int prefix = grouped_prefix(A.flags, A.gflags, tile, epoch, (unsigned)total, spun);
    if (threadIdx.x == 0) {
      if (spun) st_release_i(&C->status, ST_SPIN);
      s_prefix = prefix;
      s_status = spun ? 1 : 0;
    }
  }
  __syncthreads();
  if (s_status) return;

  // survivors to their np.repeat positions
#pragma unroll
  for (int k = 0; k < W; ++k) {
    const int li = k * TILE + threadIdx.x;
    const long long i = (long long)tile * tw + li;
    if (li < tw && i < n && c[k] > 0) {
      const long long base = (long long)s_prefix + excl[k];
      for (int q = 0; q < c[k]; ++q) {
        const long long dst = base + q;
        if (dst >= A.cap) break;
        if constexpr (EXT && NP == 0) {
          for (int p = 0; p < A.np; ++p) A.cout[p * A.cap + dst] = A.cin[p * A.cap + i];
        } else if constexpr (EXT && !PREF) {
          // external energies: the walker is copied from the input planes
          if constexpr (sizeof(R) == 4) {
            const __amdgpu_buffer_rsrc_t ri = coord_rsrc((const float*)A.cin, A.cap, NP);
            const __amdgpu_buffer_rsrc_t ro = coord_rsrc((const float*)A.cout, A.cap, NP);
#pragma unroll
            for (int p = 0; p < NP; ++p) coord_st(ro, A.cap, p, dst, coord_ld(ri, A.cap, p, i));
          } else {
#pragma unroll
            for (int p = 0; p < NP; ++p) A.cout[p * A.cap + dst] = A.cin[p * A.cap + i];
          }
        } else if constexpr (sizeof(R) == 4) {
          const __amdgpu_buffer_rsrc_t r = coord_rsrc((const float*)A.cout, A.cap, NP);
#pragma unroll
          for (int p = 0; p < NP; ++p) coord_st(r, A.cap, p, dst, x[k][p]);
        } else {
#pragma unroll
          for (int p = 0; p < NP; ++p) A.cout[p * A.cap + dst] = x[k][p];
        }
        if (A.who_out) A.who_out[dst] = A.who_in[i];
        if constexpr (EXT) {
          // the trial caches of a caller-evaluated wave function ride along
          for (int q2 = 0; q2 < A.nx; ++q2) A.xc_out[(long long)q2 * A.cap + dst] = A.xc_in[(long long)q2 * A.cap + i];
        }
      }
    }
  }
  // the last tile has its prefix only after every other tile has published its count, and a
  // tile publishes its record before its count
  if (tile == ntiles - 1) finalize_discrete<R, IS>(A, C, par, n, ntiles, epoch, step, dte, s_prefix + total);
}

// ------------------------------------------------------------ importance-sampled move
// The move of an importance-sampled step (pyvibdmc.py:549-612): drift, proposal,
// Metropolis test, then the potential and local kinetic energy at the final position
// (pyvibdmc.py:778-786).  The walker is updated in place and its local energy stored; the
// accepted proposals are counted for dt_eff, which the weights of the same step need
// (pyvibdmc.py:387-389, 603), so branching runs as the next kernel.
template <typename R, int NP, bool INJ, int MS = MS_ANY>
__global__ __launch_bounds__(TILE) void k_is_move(StepArgs<R> A) {
  __shared__ int s_i[TILE / PVD_WAVE];
  const int par = A.par;
  Ctrl* C = A.ctrl;
  const int n = C->n[par];
  const int ntiles = (n + TILE - 1) / TILE;
  const int tile = blockIdx.x;
  if (C->status != ST_OK || tile >= ntiles) return;
  const unsigned long long step = C->step[par];
  const long long i = (long long)tile * TILE + threadIdx.x;
  int a = 0;
  if (i < n) {
    R x[NP], disp[NP];
    double um, ub;
    load_walker<R, NP>(A, i, x);
    draw_step<R, NP, INJ, true>(A, i, step, disp, um, ub, false);
    bool acc;
    const R ke = is_move<R, NP, INJ, MS>(A, disp, um, x, acc);
    a = acc ? 1 : 0;
    const R v = pes_eval_s<R, NP, MS>(A.pes, x);
    A.E[i] = (double)(v + ke);
    if (acc) store_walker_in<R, NP>(A, i, x);
  }
  a = block_sum<TILE>(a, s_i);
  if (threadIdx.x == 0 && a) fetch_add_relaxed(acc_slot(A, par, tile), (unsigned)a);
}

// Displacements of components p0 .. p0+3 of walker i on the split path (run-time component
// count): injected (the reference's (n, M, d) draws) or Philox blocks tagged 0x80 / 0xC0.
template <typename R, bool INJ>
PVD_DEV void split_disp(const StepArgs<R>& A, long long i, unsigned long long step, int p0, R* z) {
  if constexpr (INJ) {
    for (int q = 0; q < 4; ++q) z[q] = p0 + q < A.np ? (R)A.inj_disp[i * A.np + p0 + q] : R(0);
  } else {
    const Philox4 r = A.rng.block((uint32_t)i, step, RNG_DISP, (uint32_t)(p0 / 4) | 0x80u);
    if constexpr (sizeof(R) == 4) {
      box_muller_f(r.v[0], r.v[1], z[0], z[1]);
      box_muller_f(r.v[2], r.v[3], z[2], z[3]);
    } else {
      double a0, a1;
      box_muller_d(r.v[0], r.v[1], r.v[2], r.v[3], a0, a1);
      z[0] = a0;
      z[1] = a1;
      const Philox4 r2 = A.rng.block((uint32_t)i, step, RNG_DISP, (uint32_t)(p0 / 4) | 0xC0u);
      box_muller_d(r2.v[0], r2.v[1], r2.v[2], r2.v[3], a0, a1);
      z[2] = a0;
      z[3] = a1;
    }
    for (int q = 0; q < 4; ++q) z[q] = (R)A.sigma[(p0 + q) < MAXNP ? p0 + q : 0] * z[q];
  }
}

// ------------------------------------------------------------------ split path: the move
// Moves walkers in place (plain or importance-sampled) so that an outside PES can be
// evaluated on the moved walkers.  IS writes the local kinetic energy per walker and
// counts acceptances.
template <typename R, int NP, bool INJ, bool IS, int MS = MS_ANY>
__global__ __launch_bounds__(TILE) void k_move(StepArgs<R> A) {
  __shared__ int s_i[TILE / PVD_WAVE];
  const int par = A.par;
  Ctrl* C = A.ctrl;
  const int n = C->n[par];
  const int ntiles = (n + TILE - 1) / TILE;
  const int tile = blockIdx.x;
  if (C->status != ST_OK || tile >= ntiles) return;
  const unsigned long long step = C->step[par];
  const long long i = (long long)tile * TILE + threadIdx.x;
  int a = 0;
  if (i < n) {
    if constexpr (IS) {
      R x[NP], disp[NP];
      double um, ub;
      load_walker<R, NP>(A, i, x);
      draw_step<R, NP, INJ, true>(A, i, step, disp, um, ub, false);
      bool acc;
      const R ke = is_move<R, NP, INJ, MS>(A, disp, um, x, acc);
      a = acc ? 1 : 0;
      A.ke[i] = (double)ke;
#pragma unroll
      for (int p = 0; p < NP; ++p) A.cout[p * A.cap + i] = x[p];
    } else {
      // runtime component count; draws in blocks of four components
      for (int p0 = 0; p0 < A.np; p0 += 4) {
        R z[4];
        split_disp<R, INJ>(A, i, step, p0, z);
        for (int q = 0; q < 4 && p0 + q < A.np; ++q) {
          const long long o = (long long)(p0 + q) * A.cap + i;
          A.cout[o] = A.cin[o] + z[q];
        }
      }
    }
  }
  if constexpr (IS) {
    a = block_sum<TILE>(a, s_i);
    if (threadIdx.x == 0 && a) fetch_add_relaxed(acc_slot(A, par, tile), (unsigned)a);
  }
}

// --------------------------------------------------------------- caller's trial function
// Importance sampling with a trial wave function evaluated outside the library: the
// reference's ImpSampManager plugins (call_trial / call_derivs, imp_samp_manager.py:92-139,
// 197-224) called by imp_move_randomly (pyvibdmc.py:549-612).  The library keeps that
// move's caches next to the walkers as FP64 planes xc[q * cap + i]:
//   q in [0, np)      dpsi/psi   (f_x)
//   q in [np, 2 np)   d2psi/psi  (psi_sec_der)
//   q = 2 np          psi        (psi_1)
// and they follow the walkers through branching (pyvibdmc.py:424-427, 353-356).  A step is:
// k_his_propose (y = x + disp + D_x dt, written for the caller), the caller's psi / derivatives
// at y, k_his_accept (Metropolis, caches, local kinetic energy), a potential, branching.
template <typename R>
PVD_DEV double his_local_kin(const StepArgs<R>& A, const double* sd, long long stride, long long i) {
  // -0.5 sum_c sum_a inv_m sd, atoms summed first (imp_samp.py:49-53)
  double tot = 0.0;
  for (int c = 0; c < A.ds; ++c) {
    double s = 0.0;
    for (int a = 0; a < A.ms; ++a) {
      const int p = a * A.ds + c;
      const double v = A.inv_m[p] * sd[p * stride + i];
      s = (a == 0) ? v : s + v;
    }
    tot = (c == 0) ? s : tot + s;
  }
  return -0.5 * tot;
}

// The proposal of pyvibdmc.py:556-564: y = (x + disp) + (D_x dt), D_x = inv_m * f_x, as
// (n, M, d) float64 rows for the caller.
template <typename R, bool INJ>
__global__ __launch_bounds__(TILE) void k_his_propose(StepArgs<R> A, double* y) {
  const int par = A.par;
  const Ctrl* C = A.ctrl;
  const int n = C->n[par];
  const long long i = (long long)blockIdx.x * TILE + threadIdx.x;
  if (C->status != ST_OK || i >= n) return;
  const unsigned long long step = C->step[par];
  const int np = A.np;
  for (int p0 = 0; p0 < np; p0 += 4) {
    R z[4];
    split_disp<R, INJ>(A, i, step, p0, z);
    for (int q = 0; q < 4 && p0 + q < np; ++q) {
      const int p = p0 + q;
      const double dx = A.inv_m[p] * A.xc_in[(long long)p * A.cap + i];
      y[i * np + p] = ((double)A.cin[(long long)p * A.cap + i] + (double)z[q]) + dx * A.dt;
    }
  }
}

// Metropolis test of pyvibdmc.py:566-605 in the literal form of imp_samp.py:28-47 (FP64:
// the ratio of psi squared times, per component, exp(-(x - y - D_y dt)^2 / 2 sigma^2) over
// exp(-(y - x - D_x dt)^2 / 2 sigma^2), multiplied over atoms and then over dimensions; zero
// when psi changes sign), then the accepted walkers take y and its caches
// (pyvibdmc.py:603-609) and every walker stores the local kinetic energy of its final
// position for the branching kernel (pyvibdmc.py:807-809).  y is the caller's array after
// its trial call (which may have perturbed it in place, as finite differences do).
template <typename R, bool INJ>
__global__ __launch_bounds__(TILE) void k_his_accept(StepArgs<R> A, const double* y, const double* psi_y,
                                                     const double* f_y, const double* sd_y) {
  __shared__ int s_i[TILE / PVD_WAVE];
  const int par = A.par;
  const Ctrl* C = A.ctrl;
  const int n = C->n[par];
  const int ntiles = (n + TILE - 1) / TILE;
  const int tile = blockIdx.x;
  if (C->status != ST_OK || tile >= ntiles) return;
  const unsigned long long step = C->step[par];
  const long long i = (long long)tile * TILE + threadIdx.x;
  const int np = A.np;
  const long long cap = A.cap;
  int a = 0;
  if (i < n) {
    const double u = INJ ? A.inj_umet[i] : uniform_d(A.rng, (uint32_t)i, step, RNG_METRO);
    const double psx = A.xc_in[2ll * np * cap + i], psy = psi_y[i];
    const double r = psy / psx;
    const double ratio = r * r;
    double prod = 1.0;
    for (int c = 0; c < A.ds; ++c) {
      double pc = 1.0;
      for (int at = 0; at < A.ms; ++at) {
        const int p = at * A.ds + c;
        const double s2 = A.sigma[p] * A.sigma[p];
        const double xp = (double)A.cin[(long long)p * cap + i], yp = y[i * np + p];
        const double a1 = (xp - yp) - (A.inv_m[p] * f_y[i * np + p]) * A.dt;
        const double a2 = (yp - xp) - (A.inv_m[p] * A.xc_in[(long long)p * cap + i]) * A.dt;
        const double t1 = exp((-1.0 * (a1 * a1)) / (2.0 * s2));
        const double t2 = exp((-1.0 * (a2 * a2)) / (2.0 * s2));
        const double tc = t1 / t2;
        pc = (at == 0) ? tc : pc * tc;
      }
      prod = (c == 0) ? pc : prod * pc;
    }
    double met = prod * ratio;
    if (psx * psy <= 0.0) met = 0.0;
    const bool acc = met > u;
    a = acc ? 1 : 0;
    R* x = const_cast<R*>(A.cin);
    double* xc = const_cast<double*>(A.xc_in);
    if (acc) {
      for (int p = 0; p < np; ++p) {
        x[(long long)p * cap + i] = (R)y[i * np + p];
        xc[(long long)p * cap + i] = f_y[i * np + p];
        xc[(long long)(np + p) * cap + i] = sd_y[i * np + p];
      }
      xc[2ll * np * cap + i] = psy;
    }
    A.ke[i] = his_local_kin(A, xc + (long long)np * cap, cap, i);
  }
  a = block_sum<TILE>(a, s_i);
  if (threadIdx.x == 0 && a) fetch_add_relaxed(acc_slot(A, par, tile), (unsigned)a);
}

// Caches of the current walkers from the caller's trial call (the first step's
// self.impsamp.drift(self._walker_coords), pyvibdmc.py:552-553, 763-767): rows (n, M, d)
// float64; x (optional) replaces the coordinates (a finite-difference call perturbs them in
// place), and the local kinetic energy of sd is stored for the first-step energies.
template <typename R>
__global__ __launch_bounds__(TILE) void k_his_cache(StepArgs<R> A, const double* x, const double* psi,
                                                    const double* f, const double* sd) {
  const int n = A.ctrl->n[A.par];
  const long long i = (long long)blockIdx.x * TILE + threadIdx.x;
  if (i >= n) return;
  const int np = A.np;
  const long long cap = A.cap;
  R* xo = const_cast<R*>(A.cin);
  double* xc = const_cast<double*>(A.xc_in);
  for (int p = 0; p < np; ++p) {
    if (x) xo[(long long)p * cap + i] = (R)x[i * np + p];
    xc[(long long)p * cap + i] = f[i * np + p];
    xc[(long long)(np + p) * cap + i] = sd[i * np + p];
  }
  xc[2ll * np * cap + i] = psi[i];
  A.ke[i] = his_local_kin(A, xc + (long long)np * cap, cap, i);
}

// The simulation's device potential on its current walkers (potential.getpot on
// self._walker_coords, pyvibdmc.py:786-793), float64 energies.
template <typename R, int NP>
__global__ __launch_bounds__(TILE) void k_pes_planes(StepArgs<R> A, double* v) {
  const int n = A.ctrl->n[A.par];
  const long long i = (long long)blockIdx.x * TILE + threadIdx.x;
  if (i >= n) return;
  R x[NP];
  load_walker<R, NP>(A, i, x);
  v[i] = (double)pes_eval<R, NP>(A.pes, x);
}

// ------------------------------------------------------------------------- continuous
// Weights update in place (pyvibdmc.py:433, 833-837); walkers below the lower threshold
// are flagged in per-wave masks for the split replay (pyvibdmc.py:436-437, 340-356).
template <typename R, int NP, bool INJ, bool IS, bool EXT, int MS = MS_ANY>
__global__ __launch_bounds__(TILE) void k_cont(StepArgs<R> A) {
  const int par = A.par;
  Ctrl* C = A.ctrl;
  const int n = C->n[par];
  const int ntiles = (n + TILE - 1) / TILE;
  const int tile = blockIdx.x;
  if (C->status != ST_OK || tile >= ntiles) return;
  const unsigned long long step = C->step[par];
  const double vref = C->vref[par];
  double dte = A.dt;
  if (IS) dte = A.dt * ((double)acc_total(A, par) / (double)n);
  const long long i = (long long)tile * TILE + threadIdx.x;
  const bool valid = i < n;
  double E = 0.0, w = 0.0;
  bool kill = false;
  if (valid) {
    if constexpr (EXT) {
      E = A.V_ext[i];
      if (IS) E = E + A.ke[i];
    } else if constexpr (IS) {
      // moved in place, and its local energy computed, by k_is_move (continuous weighting
      // keeps one walker buffer, so the coordinates need no copy)
      E = A.E[i];
    } else {
      R x[NP], disp[NP];
      double um, ub;
      load_walker<R, NP>(A, i, x);
      draw_step<R, NP, INJ, false>(A, i, step, disp, um, ub, false);
#pragma unroll
      for (int p = 0; p < NP; ++p) x[p] = x[p] + disp[p];
#pragma unroll
      for (int p = 0; p < NP; ++p) A.cout[p * A.cap + i] = x[p];
      E = (double)pes_eval_s<R, NP, MS>(A.pes, x);
    }
    w = A.w[i] * exp((-1.0 * (E - vref)) * dte);
    A.w[i] = w;
    A.E[i] = E;
    kill = A.branch && (w < A.thresh_lo);
  }
  const unsigned long long km = __ballot(kill);
  // candidates of the split replay, appended in any order (the replay sorts them): list A
  // from cand_t up, list B from cand_t2 to cand_t
  const double ct = C->cand_t, ct2 = C->cand_t2;
  const bool cnd = A.branch && ct > 0.0 && valid && w >= ct;
  const bool cnd2 = A.branch && ct2 > 0.0 && valid && !cnd && w >= ct2;
  const unsigned long long cm = __ballot(cnd), cm2 = __ballot(cnd2);
  // the tile's sums in one pass: wave partials, combined by one thread in a fixed order
  constexpr int NW = TILE / PVD_WAVE;
  __shared__ double s_r[5][NW];
  __shared__ int s_k[NW], s_c[NW + 1], s_c2[NW + 1];
  const int wv = threadIdx.x / PVD_WAVE, l = lane_id();
  const double v0 = wave_sum(w * E), v1 = wave_sum(w);
  const double v2 = wave_sum(kill ? 0.0 : w * E), v3 = wave_sum(kill ? 0.0 : w);
  const double v4 = wave_max(valid ? w : 0.0);
  if (l == 0) {
    A.kmask[(long long)tile * NW + wv] = km;
    s_r[0][wv] = v0;
    s_r[1][wv] = v1;
    s_r[2][wv] = v2;
    s_r[3][wv] = v3;
    s_r[4][wv] = v4;
    s_k[wv] = __popcll(km);
    s_c[wv] = __popcll(cm);
    s_c2[wv] = __popcll(cm2);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    TileRec r;
    r.s0 = s_r[0][0];
    r.s1 = s_r[1][0];
    r.s2 = s_r[2][0];
    r.s3 = s_r[3][0];
    r.mx = s_r[4][0];
    r.count = s_k[0];
    int nc = s_c[0], nc2 = s_c2[0];
    for (int q = 1; q < NW; ++q) {
      r.s0 += s_r[0][q];
      r.s1 += s_r[1][q];
      r.s2 += s_r[2][q];
      r.s3 += s_r[3][q];
      r.mx = s_r[4][q] > r.mx ? s_r[4][q] : r.mx;
      r.count += s_k[q];
      nc += s_c[q];
      nc2 += s_c2[q];
    }
    r.births = 0;
    r.deaths = 0;
    r.bad = 0;
    A.rec[tile] = r;
    const int sub = tile & (LIST_SUB - 1);
    s_c[NW] = nc > 0 ? (int)min(atomicAdd(list_count(A, 0, sub), (unsigned)nc), (unsigned)SUB_CAP) : 0;
    s_c2[NW] = nc2 > 0 ? (int)min(atomicAdd(list_count(A, 1, sub), (unsigned)nc2), (unsigned)SUB_CAP) : 0;
  }
  __syncthreads();
  if (cnd || cnd2) {
    const unsigned long long bm = cnd ? cm : cm2;
    const int* sc = cnd ? s_c : s_c2;
    int o = sc[NW] + __popcll(bm & ((1ull << l) - 1ull));
    for (int q = 0; q < wv; ++q) o += sc[q];
    if (o < SUB_CAP) {
      Cand c;
      c.w = w;
      c.slot = (int)i;
      c.pad = 0;
      A.cand[((cnd ? 0 : LIST_SUB) + (tile & (LIST_SUB - 1))) * SUB_CAP + o] = c;
    }
  }
}

// Exact replay of the sequential split loop (pyvibdmc.py:344-356) for all walkers below
// the threshold, in one workgroup (k_cont_replay).  It first tries the closed form from
// the candidate lists (replay_candidates, below); the interval replay described here is
// the general path for what that leaves (the first step, ties, lists that miss).
//
// The loop repeatedly takes the walker of largest weight (first index on ties), halves
// it and gives the other half to the next walker on the kill list.  The chosen weights
// never increase, and halving moves a weight from [L, H) to below H/2 <= L.  So for any
// interval [L, H) with H <= 2L that lies above everything still to be chosen, the
// entries inside it are chosen in order of (weight desc, index asc) before any piece
// they create.  The replay walks such intervals downwards, sorting each one in LDS.
//
// Above the threshold no killed walker can be a candidate, so an interval's kills are
// assigned to its sorted members in parallel.  Once the intervals reach below the
// threshold, killed walkers that have not had their turn yet are candidates too, and a
// piece that sits in a killed slot is overwritten when that slot's turn comes; there the
// interval is replayed entry by entry by one thread with that validity rule.  All
// copies read their sources before any write, and the corrections to sum(w) and
// sum(w*E) use the pre-split values.
constexpr int RP_NT = 1024;
constexpr int RP_CAP = 2048;   // interval members held in LDS
constexpr int RP_KCAP = KILL_CAP;  // kills per step
constexpr int RP_PCAP = RP_KCAP * 2;  // pending pieces per buffer
constexpr int RP_FCAP = RP_KCAP * 4;  // final pieces

// total order (weight descending, then slot, birth, source): the sorted interval does not
// depend on the order in which its members were gathered
PVD_DEV bool piece_before(const Piece& a, const Piece& b) {
  if (a.v != b.v) return a.v > b.v;
  if (a.slot != b.slot) return a.slot < b.slot;
  if (a.birth != b.birth) return a.birth < b.birth;
  return a.src < b.src;
}

// kill index of `slot` in this launch (tagged with the launch epoch), or INT_MAX
PVD_DEV int kill_pos(const unsigned long long* tag, int slot, unsigned long long epoch) {
  const unsigned long long t = tag[slot];
  return (t >> 32) == epoch ? (int)(t & 0xFFFFFFFFull) : 0x7FFFFFFF;
}

// a piece is live at kill index q unless its slot's own turn (earlier than q) overwrote it
PVD_DEV bool piece_live(const Piece& e, int q, const unsigned long long* tag, unsigned long long epoch) {
  const int kp = kill_pos(tag, e.slot, epoch);
  return !(kp < q && e.birth < kp);
}

// Outcome of the candidate replay; the interval replay runs when it is not RC_DONE.
enum { RC_DONE = 0, RC_FEW = 1, RC_MANY = 2, RC_OTHER = 3 };
// scratch: list keys and order (8 + 2 B), bucketed keys and order (8 + 2 B), bucket
// counts (4 B per bucket, one per candidate), split depth (1 B)
constexpr int RC_BUCKETS = CAND_CAP;
constexpr int RC_LDS = 25 * CAND_CAP;
constexpr int RP_LDS = RP_CAP * (int)sizeof(Piece) > RC_LDS ? RP_CAP * (int)sizeof(Piece) : RC_LDS;

PVD_DEV int dexp(double w) { return (int)((__double_as_longlong(w) >> 52) & 0x7FF) - 1023; }

// Histogram of weights for the candidate thresholds: 256 bins per octave from the lower
// threshold up (the top 8 mantissa bits), 16 octaves, the last bin open-ended.
constexpr int TH_SHIFT = 44;
constexpr int TH_BINS = 4096;
constexpr int TH_BPT = TH_BINS / RP_NT;  // bins per replay thread
PVD_DEV int th_bin(double v, long long base) {
  const long long b = (long long)((unsigned long long)__double_as_longlong(v) >> TH_SHIFT) - base;
  return b < 0 ? -1 : (int)min(b, (long long)TH_BINS - 1);
}
PVD_DEV long long th_base(double thresh) {
  return (long long)((unsigned long long)__double_as_longlong(thresh) >> TH_SHIFT);
}
PVD_DEV double th_edge(double thresh, int b) {
  return fmax(thresh, __longlong_as_double((th_base(thresh) + b) << TH_SHIFT));
}

// Scan of the histogram from the top bin down: the first bin (counting down) at which the
// running count passes `lim` (mode 0: reaches lim, mode 1: exceeds lim), or -1.  All
// threads call it; thread t holds bins TH_BINS-1-TH_BPT*t .. TH_BINS-TH_BPT*(t+1).
PVD_DEV int th_cross(const int* hist, int lim, bool exceed, int* s_i, int* s_sh) {
  const int t = threadIdx.x;
  int c[TH_BPT], loc = 0;
#pragma unroll
  for (int u = 0; u < TH_BPT; ++u) {
    c[u] = hist[TH_BINS - 1 - (TH_BPT * t + u)];
    loc += c[u];
  }
  if (t == 0) s_sh[2] = -1;
  int tot;
  int run = block_excl_scan<RP_NT>(loc, s_i, tot);
#pragma unroll
  for (int u = 0; u < TH_BPT; ++u) {
    const bool hit = exceed ? (run <= lim && run + c[u] > lim) : (run < lim && run + c[u] >= lim);
    if (hit) s_sh[2] = TH_BINS - 1 - (TH_BPT * t + u);
    run += c[u];
  }
  __syncthreads();
  const int b = s_sh[2];
  __syncthreads();
  return b;
}

// Lowest bin edge with at least `target` counts at or above it (a weight >= thresh), or 0
// when the histogram holds fewer.
PVD_DEV double th_pick(const int* hist, int target, double thresh, int* s_i, int* s_sh) {
  const int b = th_cross(hist, target, false, s_i, s_sh);
  return b < 0 ? 0.0 : th_edge(thresh, b);
}

// List A size the threshold aims at: half as many again as the kills (at least 1024), so
// that the replay needs no cut inside the list (the LDS replay cuts lists longer than it
// sorts; with more than CAND_CAP / 1.25 kills the global-memory replay takes them).
PVD_DEV int th_target(int K) { return min(max(K + K / 2, 1024), CAND_G - CAND_G / 8); }

// Closed-form replay of the sequential split loop (pyvibdmc.py:344-356) from the
// candidate list: the walkers of weight >= T0, where T0 >= the lower threshold.
//
// Each turn halves the largest entry (first slot on ties) and gives a copy with the other
// half to the next killed slot.  An original walker o of weight w_o = m_o 2^e_o (m_o in
// [1, 2)) that has been split d levels deep has therefore become 2^d entries of weight
// w_o 2^-d, all copies of o.  The octave [2^E, 2^(E+1)) holds, for each o with e_o >= E,
// its 2^(e_o-E) entries of the one value m_o 2^E: the loop takes the octaves from the top
// down, the families of an octave in decreasing m_o, and the entries of one family in
// slot order.  So a scan of the family sizes over the candidates sorted by mantissa gives,
// octave by octave, the kill indices each family takes, until the K kills are used.  Every
// kill slot then holds its family's walker with weight w_o 2^-D_o (D_o: the family's split
// depth), and only the family p where the kills run out (r of its entries, in slot order)
// needs its slots ranked.
//
// Splitting moves weight without changing any family's total, so the step's sums over
// the walkers not killed are already those after the split (vref needs nothing from here).
//
// The replay uses the walkers at or above a cut T >= T0 (T0 itself when the list fits the
// sort).  The entries of value >= T must number at least K: the chosen ones are then above
// every other walker, and above every killed walker.  Equal mantissas of two families that
// meet in an octave at or above the last one used, a list overflow or a split deeper than
// the scratch holds make it return without touching the walkers.
template <typename R>
PVD_DEV int replay_candidates(const StepArgs<R>& A, int K, int nlists, const int* subn, double T0, unsigned char* buf,
                              int* s_i, int* s_sh, int* hist, double& cut, bool& refined) {
  constexpr int NT = RP_NT, JPT = CAND_CAP / RP_NT;
  constexpr unsigned long long MANT = (1ull << 52) - 1;
  constexpr int NS = 2 * LIST_SUB;
  const int tid = threadIdx.x;
  // the sub-lists used (list A's, then list B's), read as one list of nl entries
  int pre[NS + 1];
  pre[0] = 0;
#pragma unroll
  for (int c = 0; c < NS; ++c) {
    const int k = c < nlists * LIST_SUB ? subn[c] : 0;
    if (k > SUB_CAP16) return RC_MANY;  // entries were lost, or beyond the 16-bit positions
    pre[c + 1] = pre[c] + k;
  }
  const int nl = pre[NS];
  if (nl <= 0) return RC_FEW;
  if (K > KILL_CAP) return RC_OTHER;
  if (!(T0 >= A.thresh_lo) || !(T0 > 0.0)) return RC_OTHER;
  // position of the q-th entry as (sub-list << 12) | offset (constant indices only: the
  // sub-list bounds stay in registers; an empty sub-list is passed over), and the entry
  auto pos16 = [&](int q) -> int {
    int u0 = 0, b = 0;
#pragma unroll
    for (int u = 0; u < NS; ++u)
      if (q >= pre[u]) {
        u0 = u;
        b = pre[u];
      }
    return (u0 << 12) | (q - b);
  };
  auto cand16 = [&](int x) -> const Cand& { return A.cand[(x >> 12) * SUB_CAP + (x & 4095)]; };
  const long long tb = th_base(A.thresh_lo);
  unsigned long long* key = reinterpret_cast<unsigned long long*>(buf);
  unsigned short* ord = reinterpret_cast<unsigned short*>(buf + 8 * CAND_CAP);
  unsigned long long* bkey = reinterpret_cast<unsigned long long*>(buf + 10 * CAND_CAP);
  unsigned short* bord = reinterpret_cast<unsigned short*>(buf + 18 * CAND_CAP);
  int* bcnt = reinterpret_cast<int*>(buf + 20 * CAND_CAP);
  unsigned char* dep = buf + 24 * CAND_CAP;
  int* start = reinterpret_cast<int*>(buf);  // over the keys once they are sorted
  // 1. the cut: T0 when the list fits the sort, else the lowest histogram edge with at most
  //    about 1.25 K entries at or above it
  double T = T0;
  refined = nl > CAND_CAP;
  if (refined) {
    for (int b = tid; b < TH_BINS; b += NT) hist[b] = 0;
    __syncthreads();
    for (int q = tid; q < nl; q += NT) atomicAdd(&hist[th_bin(cand16(pos16(q)).w, tb)], 1);
    __syncthreads();
    // about 1.25 K walkers (at least 1024): they alone hold K entries
    const int lim = min(CAND_CAP, max(K + K / 4, 1024));
    const int over = th_cross(hist, lim, true, s_i, s_sh);  // the bin that brings too many
    if (over < 0 || over >= TH_BINS - 1) return RC_MANY;
    T = fmax(T0, th_edge(A.thresh_lo, over + 1));
  }
  // the walkers >= T, in any order (wave ballots onto one counter)
  if (tid == 0) s_sh[3] = 0;
  __syncthreads();
  for (int q0 = 0; q0 < nl; q0 += NT) {
    const int q = q0 + tid;
    const int pq = pos16(q);
    double w = 0.0;
    bool in = false;
    if (q < nl) {
      w = cand16(pq).w;
      in = w >= T;
    }
    const unsigned long long bal = __ballot(in);
    if (bal) {
      const int lead = __ffsll((long long)bal) - 1;
      int base = 0;
      if (lane_id() == lead) base = atomicAdd(&s_sh[3], __popcll(bal));
      base = __shfl(base, lead, PVD_WAVE);
      const int o = base + __popcll(bal & ((1ull << lane_id()) - 1ull));
      if (in && o < CAND_CAP) {
        // decreasing mantissa, then decreasing exponent: in a run of equal mantissas the
        // adjacent pair of largest minimum exponent is the run's first pair
        const unsigned long long wb = (unsigned long long)__double_as_longlong(w);
        key[o] = ((MANT - (wb & MANT)) << 11) | (0x7FFull - ((wb >> 52) & 0x7FFull));
        ord[o] = (unsigned short)pq;
      }
    }
  }
  __syncthreads();
  const int C = s_sh[3];
  if (C > CAND_CAP) return RC_MANY;
  if (C == 0) return RC_FEW;
  // 2. candidates by decreasing mantissa (then exponent): counting sort on the top 12 bits
  //    of the 63-bit key (buckets in LDS), then each entry ranked inside its bucket; a
  //    bitonic sort when one bucket holds too many
  constexpr int KPT = CAND_CAP / NT;
  constexpr int BSH = 63 - 12;
  for (int b = tid; b < RC_BUCKETS; b += NT) bcnt[b] = 0;
  __syncthreads();
  int bpos[KPT];
#pragma unroll
  for (int k = 0; k < KPT; ++k) {
    const int q = tid + k * NT;
    bpos[k] = q < C ? atomicAdd(&bcnt[(int)(key[q] >> BSH)], 1) : 0;
  }
  __syncthreads();
  int bmax;
  {
    int c4[4], loc = 0, mx = 0;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      c4[u] = bcnt[tid * 4 + u];
      loc += c4[u];
      mx = max(mx, c4[u]);
    }
    int tot;
    const int ex = block_excl_scan<NT>(loc, s_i, tot);
    int st = ex;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      bcnt[tid * 4 + u] = st;  // bucket starts (block_excl_scan's barriers: counts all read)
      st += c4[u];
    }
    bmax = block_max<NT>(mx, s_i);
  }
  if (bmax <= 64) {
#pragma unroll
    for (int k = 0; k < KPT; ++k) {
      const int q = tid + k * NT;
      if (q < C) {
        const int d = bcnt[(int)(key[q] >> BSH)] + bpos[k];
        bkey[d] = key[q];
        bord[d] = ord[q];
      }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < KPT; ++k) {
      const int q = tid + k * NT;
      if (q < C) {
        const unsigned long long kq = bkey[q];
        const int bk = (int)(kq >> BSH);
        const int bs = bcnt[bk], be = bk + 1 < RC_BUCKETS ? bcnt[bk + 1] : C;
        int rank = 0;
        for (int i = bs; i < be; ++i) {
          const unsigned long long ki = bkey[i];
          rank += (ki < kq || (ki == kq && i < q)) ? 1 : 0;
        }
        key[bs + rank] = kq;
        ord[bs + rank] = bord[q];
      }
    }
    __syncthreads();
  } else {
    int P = 64;
    while (P < C) P <<= 1;
    for (int q = C + tid; q < P; q += NT) {
      key[q] = ~0ull;
      ord[q] = 0;
    }
    __syncthreads();
    for (int kk = 2; kk <= P; kk <<= 1) {
      for (int jj = kk >> 1; jj > 0; jj >>= 1) {
        for (int pi = tid; pi < P / 2; pi += NT) {
          const int q = ((pi & ~(jj - 1)) << 1) | (pi & (jj - 1));
          const int r = q | jj;
          const unsigned long long a = key[q], b = key[r];
          if (((q & kk) == 0) ? (a > b) : (a < b)) {
            key[q] = b;
            key[r] = a;
            const unsigned short t = ord[q];
            ord[q] = ord[r];
            ord[r] = t;
          }
        }
        __syncthreads();
      }
    }
  }
  // 3. this thread's families: sorted positions tid*JPT .. tid*JPT+JPT-1
  double wj[JPT];
  int oj[JPT], ej[JPT];
  unsigned long long mj[JPT];
  int etop = -0x40000000, tie_e = -0x40000000;
#pragma unroll
  for (int s = 0; s < JPT; ++s) {
    const int j = tid * JPT + s;
    wj[s] = 0.0;
    oj[s] = -1;
    ej[s] = -0x40000000;
    mj[s] = 0;
    if (j < C) {
      const Cand c = cand16(ord[j]);
      wj[s] = c.w;
      oj[s] = c.slot;
      ej[s] = dexp(c.w);
      mj[s] = (unsigned long long)__double_as_longlong(c.w) & MANT;
      etop = max(etop, ej[s]);
      // equal mantissas meet in every octave at or below the smaller exponent
      if (j + 1 < C && (key[j + 1] >> 11) == (key[j] >> 11))
        tie_e = max(tie_e, min(ej[s], dexp(cand16(ord[j + 1]).w)));
    }
  }
  etop = block_max<NT>(etop, s_i);
  tie_e = block_max<NT>(tie_e, s_i);  // barriers: the keys are no longer read
  const int eT = dexp(T);
  const unsigned long long mT = (unsigned long long)__double_as_longlong(T) & MANT;
  if (etop - eT > 40) return RC_OTHER;
  // 4. octaves from the top: kill indices of each family, until K are used; in the octave
  //    of T only the entries of value >= T count
  int base = 0, Es = 0;
  bool found = false;
  for (int E = etop; E >= eT; --E) {
    int sz[JPT], loc = 0;
#pragma unroll
    for (int s = 0; s < JPT; ++s) {
      int v = 0;
      if (ej[s] >= E && (E > eT || mj[s] >= mT)) {
        const int d = ej[s] - E;
        v = d >= 20 ? KILL_CAP + 1 : min(1 << d, KILL_CAP + 1);
      }
      sz[s] = v;
      loc += v;
    }
    int tot;
    const int ex = block_excl_scan<NT>(loc, s_i, tot);
    int stj[JPT];
    int st = base + ex;
#pragma unroll
    for (int s = 0; s < JPT; ++s) {
      stj[s] = st;
      if (tid * JPT + s < C) start[tid * JPT + s] = st;
      st += sz[s];
    }
    __syncthreads();
    const int hi = min(base + tot, K);
    for (int t = base + tid; t < hi; t += NT) {
      // the last family starting at or before t (families of size 0 start where the next does)
      int lo = 0, h = C - 1;
      while (lo < h) {
        const int mid = (lo + h + 1) >> 1;
        if (start[mid] <= t) lo = mid;
        else h = mid - 1;
      }
      A.asg[t] = lo;
    }
    if (base + tot >= K) {
#pragma unroll
      for (int s = 0; s < JPT; ++s) {
        if (sz[s] > 0 && stj[s] < K && stj[s] + sz[s] >= K) {
          s_sh[0] = tid * JPT + s;
          s_sh[1] = stj[s];
        }
      }
      Es = E;
      found = true;
      __syncthreads();
      break;
    }
    base += tot;
  }
  if (!found) return RC_FEW;
  if (tie_e >= Es) return RC_OTHER;
  const int p = s_sh[0], startp = s_sh[1];
  const int r = K - startp;
  const Cand cp = cand16(ord[p]);
  const int Dp = dexp(cp.w) - Es;
  const int m = 1 << Dp;
  if (Dp > 13) return RC_OTHER;  // the family's slots must fit the scratch
  int Dj[JPT];
#pragma unroll
  for (int s = 0; s < JPT; ++s) {
    const int j = tid * JPT + s;
    Dj[s] = (j >= C || ej[s] < Es) ? 0 : ej[s] - Es + (j < p ? 1 : 0);
    if (j < C) dep[j] = (unsigned char)Dj[s];
  }
  // 5. the slots of family p: the kill slots of its full levels, then its own
  int* sp = start;
  int cnt = 0;
  for (int t0 = 0; t0 < startp; t0 += NT) {
    const int t = t0 + tid;
    const bool mem = t < startp && A.asg[t] == p;
    int tot;
    const int ex = block_excl_scan<NT>(mem ? 1 : 0, s_i, tot);
    if (mem && cnt + ex < m) sp[cnt + ex] = A.klist[t];
    cnt += tot;
  }
  if (cnt + 1 != m) return RC_OTHER;  // not reached: family p has 2^Dp entries
  if (tid == 0) sp[m - 1] = cp.slot;
  __syncthreads();
  // 6. weights in place, and the list of copies (kill slot, source) for the copy kernel;
  //    every source is a candidate, which is never a kill slot
  for (int i = tid; i < m; i += NT) {
    const int sl = sp[i];
    int rank = 0;
    for (int k = 0; k < m; ++k) rank += sp[k] < sl ? 1 : 0;
    A.w[sl] = ldexp(cp.w, -(Dp + (rank < r ? 1 : 0)));
  }
  for (int t = tid; t < K; t += NT) {
    const int j = A.asg[t];
    const int ks = A.klist[t];
    int o;
    double wn;
    if (j == p) {
      o = cp.slot;
      wn = ldexp(cp.w, -(Dp + 1));
    } else {
      const Cand c = cand16(ord[j]);
      o = c.slot;
      wn = ldexp(c.w, -(int)dep[j]);
    }
    if (j != p || t >= startp) A.w[ks] = wn;
    A.cpy[2 * t] = ks;
    A.cpy[2 * t + 1] = o;
  }
#pragma unroll
  for (int s = 0; s < JPT; ++s) {
    const int j = tid * JPT + s;
    if (j < C && j != p && Dj[s] > 0) A.w[oj[s]] = ldexp(wj[s], -Dj[s]);
  }
  cut = ldexp(cp.w, -Dp);
  return RC_DONE;
}

// bucket counts of the global-memory replay's counting sort (in the replay's LDS buffer), and
// the largest bucket it ranks
constexpr int RCG_BUCKETS = 16384;
constexpr int RCG_BMAX = 256;
static_assert(RCG_BUCKETS * (int)sizeof(int) <= RP_LDS && RCG_BUCKETS % RP_NT == 0, "bucket counts fit the LDS buffer");

// The same closed-form replay when the kills outgrow the LDS sort (1.25 K > CAND_CAP): up to
// CAND_G candidates, sorted and scanned through an L2-resident scratch (A.rsg) instead of LDS
// and registers.  One workgroup runs it, so each pass over the scratch is bound by what one CU
// issues: the latency of its loads, and for scattered addresses about one cache line per
// cycle.  Every streaming loop issues RG_U independent loads per thread before it uses them,
// the octave scans give each thread a contiguous run of candidates (one block scan per
// octave), and the scattered weight and copy writes are the chip-wide copy kernel's.  Equal
// keys are ranked by position; where that order would matter (a tie in an octave that is
// used) the replay declines.  The sort key holds the weight's exponent and mantissa whole (the weight is
// rebuilt from it) and the candidate's walker slot travels with it, so no pass after the
// collection reads the lists again.  Bucket counts of the counting sort live in LDS (`bcnt`,
// RCG_BUCKETS ints).  It declines where the LDS form declines, and also when one mantissa
// bucket holds more than RCG_BMAX candidates.
constexpr int RG_U = 4;  // loads in flight per thread in the streaming loops

// the replay scratch: keys, bucketed keys (8 B each), walker slots, bucketed slots, family-p
// slots (4 B each), split depths (1 B), CAND_G of each
struct RsgView {
  unsigned long long* key;
  unsigned long long* bkey;
  int* ord;
  int* bord;
  int* start;
  unsigned char* dep;
};
PVD_DEV RsgView rsg_view(unsigned char* rsg) {
  RsgView v;
  v.key = reinterpret_cast<unsigned long long*>(rsg);
  v.bkey = v.key + CAND_G;
  v.ord = reinterpret_cast<int*>(v.bkey + CAND_G);
  v.bord = v.ord + CAND_G;
  v.start = v.bord + CAND_G;
  v.dep = reinterpret_cast<unsigned char*>(v.start + CAND_G);
  return v;
}

PVD_DEV double key_weight(unsigned long long k) {
  constexpr unsigned long long MANT = (1ull << 52) - 1;
  return __longlong_as_double((long long)(((0x7FFull - (k & 0x7FFull)) << 52) | (MANT - (k >> 11))));
}

template <typename R>
PVD_DEV int replay_candidates_g(const StepArgs<R>& A, int K, int nlists, const int* subn, double T0, int* bcnt,
                                int* s_i, int* s_sh, int* hist, double& cut, bool& refined) {
  constexpr int NT = RP_NT;
  constexpr unsigned long long MANT = (1ull << 52) - 1;
  constexpr int NS = 2 * LIST_SUB;
  constexpr int U = RG_U;
  const int tid = threadIdx.x;
  int pre[NS + 1];
  pre[0] = 0;
#pragma unroll
  for (int c = 0; c < NS; ++c) {
    const int k = c < nlists * LIST_SUB ? subn[c] : 0;
    if (k > SUB_CAP) return RC_MANY;  // entries were lost
    pre[c + 1] = pre[c] + k;
  }
  const int nl = pre[NS];
  if (nl <= 0) return RC_FEW;
  if (K > KILL_CAP) return RC_OTHER;
  if (!(T0 >= A.thresh_lo) || !(T0 > 0.0)) return RC_OTHER;
  auto phys = [&](int q) -> int {
    int u0 = 0, b = 0;
#pragma unroll
    for (int u = 0; u < NS; ++u)
      if (q >= pre[u]) {
        u0 = u;
        b = pre[u];
      }
    return u0 * SUB_CAP + (q - b);
  };
  const long long tb = th_base(A.thresh_lo);
  // key = (inverted mantissa << 11) | inverted exponent; ord holds the walker slot
  const RsgView rv = rsg_view(A.rsg);
  unsigned long long* key = rv.key;
  unsigned long long* bkey = rv.bkey;
  int* ord = rv.ord;
  int* bord = rv.bord;
  int* start = rv.start;
  unsigned char* dep = rv.dep;
  auto kexp = [](unsigned long long k) { return (int)(0x7FFull - (k & 0x7FFull)) - 1023; };
  auto kmant = [](unsigned long long k) { return MANT - (k >> 11); };
  // 1. the cut: T0, or, when the lists hold over a quarter more than about 1.25 K walkers
  //    (at a steady state list A holds the last step's 2 K split pieces), the lowest
  //    histogram edge with at most 1.25 K entries at or above it: every later pass is over
  //    the walkers >= the cut only
  double T = T0;
  const int lim = min(CAND_G, max(K + K / 4, 1024));
  refined = nl > lim + lim / 4;
  if (refined) {
    for (int b = tid; b < TH_BINS; b += NT) hist[b] = 0;
    __syncthreads();
    for (int q0 = 0; q0 < nl; q0 += U * NT) {
      double wq[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int q = q0 + u * NT + tid;
        wq[u] = q < nl ? A.cand[phys(q)].w : 0.0;
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (q0 + u * NT + tid < nl) atomicAdd(&hist[th_bin(wq[u], tb)], 1);
    }
    __syncthreads();
    const int over = th_cross(hist, lim, true, s_i, s_sh);
    if (over < 0 || over >= TH_BINS - 1) return RC_MANY;
    // and no higher than the edge that keeps K walkers: one bin (0.4 % of an octave) can
    // hold thousands when the weights are bunched, as in the first steps of a run
    const int reach = th_cross(hist, K, false, s_i, s_sh);
    if (reach >= 0) T = fmax(T0, th_edge(A.thresh_lo, min(over + 1, reach)));
  }
  // the walkers >= T, in any order (wave ballots onto one counter): key and walker slot, and
  // the range of the keys
  if (tid == 0) s_sh[3] = 0;
  __syncthreads();
  unsigned long long kmin = ~0ull, kmax = 0ull;
  for (int q0 = 0; q0 < nl; q0 += U * NT) {
    Cand c[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int q = q0 + u * NT + tid;
      if (q < nl) {
        c[u] = A.cand[phys(q)];
      } else {
        c[u].w = -1.0;
        c[u].slot = 0;
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const bool in = c[u].w >= T;
      const unsigned long long bal = __ballot(in);
      if (bal) {
        const int lead = __ffsll((long long)bal) - 1;
        int base = 0;
        if (lane_id() == lead) base = atomicAdd(&s_sh[3], __popcll(bal));
        base = __shfl(base, lead, PVD_WAVE);
        const int o = base + __popcll(bal & ((1ull << lane_id()) - 1ull));
        if (in && o < CAND_G) {
          const unsigned long long wb = (unsigned long long)__double_as_longlong(c[u].w);
          const unsigned long long k = ((MANT - (wb & MANT)) << 11) | (0x7FFull - ((wb >> 52) & 0x7FFull));
          key[o] = k;
          ord[o] = c[u].slot;
          kmin = k < kmin ? k : kmin;
          kmax = k > kmax ? k : kmax;
        }
      }
    }
  }
  {
    unsigned long long* s_k = reinterpret_cast<unsigned long long*>(hist);  // free after step 1
    kmax = block_max<NT>(kmax, s_k);
    kmin = ~block_max<NT>(~kmin, s_k);
  }
  const int C = s_sh[3];
  if (C > CAND_G) return RC_MANY;
  if (C == 0) return RC_FEW;
  // 2. candidates by decreasing mantissa (then exponent): counting sort on 14 bits of the key
  //    above its minimum (an order-preserving map of the keys' own range onto the buckets: the
  //    candidates of a steady state lie in a band about a tenth of an octave wide, which the
  //    top mantissa bits would crowd into a tenth of the buckets), then each entry ranked
  //    inside its bucket
  constexpr int BPT = RCG_BUCKETS / NT;
  const unsigned long long kspan = kmax - kmin;
  const int bsh = (kspan >> 14) == 0ull ? 0 : 64 - __clzll((long long)kspan) - 14;
  auto bucket = [&](unsigned long long k) { return (int)((k - kmin) >> bsh); };  // < RCG_BUCKETS
  for (int b = tid; b < RCG_BUCKETS; b += NT) bcnt[b] = 0;
  __syncthreads();
  for (int q0 = 0; q0 < C; q0 += U * NT) {
    unsigned long long kk[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int q = q0 + u * NT + tid;
      kk[u] = q < C ? key[q] : 0ull;
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (q0 + u * NT + tid < C) atomicAdd(&bcnt[bucket(kk[u])], 1);
  }
  __syncthreads();
  int bmax;
  {
    int loc = 0, mx = 0;
    for (int u = 0; u < BPT; ++u) {
      const int c = bcnt[tid * BPT + u];
      loc += c;
      mx = max(mx, c);
    }
    int tot;
    const int ex = block_excl_scan<NT>(loc, s_i, tot);
    int st = ex;
    for (int u = 0; u < BPT; ++u) {
      const int c = bcnt[tid * BPT + u];
      bcnt[tid * BPT + u] = st;  // bucket starts (each thread rewrites only its own buckets)
      st += c;
    }
    bmax = block_max<NT>(mx, s_i);
  }
  if (bmax > RCG_BMAX) return RC_OTHER;  // the in-bucket rank costs bucket size squared
  for (int q0 = 0; q0 < C; q0 += U * NT) {
    unsigned long long kk[U];
    int oo[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int q = q0 + u * NT + tid;
      kk[u] = q < C ? key[q] : 0ull;
      oo[u] = q < C ? ord[q] : 0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (q0 + u * NT + tid < C) {
        const int d = atomicAdd(&bcnt[bucket(kk[u])], 1);  // bcnt ends as the bucket ends
        bkey[d] = kk[u];
        bord[d] = oo[u];
      }
  }
  __syncthreads();
  // ranks: U entries per thread at a time, the first RK keys of each one's bucket in flight
  constexpr int RK = 4;
  for (int q0 = 0; q0 < C; q0 += U * NT) {
    unsigned long long kq[U];
    int oq[U], bs[U], be[U], rank[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int q = q0 + u * NT + tid;
      kq[u] = q < C ? bkey[q] : 0ull;
      oq[u] = q < C ? bord[q] : 0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int q = q0 + u * NT + tid;
      const int bk = q < C ? bucket(kq[u]) : 0;
      bs[u] = q < C ? (bk > 0 ? bcnt[bk - 1] : 0) : 0;
      be[u] = q < C ? bcnt[bk] : 0;
      rank[u] = 0;
    }
#pragma unroll
    for (int i = 0; i < RK; ++i)
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int x = bs[u] + i, q = q0 + u * NT + tid;
        if (x < be[u]) {
          const unsigned long long ki = bkey[x];
          rank[u] += (ki < kq[u] || (ki == kq[u] && x < q)) ? 1 : 0;
        }
      }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int q = q0 + u * NT + tid;
      for (int x0 = bs[u] + RK; x0 < be[u]; x0 += 8) {  // larger buckets: eight loads in flight
        unsigned long long k8[8];
#pragma unroll
        for (int v = 0; v < 8; ++v) k8[v] = x0 + v < be[u] ? bkey[x0 + v] : ~0ull;
#pragma unroll
        for (int v = 0; v < 8; ++v) {
          const int x = x0 + v;
          rank[u] += (x < be[u] && (k8[v] < kq[u] || (k8[v] == kq[u] && x < q))) ? 1 : 0;
        }
      }
      if (q < C) {
        key[bs[u] + rank[u]] = kq[u];
        ord[bs[u] + rank[u]] = oq[u];
      }
    }
  }
  __syncthreads();
  // 3. the top octave, and the highest octave in which two equal mantissas meet
  int etop = -0x40000000, tie_e = -0x40000000;
  for (int j0 = 0; j0 < C; j0 += U * NT) {
    unsigned long long kk[U], kn[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int j = j0 + u * NT + tid;
      kk[u] = j < C ? key[j] : 0ull;
      kn[u] = j + 1 < C ? key[j + 1] : 0ull;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int j = j0 + u * NT + tid;
      if (j < C) {
        etop = max(etop, kexp(kk[u]));
        if (j + 1 < C && (kn[u] >> 11) == (kk[u] >> 11)) tie_e = max(tie_e, min(kexp(kk[u]), kexp(kn[u])));
      }
    }
  }
  etop = block_max<NT>(etop, s_i);
  tie_e = block_max<NT>(tie_e, s_i);
  const int eT = dexp(T);
  const unsigned long long mT = (unsigned long long)__double_as_longlong(T) & MANT;
  if (etop - eT > 40) return RC_OTHER;
  // 4. octaves from the top: each family's run of kill indices, until K are used; in the
  //    octave of T only the entries of value >= T count.  Thread tid holds the candidates
  //    [j0, j1) in sorted order: its entry count, one block scan, then its families' runs.
  auto fam = [&](unsigned long long kj, int E) -> int {
    const int ej = kexp(kj);
    if (!(ej >= E && (E > eT || kmant(kj) >= mT))) return 0;
    const int d = ej - E;
    return d >= 20 ? KILL_CAP + 1 : min(1 << d, KILL_CAP + 1);
  };
  const int P = (C + NT - 1) / NT;
  const int j0 = min(C, tid * P), j1 = min(C, j0 + P);
  int base = 0, Es = 0;
  bool found = false;
  for (int E = etop; E >= eT; --E) {
    int loc = 0;
    for (int jb = j0; jb < j1; jb += U) {
      unsigned long long kk[U];
#pragma unroll
      for (int u = 0; u < U; ++u) kk[u] = jb + u < j1 ? key[jb + u] : 0ull;
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (jb + u < j1) loc = min(loc + fam(kk[u], E), KILL_CAP + 1);
    }
    // a thread's count is clamped above K: the block sums stay in int, and every later
    // thread still starts at or beyond K
    int tot;
    const int ex = block_excl_scan<NT>(loc, s_i, tot);
    int st = base + ex;
    for (int jb = j0; jb < j1 && st < K; jb += U) {
      unsigned long long kk[U];
#pragma unroll
      for (int u = 0; u < U; ++u) kk[u] = jb + u < j1 ? key[jb + u] : 0ull;
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int j = jb + u;
        const int v = j < j1 ? fam(kk[u], E) : 0;
        if (v > 0 && st < K) {
          const int hi = min(st + v, K);
          for (int t = st; t < hi; ++t) A.asg[t] = j;
          if (st + v >= K) {
            s_sh[0] = j;
            s_sh[1] = st;
          }
        }
        st += v;
      }
    }
    base += tot;
    __syncthreads();
    if (base >= K) {
      Es = E;
      found = true;
      break;
    }
  }
  if (!found) return RC_FEW;
  if (tie_e >= Es) return RC_OTHER;
  const int p = s_sh[0], startp = s_sh[1];
  const int r = K - startp;
  const double cpw = key_weight(key[p]);
  const int cps = ord[p];
  const int Dp = dexp(cpw) - Es;
  const int m = 1 << Dp;
  if (Dp > 13) return RC_OTHER;  // the family's slots must fit the scratch
  // split depth of each candidate (read by the copy kernel)
  for (int jb = j0; jb < j1; jb += U) {
    unsigned long long kk[U];
#pragma unroll
    for (int u = 0; u < U; ++u) kk[u] = jb + u < j1 ? key[jb + u] : 0ull;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int j = jb + u;
      if (j < j1) {
        const int ej = kexp(kk[u]);
        dep[j] = (unsigned char)(ej < Es ? 0 : ej - Es + (j < p ? 1 : 0));
      }
    }
  }
  // 5. the slots of family p: the kill slots of its full levels (any order: the weights
  //    below rank them by slot), then its own
  int* sp = start;
  if (tid == 0) s_sh[3] = 0;
  __syncthreads();
  for (int t0 = 0; t0 < startp; t0 += U * NT) {
    int aa[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int t = t0 + u * NT + tid;
      aa[u] = t < startp ? A.asg[t] : -1;
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (aa[u] == p) {
        const int o = atomicAdd(&s_sh[3], 1);
        if (o < m - 1) sp[o] = A.klist[t0 + u * NT + tid];
      }
  }
  __syncthreads();
  if (s_sh[3] + 1 != m) return RC_OTHER;  // not reached: family p has 2^Dp entries
  if (tid == 0) sp[m - 1] = cps;
  __syncthreads();
  // 6. the weights of family p's own slots; every other write (the kill slots' sources and
  //    weights, the split candidates' own weights) is the copy kernel's, chip-wide
  for (int i = tid; i < m; i += NT) {
    const int sl = sp[i];
    int rank = 0;
    for (int k = 0; k < m; ++k) rank += sp[k] < sl ? 1 : 0;
    A.w[sl] = ldexp(cpw, -(Dp + (rank < r ? 1 : 0)));
  }
  if (tid == 0) {
    Ctrl* Cw = A.ctrl;
    Cw->gc_n = C;
    Cw->gc_p = p;
    Cw->gc_slot = cps;
    Cw->gc_dp = Dp;
    Cw->gc_start = startp;
    Cw->gc_w = cpw;
  }
  cut = ldexp(cpw, -Dp);
  return RC_DONE;
}

template <typename R>
__global__ __launch_bounds__(RP_NT) void k_cont_replay(StepArgs<R> A) {
  __shared__ __align__(16) unsigned char s_buf[RP_LDS];
  Piece* cur = reinterpret_cast<Piece*>(s_buf);
  __shared__ int s_ncur, s_npend, s_nfin, s_fail, s_next;
  __shared__ double s_d[RP_NT / PVD_WAVE];
  __shared__ int s_i[RP_NT / PVD_WAVE];
  __shared__ int s_K, s_nc;
  __shared__ int s_subn[2 * LIST_SUB];
  __shared__ double s_a0n, s_a1n;
  __shared__ int s_sh[4];
  __shared__ int s_hist[TH_BINS];
  const int par = A.par;
  Ctrl* C = A.ctrl;
  const int np1 = par ^ 1;
  if (C->status != ST_OK) return;
  const int n = C->n[par];
  const int ntiles = (n + TILE - 1) / TILE;
  const unsigned long long step = C->step[par];
  const unsigned long long ep = (unsigned long long)C->epoch[par];
  const int tid = threadIdx.x;
  int* kl = A.klist;
  // 0. the kill list in ascending slot order (from the per-wave masks) and the fixed-order
  //    sums of the tile records, their loads issued together; without splits the sums
  //    finish the step (calc_vref, continuous: pyvibdmc.py:659-661)
  {
    const unsigned acct = A.is_ ? acc_total(A, par) : 0u;
    const int nwords = ntiles * (TILE / PVD_WAVE);
    constexpr int KW = 16;  // mask words per thread and round
    int kbase = 0;
    double a0 = 0.0, a1 = 0.0, a0n = 0.0, a1n = 0.0;
    for (int base = 0; base < nwords || base == 0; base += RP_NT * KW) {
      const int span = min(nwords - base, RP_NT * KW);
      const int per = (span + RP_NT - 1) / RP_NT;
      const int w0 = base + tid * per;
      unsigned long long m[KW];
      int pc = 0;
#pragma unroll
      for (int u = 0; u < KW; ++u) {
        m[u] = (u < per && w0 + u < base + span) ? A.kmask[w0 + u] : 0ull;
        pc += __popcll(m[u]);
      }
      if (base == 0) {
        for (int t = tid; t < ntiles; t += RP_NT) {
          const TileRec& rc = A.rec[t];
          a0 += rc.s0;
          a1 += rc.s1;
          a0n += rc.s2;
          a1n += rc.s3;
        }
      }
      int tot;
      const int ex = block_excl_scan<RP_NT>(pc, s_i, tot);
      int o = kbase + ex;
#pragma unroll
      for (int u = 0; u < KW; ++u) {
        unsigned long long mm = m[u];
        while (mm) {
          const int b = __ffsll((long long)mm) - 1;
          mm &= mm - 1ull;
          if (o < KILL_CAP) kl[o] = (w0 + u) * PVD_WAVE + b;
          ++o;
        }
      }
      kbase += tot;
    }
    const int k = kbase;
    {
      // one pass: wave partials, combined by one thread in a fixed order
      constexpr int NW = RP_NT / PVD_WAVE;
      const int wv = tid / PVD_WAVE;
      const double v0 = wave_sum(a0), v1 = wave_sum(a1), v2 = wave_sum(a0n), v3 = wave_sum(a1n);
      double* sr = reinterpret_cast<double*>(s_buf);
      if (lane_id() == 0) {
        sr[wv] = v0;
        sr[NW + wv] = v1;
        sr[2 * NW + wv] = v2;
        sr[3 * NW + wv] = v3;
      }
      __syncthreads();
      a0 = sr[0];
      a1 = sr[NW];
      a0n = sr[2 * NW];
      a1n = sr[3 * NW];
      for (int q = 1; q < NW; ++q) {
        a0 += sr[q];
        a1 += sr[NW + q];
        a0n += sr[2 * NW + q];
        a1n += sr[3 * NW + q];
      }
    }
    if (tid == 0) {
      s_K = k;
      s_a0n = a0n;
      s_a1n = a1n;
      C->ncopy = 0;
      C->gc_n = 0;
      // the candidate count of this step; the next step's weight kernel starts from zero
      s_nc = 0;
      for (int c = 0; c < LIST_SUB; ++c) {
        s_subn[c] = (int)min(*list_count(A, 0, c), (unsigned)SUB_CAP + 1u);
        s_subn[LIST_SUB + c] = (int)min(*list_count(A, 1, c), (unsigned)SUB_CAP + 1u);
        s_nc += s_subn[c];
        *list_count(A, 0, c) = 0u;
        *list_count(A, 1, c) = 0u;
      }
      C->kills[par] = k;
      C->sumw[par] = a1;
      C->sumwe[par] = a0;
      C->n[np1] = n;
      C->step[np1] = step + 1;
      C->epoch[np1] = C->epoch[par] + 1;
      if (A.is_ && (long long)step < A.tmax) {
        const double dte = A.dt * ((double)acct / (double)n);
        C->eff_t = (step == 0 ? 0.0 : C->eff_t) + dte;
        A.tr_dteff[step] = C->eff_t;
      }
      // the step's weights and every thread here have read the acceptance count (the scans
      // above hold barriers): clear it for step + 2
      if (A.is_) {
        for (int b = 0; b < ACC_SLOTS; ++b) st_relaxed_u(acc_slot(A, par, b), 0u);
      }
      if (k == 0) {
        const double vnew = a0 / a1 - (A.alpha * ((a1 - A.n0) / A.n0));
        C->vref[np1] = vnew;
        if ((long long)step < A.tmax) {
          A.tr_vref[step] = vnew;
          A.tr_pop[step] = a1;
          A.tr_births[step] = 0;
          A.tr_deaths[step] = 0;
        }
      }
    }
    __syncthreads();
  }
  const int K = s_K;
  if (K == 0) return;
  const unsigned long long tagbase = ep << 32;
  if (K > KILL_CAP) {
    if (tid == 0) {
      C->status = ST_BRANCH;
      C->status_step = (long long)step;
    }
    return;
  }
  // the step's vref from the sums after the split, and the next step's lists: the
  // threshold follows this step's cut, its factor moved by (size of list A / wanted)^0.2,
  // or, after the interval replay, is read off the histogram of all weights (t0 > 0)
  auto finish = [&](double a0, double a1, double cut, int code, double t0, bool miss) {
    if (tid == 0) {
      const double vnew = a0 / a1 - (A.alpha * ((a1 - A.n0) / A.n0));
      C->vref[np1] = vnew;
      if ((long long)step < A.tmax) {
        A.tr_vref[step] = vnew;
        A.tr_pop[step] = a1;
        A.tr_births[step] = K;
        A.tr_deaths[step] = code;  // which replay ran (and why the lists did not serve)
      }
      float g = C->cand_g;
      if (!(g > 0.0f)) g = 0.5f;
      if (t0 > 0.0 && cut > 0.0) g = (float)(t0 / cut);
      else if (C->cand_t > 0.0) g *= s_nc > 0 ? powf((float)s_nc / (float)th_target(K), 0.2f) : 0.5f;
      g = fminf(0.98f, fmaxf(0.02f, g));
      C->cand_g = g;
      C->cand_t = cut > 0.0 ? fmax(A.thresh_lo, cut * g) : 0.0;
      const int nb = miss ? 64 : max(0, C->cand_b - 1);
      C->cand_b = nb;
      C->cand_t2 = (cut > 0.0 && nb > 0) ? fmax(A.thresh_lo, 0.75 * cut * g) : C->cand_t;
    }
  };
  // 2. candidate replay from list A, else from lists A and B; the interval replay below
  //    covers what they leave
  int rcode;
  {
    double cut = 0.0;
    bool refined = false;
    // more kills than the LDS sort serves: the global-memory form; it also takes the lists
    // the LDS form declines as too long
    bool big = K + K / 4 > CAND_CAP && A.rsg != nullptr;
    auto replay = [&](int nlists, double T0) {
      return big ? replay_candidates_g(A, K, nlists, s_subn, T0, reinterpret_cast<int*>(s_buf), s_i, s_sh, s_hist,
                                       cut, refined)
                 : replay_candidates(A, K, nlists, s_subn, T0, s_buf, s_i, s_sh, s_hist, cut, refined);
    };
    int nlists = 1;
    double t_used = C->cand_t;
    rcode = replay(nlists, t_used);
    const bool miss = rcode != RC_DONE;
    if (rcode == RC_FEW && C->cand_t2 < C->cand_t) {
      __syncthreads();
      nlists = 2;
      t_used = C->cand_t2;
      rcode = replay(nlists, t_used);
    }
    if (rcode == RC_MANY && !big && A.rsg != nullptr) {
      __syncthreads();
      big = true;
      rcode = replay(nlists, t_used);
    }
    if (rcode == RC_DONE) {
      if (tid == 0) C->ncopy = K;
      finish(s_a0n, s_a1n, cut, big ? (miss ? 8 : 7) : miss ? 6 : refined ? 5 : 1, 0.0, miss);
      return;
    }
  }
  __syncthreads();
  // kill index of each killed slot, for the interval replay's validity rule
  for (int q = tid; q < K; q += RP_NT) A.killtag[kl[q]] = tagbase | (unsigned long long)q;
  __syncthreads();
  if (K > RP_KCAP) {
    if (tid == 0) {
      C->status = ST_BRANCH;
      C->status_step = (long long)step;
    }
    return;
  }
  // 3. global maximum weight; first interval [gm, 2 gm)
  double gm = 0.0;
  for (int t = tid; t < ntiles; t += RP_NT) gm = fmax(gm, A.rec[t].mx);
  gm = block_max<RP_NT>(gm, s_d);
  if (tid == 0) {
    s_npend = 0;
    s_nfin = 0;
    s_fail = 0;
  }
  __syncthreads();
  double H = 2.0 * gm;
  int next = 0;
  int pbuf = 0;  // which half of `pend` holds the live pending list
  int guard = 0;
  const int wv = tid / PVD_WAVE, l = lane_id();
  while (next < K) {
    // choose L in [H/2, H) so that the interval fits in LDS
    double L = 0.5 * H;
    for (int attempt = 0;; ++attempt) {
      int c = 0;
      for (int t = wv; t < ntiles; t += RP_NT / PVD_WAVE) {
        if (A.rec[t].mx < L) continue;
        const long long s0 = (long long)t * TILE + l;
        for (int q = 0; q < TILE / PVD_WAVE; ++q) {
          const long long ss = s0 + q * PVD_WAVE;
          if (ss < n) {
            const double w = A.w[ss];
            c += (w >= L && w < H) ? 1 : 0;
          }
        }
      }
      c = block_sum<RP_NT>(c, s_i);
      int pin = 0;
      for (int q = tid; q < s_npend; q += RP_NT) {
        const double v = A.pend[pbuf * RP_PCAP + q].v;
        pin += (v >= L && v < H) ? 1 : 0;
      }
      pin = block_sum<RP_NT>(pin, s_i);
      if (c + pin <= RP_CAP) break;
      L = 0.5 * (L + H);
      if (attempt > 60) {
        if (tid == 0) s_fail = 1;
        break;
      }
    }
    __syncthreads();
    if (s_fail || !(L > 1e-300) || !(L < H) || ++guard > 4096) {
      if (tid == 0) {
        C->status = ST_BRANCH;
        C->status_step = (long long)step;
      }
      return;
    }
    const bool degenerate = L < A.thresh_lo;
    // gather the interval: in-range pending pieces, then original walkers in range
    if (tid == 0) s_ncur = 0;
    const int npend = s_npend;
    const Piece* pin_list = A.pend + pbuf * RP_PCAP;
    Piece* pout_list = A.pend + (pbuf ^ 1) * RP_PCAP;
    __syncthreads();
    if (tid == 0) s_npend = 0;
    __syncthreads();
    // in-range pieces join the interval (its order is fixed by the sort below); the
    // others are carried over in their input order (scan offsets, deterministic)
    {
      int pbase = 0;
      for (int q0 = 0; q0 < npend; q0 += RP_NT) {
        const int q = q0 + tid;
        Piece pc;
        bool in = false, keep = false;
        if (q < npend) {
          pc = pin_list[q];
          in = pc.v >= L && pc.v < H;
          keep = !in;
        }
        if (in) cur[atomicAdd(&s_ncur, 1)] = pc;
        int tot;
        const int ex = block_excl_scan<RP_NT>(keep ? 1 : 0, s_i, tot);
        if (keep) pout_list[pbase + ex] = pc;
        pbase += tot;
      }
      if (tid == 0) s_npend = pbase;
      __syncthreads();
    }
    for (int t = wv; t < ntiles; t += RP_NT / PVD_WAVE) {
      if (A.rec[t].mx < L) continue;
      const long long s0 = (long long)t * TILE + l;
      for (int q = 0; q < TILE / PVD_WAVE; ++q) {
        const long long ss = s0 + q * PVD_WAVE;
        if (ss < n) {
          const double w = A.w[ss];
          if (w >= L && w < H) {
            Piece pc;
            pc.v = w;
            pc.slot = (int)ss;
            pc.src = (int)ss;
            pc.birth = -1;
            pc.pad = 0;
            cur[atomicAdd(&s_ncur, 1)] = pc;
          }
        }
      }
    }
    __syncthreads();
    const int ncur = s_ncur;
    // bitonic sort of cur[0, P), P = next power of two, sentinels at the end
    int P = 1;
    while (P < ncur) P <<= 1;
    for (int q = ncur + tid; q < P; q += RP_NT) {
      Piece z;
      z.v = -1.0;
      z.slot = 0x7FFFFFFF;
      z.src = -1;
      z.birth = -1;
      z.pad = 0;
      cur[q] = z;
    }
    __syncthreads();
    for (int kk = 2; kk <= P; kk <<= 1) {
      for (int jj = kk >> 1; jj > 0; jj >>= 1) {
        for (int q = tid; q < P; q += RP_NT) {
          const int ixj = q ^ jj;
          if (ixj > q) {
            const bool up = (q & kk) == 0;
            const Piece a = cur[q], b = cur[ixj];
            if (up ? piece_before(b, a) : piece_before(a, b)) {
              cur[q] = b;
              cur[ixj] = a;
            }
          }
        }
        __syncthreads();
      }
    }
    if (!degenerate) {
      // members [0, m) are chosen in order and split (children at fixed positions); the
      // rest are final (appended at scan offsets)
      const int m = min(ncur, K - next);
      const int pb = s_npend, fb = s_nfin;
      int fbase = fb;
      for (int q0 = 0; q0 < ncur; q0 += RP_NT) {
        const int q = q0 + tid;
        bool fin_here = false;
        Piece pc;
        if (q < ncur) {
          pc = cur[q];
          if (q < m) {
            Piece a = pc, b = pc;
            a.v = pc.v * 0.5;
            a.birth = next + q;
            b.v = a.v;
            b.birth = next + q;
            b.slot = kl[next + q];
            const int o = pb + 2 * q;
            if (o + 2 <= RP_PCAP) {
              pout_list[o] = a;
              pout_list[o + 1] = b;
            } else {
              s_fail = 1;
            }
          } else {
            fin_here = pc.birth >= 0;
          }
        }
        int tot;
        const int ex = block_excl_scan<RP_NT>(fin_here ? 1 : 0, s_i, tot);
        if (fin_here) {
          if (fbase + ex < RP_FCAP) A.fin[fbase + ex] = pc;
          else s_fail = 1;
        }
        fbase += tot;
      }
      __syncthreads();
      if (tid == 0) {
        s_npend = pb + 2 * m;
        s_nfin = fbase;
      }
      next += m;
    } else {
      // entry-by-entry replay of this interval by one thread
      if (tid == 0) {
        int ptr = 0, q = next;
        for (; q < K; ++q) {
          while (ptr < ncur && !piece_live(cur[ptr], q, A.killtag, ep)) ++ptr;
          if (ptr >= ncur) break;
          const Piece e = cur[ptr++];
          const int ks = kl[q];
          Piece a = e;
          a.v = e.v * 0.5;
          a.birth = q;
          if (e.slot == ks) {
            if (s_npend + 1 <= RP_PCAP) pout_list[s_npend++] = a;
            else s_fail = 1;
          } else {
            Piece b = a;
            b.slot = ks;
            if (s_npend + 2 <= RP_PCAP) {
              pout_list[s_npend++] = a;
              pout_list[s_npend++] = b;
            } else {
              s_fail = 1;
            }
          }
        }
        for (; ptr < ncur; ++ptr) {
          if (cur[ptr].birth >= 0) {
            if (s_nfin < RP_FCAP) A.fin[s_nfin++] = cur[ptr];
            else s_fail = 1;
          }
        }
        s_next = q;
      }
      __syncthreads();
      next = s_next;
    }
    __syncthreads();
    if (s_fail) {
      if (tid == 0) {
        C->status = ST_BRANCH;
        C->status_step = (long long)step;
      }
      return;
    }
    H = L;
    pbuf ^= 1;
  }
  // every still-pending piece is final
  {
    const int npend = s_npend;
    const int fb = s_nfin;
    const Piece* pl = A.pend + pbuf * RP_PCAP;
    for (int q = tid; q < npend; q += RP_NT) {
      if (fb + q < RP_FCAP) A.fin[fb + q] = pl[q];
      else s_fail = 1;
    }
    __syncthreads();
    if (tid == 0) s_nfin = fb + npend;
    __syncthreads();
    if (s_fail) {
      if (tid == 0) {
        C->status = ST_BRANCH;
        C->status_step = (long long)step;
      }
      return;
    }
  }
  // read phase: sources and corrections from the pre-split state (pieces overwritten by
  // their slot's own turn are dropped)
  const int nfin = s_nfin;
  const int np = A.np;
  const long long rsw = np + 2 + A.nx;  // staged doubles per copy
  double lsw = 0.0, lswe = 0.0;
  for (int q = tid; q < nfin; q += RP_NT) {
    Piece pc = A.fin[q];
    if (!piece_live(pc, K, A.killtag, ep)) {
      pc.slot = -1;
      A.fin[q] = pc;
      continue;
    }
    const int s = pc.slot;
    const double wold = A.w[s];
    const double esrc = A.E[pc.src];
    lsw += pc.v - wold;
    lswe += pc.v * esrc - wold * A.E[s];
    if (s != pc.src) {
      double* st = A.rstage + (long long)q * rsw;
      for (int p = 0; p < np; ++p) st[p] = (double)A.cout[p * A.cap + pc.src];
      st[np] = esrc;
      st[np + 1] = A.who_out ? (double)A.who_out[pc.src] : 0.0;
      for (int p = 0; p < A.nx; ++p) st[np + 2 + p] = A.xc_in[(long long)p * A.cap + pc.src];
    }
  }
  lsw = block_sum<RP_NT>(lsw, s_d);
  lswe = block_sum<RP_NT>(lswe, s_d);  // contains a barrier: all reads are done
  for (int q = tid; q < nfin; q += RP_NT) {
    const Piece pc = A.fin[q];
    if (pc.slot < 0) continue;
    A.w[pc.slot] = pc.v;
    if (pc.slot != pc.src) {
      const double* st = A.rstage + (long long)q * rsw;
      for (int p = 0; p < np; ++p) A.cout[p * A.cap + pc.slot] = (R)st[p];
      A.E[pc.slot] = st[np];
      if (A.who_out) A.who_out[pc.slot] = (int)st[np + 1];
      for (int p = 0; p < A.nx; ++p) A.xc_out[(long long)p * A.cap + pc.slot] = st[np + 2 + p];
    }
  }
  // the next threshold from every weight after the split (this path is the exception)
  __syncthreads();
  {
    const long long tb = th_base(A.thresh_lo);
    for (int b = tid; b < TH_BINS; b += RP_NT) s_hist[b] = 0;
    __syncthreads();
    for (int i = tid; i < n; i += RP_NT) {
      const double w = A.w[i];
      if (w >= A.thresh_lo) atomicAdd(&s_hist[th_bin(w, tb)], 1);
    }
    __syncthreads();
  }
  finish(C->sumwe[par] + lswe, C->sumw[par] + lsw, H, 1 + rcode,
         th_pick(s_hist, th_target(K), A.thresh_lo, s_i, s_sh), true);
}

// The split copies the candidate replay lists: kill slot <- source walker (coordinates,
// energy, descendant index).  Sources are never kill slots, so the copies are independent
// and spread over the whole chip.  After the global-memory replay (gc_n > 0) the copy of
// kill index t takes its family from asg[t] and its slot from klist[t], and writes the kill
// slot's weight; the threads also write the split candidates' own weights (grid-strided).
// Kill slots, split candidates and family p's slots are distinct walkers, so every weight
// written here has one writer.
template <typename R>
__global__ __launch_bounds__(TILE) void k_cont_copy(StepArgs<R> A) {
  constexpr int B = 9;  // components per batch: all loads of a batch before its stores
  const Ctrl* C = A.ctrl;
  const int t = blockIdx.x * TILE + threadIdx.x;  // < KILL_CAP: the list loads are in bounds
  if (C->status != ST_OK) return;
  const int gn = C->gc_n;
  int ks, o;
  if (gn > 0) {
    const RsgView rv = rsg_view(A.rsg);
    const int gp = C->gc_p;
    for (int j = t; j < gn; j += (int)(gridDim.x * TILE)) {  // the candidates may outnumber the threads
      if (j == gp) continue;
      const int d = rv.dep[j];
      if (d > 0) A.w[rv.ord[j]] = ldexp(key_weight(rv.key[j]), -d);
    }
    if (t >= C->ncopy) return;
    ks = A.klist[t];
    const int j = A.asg[t];
    if (j == gp) {
      o = C->gc_slot;
      if (t >= C->gc_start) A.w[ks] = ldexp(C->gc_w, -(C->gc_dp + 1));
    } else {
      o = rv.ord[j];
      A.w[ks] = ldexp(key_weight(rv.key[j]), -(int)rv.dep[j]);
    }
  } else {
    if (t >= C->ncopy) return;
    ks = A.cpy[2 * t];
    o = A.cpy[2 * t + 1];
  }
  const double e = A.E[o];
  const int who = A.who_out ? A.who_out[o] : 0;
  for (int p0 = 0; p0 < A.np; p0 += B) {
    R x[B];
#pragma unroll
    for (int q = 0; q < B; ++q)
      if (p0 + q < A.np) x[q] = A.cout[(p0 + q) * A.cap + o];
#pragma unroll
    for (int q = 0; q < B; ++q)
      if (p0 + q < A.np) A.cout[(p0 + q) * A.cap + ks] = x[q];
  }
  A.E[ks] = e;
  if (A.who_out) A.who_out[ks] = who;
  for (int q = 0; q < A.nx; ++q) A.xc_out[(long long)q * A.cap + ks] = A.xc_in[(long long)q * A.cap + o];
}

}  // namespace pvd
