// C ABI of the DMC propagation library: simulation handle, launch dispatch and the
// auxiliary kernels (layout transposes, first-step vref, descendant weights, stand-alone
// potential / trial evaluation).  Declarations and conventions: include/pyvibdmc_amd.h.
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/pyvibdmc_amd.h"
#include "pvd_kernels.hpp"

using namespace pvd;

static_assert(sizeof(pvd_pes_desc) == sizeof(PesParams), "pes layout");
static_assert(sizeof(pvd_trial_desc) == sizeof(TrialParams), "trial layout");
static_assert(PVD_MAX_COMP == MAXNP, "component bound");

namespace {

thread_local std::string g_err;

enum : int { E_BAD = 7, E_HIP = 8 };

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(x)                                                                         \
  do {                                                                                    \
    hipError_t e_ = (x);                                                                  \
    if (e_ != hipSuccess) return fail(E_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

const char* status_text(int st) {
  switch (st) {
    case ST_POP:
    case ST_WEIGHT:
      return "Massive walker birth or death event!!!!!!! Dying...";
    case ST_CAPACITY:
      return "walker capacity exceeded";
    case ST_BRANCH:
      return "continuous split outside the supported range (too many walkers below the "
             "threshold in one step, or weights below the threshold chosen for splitting)";
    case ST_SPIN:
      return "a bounded device wait expired";
    default:
      return "ok";
  }
}

// ------------------------------------------------------------------ auxiliary kernels
template <typename R>
__global__ void k_aos_to_planes(const double* aos, R* planes, long long n, int np, long long cap) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int p = 0; p < np; ++p) planes[p * cap + i] = (R)aos[i * np + p];
}

template <typename R>
__global__ void k_planes_to_aos(const R* planes, double* aos, long long n, int np, long long cap) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int p = 0; p < np; ++p) aos[i * np + p] = (double)planes[p * cap + i];
}

template <typename R>
__global__ void k_copy_planes(const R* src, R* dst, long long n, int np, long long cap) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int p = 0; p < np; ++p) dst[p * cap + i] = src[p * cap + i];
}

__global__ void k_iota(int* who, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) who[i] = (int)i;
}

__global__ void k_desc_count(const int* who, long long n, int* cnt) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) atomicAdd(&cnt[who[i]], 1);
}

// continuous descendant weights: single-member groups copy their weight exactly; larger
// groups are summed in 2^-44 fixed point with integer atomics (order independent).
constexpr double FIX_SCALE = 17592186044416.0;  // 2^44
__global__ void k_desc_cont(const int* who, const double* w, long long n, const int* cnt,
                            double* desc, unsigned long long* acc) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int q = who[i];
  if (cnt[q] == 1) {
    desc[q] = w[i];
  } else {
    const long long f = llrint(w[i] * FIX_SCALE);
    atomicAdd(&acc[q], (unsigned long long)f);
  }
}

__global__ void k_desc_finish(const int* cnt, const unsigned long long* acc, double* desc,
                              long long np, int continuous) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= np) return;
  const int c = cnt[q];
  if (!continuous) {
    desc[q] = (double)c;
  } else if (c == 0) {
    desc[q] = 0.0;
  } else if (c > 1) {
    desc[q] = (double)(long long)acc[q] / FIX_SCALE;
  }
}

// first-step energies and vref (pyvibdmc.py:760-769, 651-661)
template <typename R, int NP, bool IS, bool EXT>
__global__ __launch_bounds__(TILE) void k_start(StepArgs<R> A, int continuous) {
  __shared__ double s_d[TILE / PVD_WAVE];
  __shared__ unsigned s_last;
  const int par = A.par;
  Ctrl* C = A.ctrl;
  const int n = C->n[par];
  const int ntiles = (n + TILE - 1) / TILE;
  const int tile = blockIdx.x;
  if (tile >= ntiles) return;
  const long long i = (long long)tile * TILE + threadIdx.x;
  double E = 0.0, w = 0.0;
  if (i < n) {
    if constexpr (EXT) {
      E = A.V_ext[i];
    }
    if constexpr (!EXT || IS) {
      R x[NP];
      load_walker<R, NP>(A, i, x);
      if constexpr (!EXT) E = (double)pes_eval<R, NP>(A.pes, x);
      if constexpr (IS) {
        TrialOut<R, NP> t;
        trial_eval<R, NP>(A.trial, x, t);
        const R ke = local_kin<R, NP>(A, t.d2);
        if constexpr (EXT) {
          E = E + (double)ke;
        } else {
          E = (double)((R)E + ke);
        }
      }
    }
    w = continuous ? A.w[i] : 1.0;
  }
  const double swe = block_sum<TILE>(w * E, s_d);
  const double sw = block_sum<TILE>(w, s_d);
  if (threadIdx.x == 0) {
    A.rec[tile].s0 = swe;
    A.rec[tile].s1 = sw;
    fence_release();
    s_last = fetch_add_u(&C->done[par], 1u);
  }
  __syncthreads();
  if (s_last != (unsigned)(ntiles - 1)) return;
  fence_acquire();
  double a0 = 0.0, a1 = 0.0;
  for (int t = threadIdx.x; t < ntiles; t += TILE) {
    a0 += A.rec[t].s0;
    a1 += A.rec[t].s1;
  }
  a0 = block_sum<TILE>(a0, s_d);
  a1 = block_sum<TILE>(a1, s_d);
  if (threadIdx.x == 0) {
    C->done[par] = 0u;
    if (!continuous) {
      C->vref[par] = a0 / (double)n - (A.alpha * ((double)(n - (int)A.n0) / A.n0));
    } else {
      C->vref[par] = a0 / a1 - (A.alpha * ((a1 - A.n0) / A.n0));
    }
  }
}

// stand-alone kernels on AoS walkers
template <typename R, int NP>
__global__ void k_pes_eval(PesParams P, const R* x, long long n, R* v) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  R xx[NP];
#pragma unroll
  for (int p = 0; p < NP; ++p) xx[p] = x[i * NP + p];
  v[i] = pes_eval<R, NP>(P, xx);
}

template <typename R, int NP>
__global__ void k_pes_vjp(PesParams P, const R* x, long long n, const R* go, R* gx) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  R xx[NP], g[NP];
#pragma unroll
  for (int p = 0; p < NP; ++p) xx[p] = x[i * NP + p];
  pes_grad<R, NP>(P, xx, g);
  const R s = go[i];
#pragma unroll
  for (int p = 0; p < NP; ++p) gx[i * NP + p] = s * g[p];
}

template <typename R, int NP>
__global__ void k_trial_eval(TrialParams T, const R* x, long long n, R* psi, R* lnpsi, R* d1, R* d2) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  R xx[NP];
#pragma unroll
  for (int p = 0; p < NP; ++p) xx[p] = x[i * NP + p];
  TrialOut<R, NP> o;
  trial_eval<R, NP>(T, xx, o);
  if (psi) psi[i] = o.psi;
  if (lnpsi) lnpsi[i] = o.lnpsi;
  if (d1)
#pragma unroll
    for (int p = 0; p < NP; ++p) d1[i * NP + p] = o.d1[p];
  if (d2)
#pragma unroll
    for (int p = 0; p < NP; ++p) d2[i * NP + p] = o.d2[p];
}

template <typename R, int NP>
__global__ void k_trial_vjp(TrialParams T, const R* x, long long n, const R* go, R* gx) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  R xx[NP];
#pragma unroll
  for (int p = 0; p < NP; ++p) xx[p] = x[i * NP + p];
  TrialOut<R, NP> o;
  trial_eval<R, NP>(T, xx, o);
  const R s = go[i];
#pragma unroll
  for (int p = 0; p < NP; ++p) gx[i * NP + p] = s * o.d1[p];
}

unsigned nblk(long long n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace

// ------------------------------------------------------------------ simulation state
struct pvd_sim {
  pvd_config cfg;
  int np, M, d, prec, cont, is;
  int mspec;  // ModelSpec of the fused step kernels (0 = runtime model parameters)
  long long cap;
  void* coords[2];
  int* who[2];
  bool who_on;
  double* w;
  double* E;
  double* ke;
  unsigned long long* kmask;
  Ctrl* ctrl;
  uint64_t* flags;
  uint64_t* gflags;
  TileRec* rec;
  uint64_t* drec;
  unsigned* accs;
  double* tr_vref;
  double* tr_pop;
  double* tr_dteff;
  int* tr_births;
  int* tr_deaths;
  long long tmax;
  Piece* pend;
  Piece* fin;
  int* klist;
  unsigned long long* killtag;
  double* rstage;
  Cand* cand;
  int* asg;
  int* cpy;
  double* stage;
  void* parent;
  double* parent_w;
  int n_parent;
  double* desc;
  int* cnt;
  unsigned long long* dacc;
  long long next_step;
  unsigned epoch_next;
  int status_seen;
  int use_graphs;
  int cus;                      // compute units of the device
  const void* slot_fn[16];      // resident-workgroup counts of the discrete kernels used so far
  int slot_val[16];
  int n_slot;
  hipGraphExec_t graph[2];  // replayed step graphs, without / with descendant tracking
  // surrogate potential (external-potential simulations only)
  int surr_on;
  pvd_mlp surr;             // network description; its device parameters belong to the caller
  pvd_pes_desc surr_fb;     // analytic fallback and judge potential
  int judge_every, judge_stride;
  double* vsur;             // surrogate V of the current walkers (cap)
  double* jlog;             // judge records, 4 doubles each
  int jcap, jn;
};

namespace {

size_t esize(const pvd_sim* s) { return s->prec == PVD_FP64 ? 8 : 4; }

// buffer holding the current walkers (discrete: by step parity; continuous: in place)
int cur_buf(const pvd_sim* s) { return s->cont ? 0 : (int)(s->next_step & 1); }

template <typename R>
StepArgs<R> make_args(pvd_sim* s) {
  StepArgs<R> A;
  std::memset(&A, 0, sizeof(A));
  const int par = (int)(s->next_step & 1);
  const int cb = cur_buf(s);
  A.cin = (const R*)s->coords[cb];
  A.cout = (R*)(s->cont ? s->coords[0] : s->coords[cb ^ 1]);
  A.cap = s->cap;
  A.np = s->np;
  if (s->who_on) {
    A.who_in = s->who[cb];
    A.who_out = s->cont ? s->who[0] : s->who[cb ^ 1];
  }
  A.w = s->w;
  A.E = s->E;
  A.ke = s->ke;
  A.kmask = s->kmask;
  A.ctrl = s->ctrl;
  A.flags = s->flags;
  A.gflags = s->gflags;
  A.rec = s->rec;
  A.drec = s->drec;
  A.accs = s->accs;
  A.tr_vref = s->tr_vref;
  A.tr_pop = s->tr_pop;
  A.tr_births = s->tr_births;
  A.tr_deaths = s->tr_deaths;
  A.tr_dteff = s->tr_dteff;
  A.tmax = s->tmax;
  A.pend = s->pend;
  A.fin = s->fin;
  A.klist = s->klist;
  A.killtag = s->killtag;
  A.rstage = s->rstage;
  A.cand = s->cand;
  A.asg = s->asg;
  A.cpy = s->cpy;
  std::memcpy(&A.pes, &s->cfg.pes, sizeof(PesParams));
  std::memcpy(&A.trial, &s->cfg.trial, sizeof(TrialParams));
  for (int p = 0; p < MAXNP; ++p) {
    A.sigma[p] = s->cfg.sigma[p];
    A.inv_m[p] = s->cfg.inv_mass[p];
    A.h2s[p] = s->cfg.sigma[p] > 0.0 ? (float)(0.5 / (s->cfg.sigma[p] * s->cfg.sigma[p])) : 0.0f;
  }
  A.dt = s->cfg.delta_t;
  A.alpha = s->cfg.alpha;
  A.n0 = (double)s->cfg.num_walkers;
  A.pop_lo = (double)s->cfg.num_walkers - (double)s->cfg.num_walkers * 0.5;
  A.pop_hi = (double)s->cfg.num_walkers + (double)s->cfg.num_walkers * 0.5;
  A.thresh_lo = s->cfg.thresh_lower;
  A.rng.k0 = (uint32_t)s->cfg.seed;
  A.rng.k1 = (uint32_t)(s->cfg.seed >> 32);
  A.par = par;
  const int be = s->cfg.branch_every > 0 ? s->cfg.branch_every : 1;
  A.branch = (s->next_step % be) == 0 ? 1 : 0;
  A.ms = s->M;
  A.ds = s->d;
  A.is_ = s->is;
  return A;
}

// The compiled model specialisation for a configuration: MS_PES_WATER for the [H, H, O]
// Morse+bend surface in 3 dimensions, plus MS_TRIAL_WATER when importance sampling uses the
// matching two-bond / one-angle Gaussian trial; MS_PES_CLUSTER / MS_TRIAL_CLUSTER for the
// [O, H, H, O, H, H] cluster surface and its per-monomer trial.  Other layouts keep the
// runtime parameters.
int model_spec(const pvd_config& c, int np) {
  const pvd_pes_desc& p = c.pes;
  const pvd_trial_desc& t = c.trial;
  if (np == 18 && c.n_atoms == 6 && c.n_dim == 3) {
    const bool pc = p.kind == PES_CLUSTER6;
    const bool tc = t.kind == TRIAL_GAUSS && t.nb == 4 && t.na == 2 && t.bi[0] == 1 && t.bj[0] == 0 &&
                    t.bi[1] == 2 && t.bj[1] == 0 && t.bi[2] == 4 && t.bj[2] == 3 && t.bi[3] == 5 && t.bj[3] == 3 &&
                    t.ai[0] == 1 && t.aj[0] == 0 && t.ak[0] == 2 && t.ai[1] == 4 && t.aj[1] == 3 && t.ak[1] == 5;
    return (pc ? MS_PES_CLUSTER : MS_ANY) | (tc ? MS_TRIAL_CLUSTER : MS_ANY);
  }
  if (np != 9 || c.n_atoms != 3 || c.n_dim != 3) return MS_ANY;
  const bool pw = p.kind == PES_MORSE_BEND && p.idx[0] == 0 && p.idx[1] == 1 && p.idx[2] == 2;
  const bool tw = t.kind == TRIAL_GAUSS && t.nb == 2 && t.na == 1 && t.bi[0] == 0 && t.bj[0] == 2 &&
                  t.bi[1] == 1 && t.bj[1] == 2 && t.ai[0] == 0 && t.aj[0] == 2 && t.ak[0] == 1;
  return (pw ? MS_PES_WATER : MS_ANY) | (tw ? MS_TRIAL_WATER : MS_ANY);
}

// ---- launch helpers, instantiated over the component count ----------------------------
#define PVD_NP_SWITCH(np, ...)                                  \
  switch (np) {                                                 \
    case 1: { constexpr int NP = 1; __VA_ARGS__; } break;       \
    case 3: { constexpr int NP = 3; __VA_ARGS__; } break;       \
    case 9: { constexpr int NP = 9; __VA_ARGS__; } break;       \
    case 18: { constexpr int NP = 18; __VA_ARGS__; } break;     \
    default: return fail(E_BAD, "fused kernels support 1, 3, 9 or 18 components per walker"); \
  }

// component count and model specialisation (the specialisations exist for 9 and 18
// components)
#define PVD_STEP_SWITCH(sim, ...)                                               \
  PVD_NP_SWITCH((sim)->np, if constexpr (NP == 18) {                            \
    if ((sim)->mspec == (MS_PES_CLUSTER | MS_TRIAL_CLUSTER)) {                  \
      constexpr int MS = MS_PES_CLUSTER | MS_TRIAL_CLUSTER;                     \
      __VA_ARGS__;                                                              \
    } else if ((sim)->mspec == MS_TRIAL_CLUSTER) {                              \
      constexpr int MS = MS_TRIAL_CLUSTER;                                      \
      __VA_ARGS__;                                                              \
    } else if ((sim)->mspec == MS_PES_CLUSTER) {                                \
      constexpr int MS = MS_PES_CLUSTER;                                        \
      __VA_ARGS__;                                                              \
    } else {                                                                    \
      constexpr int MS = MS_ANY;                                                \
      __VA_ARGS__;                                                              \
    }                                                                           \
  } else if constexpr (NP == 9) {                                               \
    if ((sim)->mspec == (MS_PES_WATER | MS_TRIAL_WATER)) {                      \
      constexpr int MS = MS_PES_WATER | MS_TRIAL_WATER;                         \
      __VA_ARGS__;                                                              \
    } else if ((sim)->mspec == MS_TRIAL_WATER) {                                \
      constexpr int MS = MS_TRIAL_WATER;                                        \
      __VA_ARGS__;                                                              \
    } else if ((sim)->mspec == MS_PES_WATER) {                                  \
      constexpr int MS = MS_PES_WATER;                                          \
      __VA_ARGS__;                                                              \
    } else {                                                                    \
      constexpr int MS = MS_ANY;                                                \
      __VA_ARGS__;                                                              \
    }                                                                           \
  } else {                                                                      \
    constexpr int MS = MS_ANY;                                                  \
    __VA_ARGS__;                                                                \
  })

// The kernels of one fused step, in launch order, for the current parity.  Used both for
// direct launches and as the node list of a replayed graph.
struct Launch {
  const void* fn;
  dim3 grid, block;
};

// Resident workgroups of a discrete step kernel on this device (the tiling unit), cached.
int disc_slots(pvd_sim* s, const void* fn) {
  for (int i = 0; i < s->n_slot; ++i)
    if (s->slot_fn[i] == fn) return s->slot_val[i];
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, TILE, 0) != hipSuccess || per_cu < 1) per_cu = 1;
  const int v = per_cu * std::max(1, s->cus);
  if (s->n_slot < 16) {
    s->slot_fn[s->n_slot] = fn;
    s->slot_val[s->n_slot] = v;
    ++s->n_slot;
  }
  return v;
}

// Grid of a discrete step kernel: enough workgroups for the tiling of any population up to
// the capacity (discrete_tiles); workgroups past the tile count exit at once.
unsigned disc_grid(const pvd_sim* s, int slots, int W) {
  const long long per_round = (long long)slots * TILE * W;
  const long long rounds = (s->cap + per_round - 1) / per_round;
  return (unsigned)std::max(1ll, rounds * slots);
}

template <typename R, bool INJ>
int step_kernels(pvd_sim* s, StepArgs<R>& A, Launch* L, int& nl) {
  const unsigned g = nblk(s->cap, TILE);
  nl = 0;
  const int pk = s->cfg.pes.kind;
  if (pk == PES_EXTERNAL) return fail(E_BAD, "fused step needs a device potential");
  auto add = [&](const void* fn, unsigned grid, unsigned block) { L[nl++] = Launch{fn, dim3(grid), dim3(block)}; };
#define PVD_ADD(...) PVD_STEP_SWITCH(s, add((const void*)(__VA_ARGS__), g, TILE))
  if (s->is) PVD_ADD(k_is_move<R, NP, INJ, MS>);
  if (!s->cont) {
    auto add_disc = [&](const void* fn, int W) {
      A.slots = disc_slots(s, fn);
      add(fn, disc_grid(s, A.slots, W), TILE);
    };
    if (s->is) {
      PVD_STEP_SWITCH(s, add_disc((const void*)k_discrete<R, NP, INJ, true, false, MS>, disc_wpt<R, NP, false, true>()));
    } else {
      PVD_STEP_SWITCH(s, add_disc((const void*)k_discrete<R, NP, INJ, false, false, (MS & MS_PES_ANY_SPEC)>,
                                  disc_wpt<R, NP, false>()));
    }
  } else {
    if (s->is) {
      PVD_ADD(k_cont<R, NP, INJ, true, false, MS>);
    } else {
      PVD_ADD(k_cont<R, NP, INJ, false, false, (MS & MS_PES_ANY_SPEC)>);
    }
    add((const void*)k_cont_replay<R>, 1, RP_NT);
    add((const void*)k_cont_copy<R>, KILL_CAP / TILE, TILE);
  }
#undef PVD_ADD
  return 0;
}

template <typename R, bool INJ>
int launch_fused_step(pvd_sim* s, StepArgs<R>& A, hipStream_t st) {
  Launch L[8];
  int nl = 0;
  int rc = step_kernels<R, INJ>(s, A, L, nl);
  if (rc) return rc;
  void* args[] = {&A};
  for (int i = 0; i < nl; ++i) HIPCHK(hipLaunchKernel(L[i].fn, L[i].grid, L[i].block, args, 0, st));
  return 0;
}

// A graph of PVD_GRAPH_STEPS consecutive steps starting at an even step: the kernels are
// parameter-identical from one even/odd pair to the next (all step-dependent state is read
// from the device control block), so one instantiated graph replays any number of times.
constexpr int PVD_GRAPH_STEPS = 64;

template <typename R>
int build_graph(pvd_sim* s, hipGraphExec_t* out) {
  hipGraph_t g;
  HIPCHK(hipGraphCreate(&g, 0));
  const long long saved = s->next_step;
  const long long base = saved & ~1ll;
  hipGraphNode_t prev = nullptr;
  int rc = 0;
  for (int k = 0; k < PVD_GRAPH_STEPS && rc == 0; ++k) {
    s->next_step = base + k;
    StepArgs<R> A = make_args<R>(s);
    Launch L[8];
    int nl = 0;
    rc = step_kernels<R, false>(s, A, L, nl);
    for (int i = 0; i < nl && rc == 0; ++i) {
      hipKernelNodeParams p;
      std::memset(&p, 0, sizeof(p));
      void* args[] = {&A};
      p.func = const_cast<void*>(L[i].fn);
      p.gridDim = L[i].grid;
      p.blockDim = L[i].block;
      p.sharedMemBytes = 0;
      p.kernelParams = args;
      p.extra = nullptr;
      hipGraphNode_t node;
      hipError_t e = hipGraphAddKernelNode(&node, g, prev ? &prev : nullptr, prev ? 1 : 0, &p);
      if (e != hipSuccess) rc = fail(E_HIP, std::string("hipGraphAddKernelNode: ") + hipGetErrorString(e));
      prev = node;
    }
  }
  s->next_step = saved;
  if (rc == 0) {
    hipError_t e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    if (e != hipSuccess) rc = fail(E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  }
  (void)hipGraphDestroy(g);
  return rc;
}

template <typename R>
int do_move(pvd_sim* s, const double* disp, const double* um, hipStream_t st);
template <typename R>
int do_branch_ext(pvd_sim* s, const double* v, const double* ub, hipStream_t st);

// One step with the surrogate potential: move in place, network energy (+ analytic fallback)
// on the device planes, judge on judge steps, branch.
template <typename R>
int surrogate_step(pvd_sim* s, hipStream_t st) {
  int rc = do_move<R>(s, nullptr, nullptr, st);
  if (rc) return rc;
  const int cb = cur_buf(s);
  const int32_t* n_dev = &s->ctrl->n[(int)(s->next_step & 1)];
  const int prec = s->prec == PVD_FP64 ? PVD_FP64 : PVD_FP32;
  rc = pvd_mlp_energy(&s->surr, prec, s->M, s->coords[cb], s->cap, n_dev, s->cap, s->vsur, &s->surr_fb, st);
  if (rc) return fail(E_BAD, std::string("surrogate energy: ") + pvd_mlp_last_error());
  if (s->judge_every > 0 && (s->next_step + 1) % s->judge_every == 0 && s->jn < s->jcap) {
    rc = pvd_mlp_judge(&s->surr_fb, prec, s->M, s->coords[cb], s->cap, n_dev, s->cap, s->vsur, s->judge_stride,
                       s->next_step, s->jlog + 4ll * s->jn, st);
    if (rc) return fail(E_BAD, std::string("surrogate judge: ") + pvd_mlp_last_error());
    ++s->jn;
  }
  return do_branch_ext<R>(s, s->vsur, nullptr, st);
}

template <typename R>
int run_steps(pvd_sim* s, int n_steps, const double* disp, const double* um, const double* ub,
              hipStream_t st) {
  if (s->surr_on) {
    if (disp) return fail(E_BAD, "injected draws with a surrogate potential: use move / branch_external");
    for (int k = 0; k < n_steps; ++k) {
      const int rc = surrogate_step<R>(s, st);
      if (rc) return rc;
    }
    return 0;
  }
  int k = 0;
  const bool graphs = !disp && s->use_graphs && (s->cfg.branch_every <= 1) && n_steps >= PVD_GRAPH_STEPS + 1;
  if (graphs) {
    if (s->next_step & 1) {
      StepArgs<R> A = make_args<R>(s);
      int rc = launch_fused_step<R, false>(s, A, st);
      if (rc) return rc;
      s->next_step += 1;
      s->epoch_next += 1;
      ++k;
    }
    const int slot = s->who_on ? 1 : 0;
    if (!s->graph[slot]) {
      int rc = build_graph<R>(s, &s->graph[slot]);
      if (rc) return rc;
    }
    for (; n_steps - k >= PVD_GRAPH_STEPS; k += PVD_GRAPH_STEPS) {
      HIPCHK(hipGraphLaunch(s->graph[slot], st));
      s->next_step += PVD_GRAPH_STEPS;
      s->epoch_next += PVD_GRAPH_STEPS;
    }
  }
  for (; k < n_steps; ++k) {
    StepArgs<R> A = make_args<R>(s);
    int rc;
    if (disp) {
      A.inj_disp = disp;
      A.inj_umet = um;
      A.inj_ubr = ub;
      rc = launch_fused_step<R, true>(s, A, st);
    } else {
      rc = launch_fused_step<R, false>(s, A, st);
    }
    if (rc) return rc;
    s->next_step += 1;
    s->epoch_next += 1;
  }
  return 0;
}

// Eager steps with an event pair on every kernel dispatch (mean duration per kernel slot).
template <typename R>
int run_steps_timed(pvd_sim* s, int n_steps, hipStream_t st, float* avg_ms, int32_t* n_kernels) {
  Launch L[8];
  int nl = 0;
  StepArgs<R> A0 = make_args<R>(s);
  int rc = step_kernels<R, false>(s, A0, L, nl);
  if (rc) return rc;
  std::vector<hipEvent_t> ev((size_t)2 * nl * n_steps, nullptr);
  auto release = [&]() {
    for (hipEvent_t e : ev)
      if (e) (void)hipEventDestroy(e);
  };
  for (hipEvent_t& e : ev) {
    if (hipEventCreate(&e) != hipSuccess) {
      release();
      return fail(E_HIP, "hipEventCreate failed");
    }
  }
  for (int k = 0; k < n_steps; ++k) {
    StepArgs<R> A = make_args<R>(s);
    A.slots = A0.slots;
    void* args[] = {&A};
    for (int i = 0; i < nl; ++i) {
      // start / stop events recorded by the dispatch itself (kernel execution only)
      hipEvent_t a = ev[(size_t)2 * (k * nl + i)], b = ev[(size_t)2 * (k * nl + i) + 1];
      hipError_t e = hipExtLaunchKernel(L[i].fn, L[i].grid, L[i].block, args, 0, st, a, b, 0);
      if (e != hipSuccess) {
        release();
        return fail(E_HIP, std::string("timed launch: ") + hipGetErrorString(e));
      }
    }
    s->next_step += 1;
    s->epoch_next += 1;
  }
  hipError_t e = hipStreamSynchronize(st);
  for (int i = 0; i < nl && e == hipSuccess; ++i) {
    double sum = 0.0;
    for (int k = 0; k < n_steps && e == hipSuccess; ++k) {
      float ms = 0.f;
      e = hipEventElapsedTime(&ms, ev[(size_t)2 * (k * nl + i)], ev[(size_t)2 * (k * nl + i) + 1]);
      sum += ms;
    }
    avg_ms[i] = n_steps > 0 ? (float)(sum / n_steps) : 0.f;
  }
  release();
  if (e != hipSuccess) return fail(E_HIP, std::string("timed launch events: ") + hipGetErrorString(e));
  *n_kernels = nl;
  return 0;
}

template <typename R>
int do_move(pvd_sim* s, const double* disp, const double* um, hipStream_t st) {
  StepArgs<R> A = make_args<R>(s);
  // the move happens in place in the current buffer
  A.cout = (R*)A.cin;
  A.inj_disp = disp;
  A.inj_umet = um;
  const unsigned g = nblk(s->cap, TILE);
  if (s->is) {
    if (disp) {
      PVD_STEP_SWITCH(s, hipLaunchKernelGGL((k_move<R, NP, true, true, (MS & MS_TRIAL_ANY_SPEC)>), dim3(g), dim3(TILE), 0, st, A));
    } else {
      PVD_STEP_SWITCH(s, hipLaunchKernelGGL((k_move<R, NP, false, true, (MS & MS_TRIAL_ANY_SPEC)>), dim3(g), dim3(TILE), 0, st, A));
    }
  } else {
    if (disp) hipLaunchKernelGGL((k_move<R, 1, true, false>), dim3(g), dim3(TILE), 0, st, A);
    else hipLaunchKernelGGL((k_move<R, 1, false, false>), dim3(g), dim3(TILE), 0, st, A);
  }
  HIPCHK(hipGetLastError());
  return 0;
}

template <typename R>
int do_branch_ext(pvd_sim* s, const double* v, const double* ub, hipStream_t st) {
  StepArgs<R> A = make_args<R>(s);
  A.V_ext = v;
  A.inj_ubr = ub;
  unsigned g = nblk(s->cap, TILE);
  if (!s->cont) {
    // compile-time component count where one is instantiated, else the run-time form (NP = 0)
    const void* fn = nullptr;
    auto pick = [&](auto np_tag) {
      constexpr int NP = decltype(np_tag)::value;
      fn = ub ? (s->is ? (const void*)k_discrete<R, NP, true, true, true> : (const void*)k_discrete<R, NP, true, false, true>)
              : (s->is ? (const void*)k_discrete<R, NP, false, true, true> : (const void*)k_discrete<R, NP, false, false, true>);
    };
    switch (s->np) {
      case 1: pick(std::integral_constant<int, 1>()); break;
      case 3: pick(std::integral_constant<int, 3>()); break;
      case 9: pick(std::integral_constant<int, 9>()); break;
      case 18: pick(std::integral_constant<int, 18>()); break;
      default: pick(std::integral_constant<int, 0>()); break;
    }
    A.slots = disc_slots(s, fn);
    g = disc_grid(s, A.slots, disc_wpt<R, 1, true>());
    void* args[] = {&A};
    HIPCHK(hipLaunchKernel(fn, dim3(g), dim3(TILE), args, 0, st));
  } else {
    if (s->is) hipLaunchKernelGGL((k_cont<R, 1, false, true, true>), dim3(g), dim3(TILE), 0, st, A);
    else hipLaunchKernelGGL((k_cont<R, 1, false, false, true>), dim3(g), dim3(TILE), 0, st, A);
    hipLaunchKernelGGL((k_cont_replay<R>), dim3(1), dim3(RP_NT), 0, st, A);
    hipLaunchKernelGGL((k_cont_copy<R>), dim3(KILL_CAP / TILE), dim3(TILE), 0, st, A);
  }
  HIPCHK(hipGetLastError());
  s->next_step += 1;
  s->epoch_next += 1;
  return 0;
}

template <typename R>
int do_start(pvd_sim* s, const double* v, hipStream_t st) {
  StepArgs<R> A = make_args<R>(s);
  A.V_ext = v;
  const unsigned g = nblk(s->cap, TILE);
  if (v) {
    if (s->is) {
      PVD_NP_SWITCH(s->np, hipLaunchKernelGGL((k_start<R, NP, true, true>), dim3(g), dim3(TILE), 0, st, A, s->cont));
    } else {
      hipLaunchKernelGGL((k_start<R, 1, false, true>), dim3(g), dim3(TILE), 0, st, A, s->cont);
    }
  } else {
    if (s->is) {
      PVD_NP_SWITCH(s->np, hipLaunchKernelGGL((k_start<R, NP, true, false>), dim3(g), dim3(TILE), 0, st, A, s->cont));
    } else {
      PVD_NP_SWITCH(s->np, hipLaunchKernelGGL((k_start<R, NP, false, false>), dim3(g), dim3(TILE), 0, st, A, s->cont));
    }
  }
  HIPCHK(hipGetLastError());
  return 0;
}

int read_ctrl(pvd_sim* s, Ctrl& c, hipStream_t st) {
  HIPCHK(hipMemcpyAsync(&c, s->ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  return 0;
}

template <typename R>
int pes_eval_t(const pvd_pes_desc* pes, int np, const void* x, long long n, void* v, void* g_out,
               void* gx, hipStream_t st) {
  PesParams P;
  std::memcpy(&P, pes, sizeof(P));
  if (g_out) {
    PVD_NP_SWITCH(np, hipLaunchKernelGGL((k_pes_vjp<R, NP>), dim3(nblk(n)), dim3(256), 0, st, P,
                                         (const R*)x, n, (const R*)g_out, (R*)gx));
  } else {
    PVD_NP_SWITCH(np, hipLaunchKernelGGL((k_pes_eval<R, NP>), dim3(nblk(n)), dim3(256), 0, st, P,
                                         (const R*)x, n, (R*)v));
  }
  HIPCHK(hipGetLastError());
  return 0;
}

template <typename R>
int trial_eval_t(const pvd_trial_desc* tr, int np, const void* x, long long n, void* psi, void* lnpsi,
                 void* d1, void* d2, const void* g_out, void* gx, hipStream_t st) {
  TrialParams T;
  std::memcpy(&T, tr, sizeof(T));
  if (g_out) {
    PVD_NP_SWITCH(np, hipLaunchKernelGGL((k_trial_vjp<R, NP>), dim3(nblk(n)), dim3(256), 0, st, T,
                                         (const R*)x, n, (const R*)g_out, (R*)gx));
  } else {
    PVD_NP_SWITCH(np, hipLaunchKernelGGL((k_trial_eval<R, NP>), dim3(nblk(n)), dim3(256), 0, st, T,
                                         (const R*)x, n, (R*)psi, (R*)lnpsi, (R*)d1, (R*)d2));
  }
  HIPCHK(hipGetLastError());
  return 0;
}

void free_sim(pvd_sim* s) {
  if (!s) return;
  for (int i = 0; i < 2; ++i)
    if (s->graph[i]) (void)hipGraphExecDestroy(s->graph[i]);
  void* ptrs[] = {s->coords[0], s->coords[1], s->who[0], s->who[1], s->w, s->E, s->ke, s->kmask,
                  s->ctrl, s->flags, s->gflags, s->rec, s->drec, s->accs, s->tr_vref, s->tr_pop, s->tr_dteff, s->tr_births,
                  s->tr_deaths, s->pend, s->fin, s->klist, s->killtag, s->rstage, s->cand, s->asg, s->cpy, s->stage, s->parent, s->parent_w, s->desc,
                  s->cnt, s->dacc, s->vsur, s->jlog};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  delete s;
}

}  // namespace

// ===================================================================== C entry points
extern "C" {

const char* pvd_version(void) { return "pyvibdmc_amd 0.1 (gfx950)"; }

const char* pvd_last_error(void) { return g_err.c_str(); }

int pvd_sim_create(const pvd_config* cfg, pvd_sim** out) {
  if (!cfg || !out) return fail(E_BAD, "null argument");
  *out = nullptr;
  const int np = cfg->n_atoms * cfg->n_dim;
  if (cfg->num_walkers <= 0 || cfg->n_atoms <= 0 || cfg->n_dim <= 0 || np > PVD_MAX_COMP)
    return fail(E_BAD, "bad walker shape");
  if (cfg->precision != PVD_FP64 && cfg->precision != PVD_FP32) return fail(E_BAD, "bad precision");
  if (cfg->weighting != PVD_DISCRETE && cfg->weighting != PVD_CONTINUOUS)
    return fail(E_BAD, "bad weighting");
  if (cfg->trial.kind != TRIAL_NONE && np > PVD_FUSED_MAX_COMP)
    return fail(E_BAD, "importance sampling supports at most 18 components per walker");
  pvd_sim* s = new pvd_sim();
  std::memset((void*)s, 0, sizeof(pvd_sim));
  s->cfg = *cfg;
  s->np = np;
  s->M = cfg->n_atoms;
  s->d = cfg->n_dim;
  s->prec = cfg->precision;
  s->cont = cfg->weighting == PVD_CONTINUOUS;
  s->is = cfg->trial.kind != TRIAL_NONE;
  s->mspec = model_spec(*cfg, np);
  long long cap = cfg->capacity;
  const long long need = (long long)std::floor(1.5 * cfg->num_walkers) + 1;
  if (cap <= 0) cap = need;
  cap = ((cap + TILE - 1) / TILE) * TILE;
  if (cap >= (1ll << 30)) {
    delete s;
    return fail(E_BAD, "capacity must stay below 2^30 walkers");
  }
  if (cfg->precision == PVD_FP32 && cap * (long long)np * 4 >= (1ll << 32)) {
    delete s;
    return fail(E_BAD, "FP32 walker buffers are addressed with 32-bit offsets: capacity * components * 4 "
                       "bytes must stay below 4 GiB");
  }
  s->cap = cap;
  s->tmax = cfg->num_timesteps > 0 ? cfg->num_timesteps : 1;
  {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) == hipSuccess && hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) ==
                                                 hipSuccess)
      s->cus = cus;
    if (s->cus < 1) s->cus = 1;
  }
  // tiles of the discrete kernels: at most cap / TILE plus one round of resident workgroups
  // (at most 8 per CU) - see discrete_tiles and disc_grid
  const long long ntiles = cap / TILE + 8ll * s->cus;
  const size_t es = esize(s);
#define ALLOC(ptr, bytes)                                    \
  do {                                                       \
    hipError_t e_ = hipMalloc((void**)&(ptr), (bytes));      \
    if (e_ != hipSuccess) {                                  \
      free_sim(s);                                           \
      return fail(E_HIP, std::string("hipMalloc: ") + hipGetErrorString(e_)); \
    }                                                        \
  } while (0)
  ALLOC(s->coords[0], es * np * cap);
  ALLOC(s->coords[1], es * np * cap);
  ALLOC(s->who[0], sizeof(int) * cap);
  ALLOC(s->who[1], sizeof(int) * cap);
  ALLOC(s->w, sizeof(double) * cap);
  ALLOC(s->E, sizeof(double) * cap);
  ALLOC(s->ke, sizeof(double) * cap);
  ALLOC(s->kmask, sizeof(unsigned long long) * ntiles * (TILE / PVD_WAVE));
  ALLOC(s->ctrl, sizeof(Ctrl));
  ALLOC(s->flags, sizeof(uint64_t) * ntiles);
  ALLOC(s->gflags, sizeof(uint64_t) * (ntiles / LB_GROUP + 1));
  ALLOC(s->rec, sizeof(TileRec) * ntiles);
  ALLOC(s->drec, sizeof(uint64_t) * 3 * ntiles);
  ALLOC(s->accs, sizeof(unsigned) * 2 * ACC_SLOTS * ACC_STRIDE);
  ALLOC(s->tr_vref, sizeof(double) * s->tmax);
  ALLOC(s->tr_pop, sizeof(double) * s->tmax);
  ALLOC(s->tr_dteff, sizeof(double) * s->tmax);
  ALLOC(s->tr_births, sizeof(int) * s->tmax);
  ALLOC(s->tr_deaths, sizeof(int) * s->tmax);
  if (s->cont) {
    ALLOC(s->pend, sizeof(Piece) * RP_PCAP * 2);
    ALLOC(s->fin, sizeof(Piece) * RP_FCAP);
    ALLOC(s->klist, sizeof(int) * KILL_CAP);
    ALLOC(s->cand, sizeof(Cand) * 2 * LIST_CAP);
    ALLOC(s->asg, sizeof(int) * KILL_CAP);
    ALLOC(s->cpy, sizeof(int) * 2 * KILL_CAP);
    ALLOC(s->killtag, sizeof(unsigned long long) * cap);
    ALLOC(s->rstage, sizeof(double) * (size_t)RP_FCAP * (np + 2));
  }
  ALLOC(s->stage, sizeof(double) * np * cap);
#undef ALLOC
  hipError_t e = hipMemset(s->flags, 0, sizeof(uint64_t) * ntiles);
  if (e == hipSuccess) e = hipMemset(s->gflags, 0, sizeof(uint64_t) * (ntiles / LB_GROUP + 1));
  if (e == hipSuccess) e = hipMemset(s->drec, 0, sizeof(uint64_t) * 3 * ntiles);
  if (e == hipSuccess) e = hipMemset(s->accs, 0, sizeof(unsigned) * 2 * ACC_SLOTS * ACC_STRIDE);
  if (e == hipSuccess) e = hipMemset(s->ctrl, 0, sizeof(Ctrl));
  if (e == hipSuccess) e = hipMemset(s->tr_vref, 0, sizeof(double) * s->tmax);
  if (e == hipSuccess) e = hipMemset(s->tr_pop, 0, sizeof(double) * s->tmax);
  if (e == hipSuccess) e = hipMemset(s->tr_dteff, 0, sizeof(double) * s->tmax);
  if (e == hipSuccess) e = hipMemset(s->tr_births, 0, sizeof(int) * s->tmax);
  if (e == hipSuccess) e = hipMemset(s->tr_deaths, 0, sizeof(int) * s->tmax);
  if (e == hipSuccess && s->killtag) e = hipMemset(s->killtag, 0, sizeof(unsigned long long) * cap);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) {
    free_sim(s);
    return fail(E_HIP, std::string("init: ") + hipGetErrorString(e));
  }
  s->epoch_next = 1;
  {
    const char* e = std::getenv("PVD_GRAPHS");
    s->use_graphs = !(e && e[0] == '0');
  }
  *out = s;
  return 0;
}

int pvd_sim_destroy(pvd_sim* sim) {
  if (!sim) return 0;
  (void)hipDeviceSynchronize();
  free_sim(sim);
  return 0;
}

int pvd_sim_set_walkers(pvd_sim* s, const double* coords, int32_t n, const double* weights,
                        int64_t step0, void* stream) {
  if (!s || !coords) return fail(E_BAD, "null argument");
  if (n <= 0 || n > s->cap) return fail(E_BAD, "walker count outside [1, capacity]");
  if (step0 < 0) return fail(E_BAD, "negative step");
  hipStream_t st = (hipStream_t)stream;
  s->next_step = step0;
  const int cb = cur_buf(s);
  HIPCHK(hipMemcpyAsync(s->stage, coords, sizeof(double) * (size_t)n * s->np, hipMemcpyHostToDevice, st));
  if (s->prec == PVD_FP64)
    hipLaunchKernelGGL((k_aos_to_planes<double>), dim3(nblk(n)), dim3(256), 0, st, s->stage,
                       (double*)s->coords[cb], (long long)n, s->np, s->cap);
  else
    hipLaunchKernelGGL((k_aos_to_planes<float>), dim3(nblk(n)), dim3(256), 0, st, s->stage,
                       (float*)s->coords[cb], (long long)n, s->np, s->cap);
  HIPCHK(hipGetLastError());
  if (s->cont) {
    if (weights) {
      HIPCHK(hipMemcpyAsync(s->w, weights, sizeof(double) * n, hipMemcpyHostToDevice, st));
    } else {
      return fail(E_BAD, "continuous weighting needs weights");
    }
  }
  Ctrl c;
  std::memset(&c, 0, sizeof(c));
  const int par = (int)(step0 & 1);
  c.n[par] = n;
  c.step[par] = (unsigned long long)step0;
  s->epoch_next += 2;
  c.epoch[par] = s->epoch_next;
  HIPCHK(hipMemcpyAsync(s->ctrl, &c, sizeof(c), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemsetAsync(s->accs, 0, sizeof(unsigned) * 2 * ACC_SLOTS * ACC_STRIDE, st));
  HIPCHK(hipStreamSynchronize(st));
  s->who_on = false;
  return 0;
}

int pvd_sim_get_walkers(pvd_sim* s, double* coords, double* weights, int64_t max_walkers, int32_t* n_out,
                        void* stream) {
  if (!s || !coords) return fail(E_BAD, "null argument");
  hipStream_t st = (hipStream_t)stream;
  Ctrl c;
  int rc = read_ctrl(s, c, st);
  if (rc) return rc;
  long long step = (long long)s->next_step;
  if (c.status != ST_OK) step = c.status_step;
  const int par = (int)(step & 1);
  const int n = c.n[par];
  if (n > max_walkers) return fail(E_BAD, "output buffer too small");
  const int cb = s->cont ? 0 : par;
  if (s->prec == PVD_FP64)
    hipLaunchKernelGGL((k_planes_to_aos<double>), dim3(nblk(n)), dim3(256), 0, st,
                       (const double*)s->coords[cb], s->stage, (long long)n, s->np, s->cap);
  else
    hipLaunchKernelGGL((k_planes_to_aos<float>), dim3(nblk(n)), dim3(256), 0, st,
                       (const float*)s->coords[cb], s->stage, (long long)n, s->np, s->cap);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(coords, s->stage, sizeof(double) * (size_t)n * s->np, hipMemcpyDeviceToHost, st));
  if (weights && s->cont)
    HIPCHK(hipMemcpyAsync(weights, s->w, sizeof(double) * n, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  if (n_out) *n_out = n;
  return 0;
}

int pvd_sim_start(pvd_sim* s, void* stream) {
  if (!s) return fail(E_BAD, "null argument");
  if (s->surr_on) {
    // first-step energies from the surrogate on the device
    const int cb = cur_buf(s);
    const int prec = s->prec == PVD_FP64 ? PVD_FP64 : PVD_FP32;
    const int rc = pvd_mlp_energy(&s->surr, prec, s->M, s->coords[cb], s->cap, &s->ctrl->n[(int)(s->next_step & 1)],
                                  s->cap, s->vsur, &s->surr_fb, stream);
    if (rc) return fail(E_BAD, std::string("surrogate energy: ") + pvd_mlp_last_error());
    return s->prec == PVD_FP64 ? do_start<double>(s, s->vsur, (hipStream_t)stream)
                               : do_start<float>(s, s->vsur, (hipStream_t)stream);
  }
  if (s->cfg.pes.kind == PES_EXTERNAL) return fail(E_BAD, "external potential: use pvd_sim_start_external");
  return s->prec == PVD_FP64 ? do_start<double>(s, nullptr, (hipStream_t)stream)
                             : do_start<float>(s, nullptr, (hipStream_t)stream);
}

int pvd_sim_start_external(pvd_sim* s, const double* v, void* stream) {
  if (!s || !v) return fail(E_BAD, "null argument");
  return s->prec == PVD_FP64 ? do_start<double>(s, v, (hipStream_t)stream)
                             : do_start<float>(s, v, (hipStream_t)stream);
}

int pvd_sim_run(pvd_sim* s, int32_t n_steps, void* stream) {
  if (!s) return fail(E_BAD, "null argument");
  if (n_steps < 0) return fail(E_BAD, "negative step count");
  return s->prec == PVD_FP64 ? run_steps<double>(s, n_steps, nullptr, nullptr, nullptr, (hipStream_t)stream)
                             : run_steps<float>(s, n_steps, nullptr, nullptr, nullptr, (hipStream_t)stream);
}

int pvd_sim_run_timed(pvd_sim* s, int32_t n_steps, void* stream, float* avg_ms, int32_t* n_kernels) {
  if (!s || !avg_ms || !n_kernels) return fail(E_BAD, "null argument");
  if (n_steps < 0) return fail(E_BAD, "negative step count");
  return s->prec == PVD_FP64 ? run_steps_timed<double>(s, n_steps, (hipStream_t)stream, avg_ms, n_kernels)
                             : run_steps_timed<float>(s, n_steps, (hipStream_t)stream, avg_ms, n_kernels);
}

int pvd_sim_step_injected(pvd_sim* s, const double* disp, const double* um, const double* ub, void* stream) {
  if (!s || !disp) return fail(E_BAD, "null argument");
  if (s->is && !um) return fail(E_BAD, "importance sampling needs Metropolis uniforms");
  const int be = s->cfg.branch_every > 0 ? s->cfg.branch_every : 1;
  if (!s->cont && (s->next_step % be) == 0 && !ub) return fail(E_BAD, "branching step needs uniforms");
  return s->prec == PVD_FP64 ? run_steps<double>(s, 1, disp, um, ub, (hipStream_t)stream)
                             : run_steps<float>(s, 1, disp, um, ub, (hipStream_t)stream);
}

int pvd_sim_move(pvd_sim* s, const double* disp, const double* um, void* stream) {
  if (!s) return fail(E_BAD, "null argument");
  if (s->is && disp && !um) return fail(E_BAD, "importance sampling needs Metropolis uniforms");
  return s->prec == PVD_FP64 ? do_move<double>(s, disp, um, (hipStream_t)stream)
                             : do_move<float>(s, disp, um, (hipStream_t)stream);
}

int pvd_sim_branch_external(pvd_sim* s, const double* v, const double* ub, void* stream) {
  if (!s || !v) return fail(E_BAD, "null argument");
  return s->prec == PVD_FP64 ? do_branch_ext<double>(s, v, ub, (hipStream_t)stream)
                             : do_branch_ext<float>(s, v, ub, (hipStream_t)stream);
}

int pvd_sim_planes(pvd_sim* s, void** planes, int64_t* stride, int32_t* n, void* stream) {
  if (!s) return fail(E_BAD, "null argument");
  Ctrl c;
  int rc = read_ctrl(s, c, (hipStream_t)stream);
  if (rc) return rc;
  const int cb = cur_buf(s);
  if (planes) *planes = s->coords[cb];
  if (stride) *stride = s->cap;
  if (n) *n = c.n[(int)(s->next_step & 1)];
  return 0;
}

int pvd_sim_device_view(pvd_sim* s, void** planes, int64_t* stride, int32_t** n_dev) {
  if (!s) return fail(E_BAD, "null argument");
  const int cb = cur_buf(s);
  if (planes) *planes = s->coords[cb];
  if (stride) *stride = s->cap;
  if (n_dev) *n_dev = &s->ctrl->n[(int)(s->next_step & 1)];
  return 0;
}

int pvd_sim_set_surrogate(pvd_sim* s, const pvd_mlp* net, const pvd_pes_desc* fallback, int32_t judge_every,
                          int32_t judge_stride) {
  if (!s) return fail(E_BAD, "null argument");
  if (!net) {
    s->surr_on = 0;
    return 0;
  }
  if (s->cfg.pes.kind != PES_EXTERNAL) return fail(E_BAD, "the simulation has a built-in device potential");
  if (!fallback) return fail(E_BAD, "the surrogate needs an analytic fallback potential");
  if (s->d != 3 || (s->M != 3 && s->M != 6)) return fail(E_BAD, "the surrogate supports 3 or 6 atoms in 3-D");
  if (judge_every < 0 || (judge_every > 0 && judge_stride < 1)) return fail(E_BAD, "bad judge schedule");
  if (!s->vsur) HIPCHK(hipMalloc((void**)&s->vsur, sizeof(double) * s->cap));
  if (judge_every > 0 && !s->jlog) {
    s->jcap = (int)std::min<long long>(1 << 20, s->tmax / judge_every + 2);
    HIPCHK(hipMalloc((void**)&s->jlog, sizeof(double) * 4 * s->jcap));
  }
  s->surr = *net;
  s->surr_fb = *fallback;
  s->judge_every = judge_every;
  s->judge_stride = judge_stride;
  s->jn = 0;
  s->surr_on = 1;
  return 0;
}

int pvd_sim_judge_log(pvd_sim* s, double* out, int32_t max_records, int32_t* n, void* stream) {
  if (!s || !n) return fail(E_BAD, "null argument");
  const int k = std::min(s->jn, std::max(0, (int)max_records));
  if (k > 0 && out) {
    HIPCHK(hipMemcpyAsync(out, s->jlog, sizeof(double) * 4 * k, hipMemcpyDeviceToHost, (hipStream_t)stream));
    HIPCHK(hipStreamSynchronize((hipStream_t)stream));
  }
  *n = s->jn;
  return 0;
}

int pvd_sim_status(pvd_sim* s, int32_t* status, int64_t* fail_step, int32_t* n, double* vref,
                   int64_t* next_step, void* stream) {
  if (!s) return fail(E_BAD, "null argument");
  Ctrl c;
  int rc = read_ctrl(s, c, (hipStream_t)stream);
  if (rc) return rc;
  const long long step = c.status != ST_OK ? c.status_step : (long long)s->next_step;
  const int par = (int)(step & 1);
  if (status) *status = c.status;
  if (fail_step) *fail_step = c.status != ST_OK ? c.status_step : -1;
  if (n) *n = c.n[par];
  if (vref) *vref = c.vref[par];
  if (next_step) *next_step = step;
  if (c.status != ST_OK) return fail(c.status, status_text(c.status));
  return 0;
}

int pvd_sim_trajectory(pvd_sim* s, int64_t t0, int32_t count, double* vref, double* pop, int32_t* births,
                       int32_t* deaths, double* eff_t) {
  if (!s) return fail(E_BAD, "null argument");
  if (t0 < 0 || count < 0 || t0 + count > s->tmax) return fail(E_BAD, "trajectory range");
  HIPCHK(hipDeviceSynchronize());
  if (vref) HIPCHK(hipMemcpy(vref, s->tr_vref + t0, sizeof(double) * count, hipMemcpyDeviceToHost));
  if (pop) HIPCHK(hipMemcpy(pop, s->tr_pop + t0, sizeof(double) * count, hipMemcpyDeviceToHost));
  if (births) HIPCHK(hipMemcpy(births, s->tr_births + t0, sizeof(int) * count, hipMemcpyDeviceToHost));
  if (deaths) HIPCHK(hipMemcpy(deaths, s->tr_deaths + t0, sizeof(int) * count, hipMemcpyDeviceToHost));
  if (eff_t) HIPCHK(hipMemcpy(eff_t, s->tr_dteff + t0, sizeof(double) * count, hipMemcpyDeviceToHost));
  return 0;
}

int pvd_sim_desc_begin(pvd_sim* s, void* stream) {
  if (!s) return fail(E_BAD, "null argument");
  hipStream_t st = (hipStream_t)stream;
  Ctrl c;
  int rc = read_ctrl(s, c, st);
  if (rc) return rc;
  const int n = c.n[(int)(s->next_step & 1)];
  const int cb = cur_buf(s);
  const size_t es = esize(s);
  if (!s->parent) {
    HIPCHK(hipMalloc(&s->parent, es * s->np * s->cap));
    HIPCHK(hipMalloc((void**)&s->parent_w, sizeof(double) * s->cap));
    HIPCHK(hipMalloc((void**)&s->desc, sizeof(double) * s->cap));
    HIPCHK(hipMalloc((void**)&s->cnt, sizeof(int) * s->cap));
    HIPCHK(hipMalloc((void**)&s->dacc, sizeof(unsigned long long) * s->cap));
  }
  hipLaunchKernelGGL(k_iota, dim3(nblk(n)), dim3(256), 0, st, s->who[cb], (long long)n);
  if (s->prec == PVD_FP64)
    hipLaunchKernelGGL((k_copy_planes<double>), dim3(nblk(n)), dim3(256), 0, st, (const double*)s->coords[cb],
                       (double*)s->parent, (long long)n, s->np, s->cap);
  else
    hipLaunchKernelGGL((k_copy_planes<float>), dim3(nblk(n)), dim3(256), 0, st, (const float*)s->coords[cb],
                       (float*)s->parent, (long long)n, s->np, s->cap);
  if (s->cont) HIPCHK(hipMemcpyAsync(s->parent_w, s->w, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
  HIPCHK(hipGetLastError());
  s->n_parent = n;
  s->who_on = true;
  return 0;
}

int pvd_sim_desc_end(pvd_sim* s, double* desc_wts, double* parent, double* parent_wts, int32_t* n_parent,
                     void* stream) {
  if (!s) return fail(E_BAD, "null argument");
  if (!s->who_on) return fail(E_BAD, "no descendant-weighting window is open");
  hipStream_t st = (hipStream_t)stream;
  Ctrl c;
  int rc = read_ctrl(s, c, st);
  if (rc) return rc;
  const int n = c.n[(int)(s->next_step & 1)];
  const int cb = cur_buf(s);
  const long long npar = s->n_parent;
  HIPCHK(hipMemsetAsync(s->cnt, 0, sizeof(int) * npar, st));
  HIPCHK(hipMemsetAsync(s->dacc, 0, sizeof(unsigned long long) * npar, st));
  hipLaunchKernelGGL(k_desc_count, dim3(nblk(n)), dim3(256), 0, st, s->who[cb], (long long)n, s->cnt);
  if (s->cont)
    hipLaunchKernelGGL(k_desc_cont, dim3(nblk(n)), dim3(256), 0, st, s->who[cb], s->w, (long long)n, s->cnt,
                       s->desc, s->dacc);
  hipLaunchKernelGGL(k_desc_finish, dim3(nblk(npar)), dim3(256), 0, st, s->cnt, s->dacc, s->desc, npar, s->cont);
  HIPCHK(hipGetLastError());
  if (desc_wts) HIPCHK(hipMemcpyAsync(desc_wts, s->desc, sizeof(double) * npar, hipMemcpyDeviceToHost, st));
  if (parent) {
    if (s->prec == PVD_FP64)
      hipLaunchKernelGGL((k_planes_to_aos<double>), dim3(nblk(npar)), dim3(256), 0, st, (const double*)s->parent,
                         s->stage, npar, s->np, s->cap);
    else
      hipLaunchKernelGGL((k_planes_to_aos<float>), dim3(nblk(npar)), dim3(256), 0, st, (const float*)s->parent,
                         s->stage, npar, s->np, s->cap);
    HIPCHK(hipMemcpyAsync(parent, s->stage, sizeof(double) * npar * s->np, hipMemcpyDeviceToHost, st));
  }
  if (parent_wts && s->cont)
    HIPCHK(hipMemcpyAsync(parent_wts, s->parent_w, sizeof(double) * npar, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  if (n_parent) *n_parent = (int)npar;
  s->who_on = false;
  return 0;
}

int pvd_pes_eval(const pvd_pes_desc* pes, int32_t precision, int32_t n_atoms, int32_t n_dim, const void* coords,
                 int64_t n, void* v_out, void* stream) {
  if (!pes || !coords || !v_out) return fail(E_BAD, "null argument");
  if (n <= 0) return 0;
  const int np = n_atoms * n_dim;
  return precision == PVD_FP64 ? pes_eval_t<double>(pes, np, coords, n, v_out, nullptr, nullptr, (hipStream_t)stream)
                               : pes_eval_t<float>(pes, np, coords, n, v_out, nullptr, nullptr, (hipStream_t)stream);
}

int pvd_pes_vjp(const pvd_pes_desc* pes, int32_t precision, int32_t n_atoms, int32_t n_dim, const void* coords,
                int64_t n, const void* grad_out, void* grad_coords, void* stream) {
  if (!pes || !coords || !grad_out || !grad_coords) return fail(E_BAD, "null argument");
  if (n <= 0) return 0;
  const int np = n_atoms * n_dim;
  return precision == PVD_FP64
             ? pes_eval_t<double>(pes, np, coords, n, nullptr, (void*)grad_out, grad_coords, (hipStream_t)stream)
             : pes_eval_t<float>(pes, np, coords, n, nullptr, (void*)grad_out, grad_coords, (hipStream_t)stream);
}

int pvd_trial_eval(const pvd_trial_desc* trial, int32_t precision, int32_t n_atoms, int32_t n_dim,
                   const void* coords, int64_t n, void* psi, void* lnpsi, void* dlnpsi, void* d2, void* stream) {
  if (!trial || !coords) return fail(E_BAD, "null argument");
  if (n <= 0) return 0;
  const int np = n_atoms * n_dim;
  return precision == PVD_FP64 ? trial_eval_t<double>(trial, np, coords, n, psi, lnpsi, dlnpsi, d2, nullptr,
                                                      nullptr, (hipStream_t)stream)
                               : trial_eval_t<float>(trial, np, coords, n, psi, lnpsi, dlnpsi, d2, nullptr,
                                                     nullptr, (hipStream_t)stream);
}

int pvd_trial_vjp(const pvd_trial_desc* trial, int32_t precision, int32_t n_atoms, int32_t n_dim,
                  const void* coords, int64_t n, const void* grad_out, void* grad_coords, void* stream) {
  if (!trial || !coords || !grad_out || !grad_coords) return fail(E_BAD, "null argument");
  if (n <= 0) return 0;
  const int np = n_atoms * n_dim;
  return precision == PVD_FP64 ? trial_eval_t<double>(trial, np, coords, n, nullptr, nullptr, nullptr, nullptr,
                                                      grad_out, grad_coords, (hipStream_t)stream)
                               : trial_eval_t<float>(trial, np, coords, n, nullptr, nullptr, nullptr, nullptr,
                                                     grad_out, grad_coords, (hipStream_t)stream);
}

}  // extern "C"
