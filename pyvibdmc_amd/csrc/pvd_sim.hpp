// Simulation handle of the C ABI and the host-side pieces shared by the translation units
// of the library: the handle itself, launch arguments of the step kernels, error
// reporting, and the per-precision step operations (defined in pvd_sim_step.hpp and
// instantiated once per element type in pvd_step_f32.hip / pvd_step_f64.hip, so the two
// precisions compile in parallel).
#pragma once
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

namespace pvd {

enum : int { E_BAD = 7, E_HIP = 8 };

// records `msg` as the thread's last error (pvd_last_error) and returns `code`
int fail(int code, const std::string& msg);

#define HIPCHK(x)                                                                         \
  do {                                                                                    \
    hipError_t e_ = (x);                                                                  \
    if (e_ != hipSuccess) return fail(E_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

inline unsigned nblk(long long n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace pvd

// ------------------------------------------------------------------ simulation state
// ------------------------------------------------------------------ simulation state
struct pvd_sim {
  pvd_config cfg;
  int np, M, d, prec, cont, is;
  int htrial;      // trial wave function evaluated by the caller
  int nx;          // its cache planes (2 np + 1), 0 without one
  double* xc[2];   // the cache planes, by step parity like the coordinates
  int mspec;  // ModelSpec of the fused step kernels (0 = runtime model parameters)
  long long cap;
  void* coords[2];
  int* who[2];
  bool who_on;
  double* w;
  double* E;
  double* ke;
  unsigned long long* kmask;
  pvd::Ctrl* ctrl;
  uint64_t* flags;
  uint64_t* gflags;
  pvd::TileRec* rec;
  uint64_t* drec;
  unsigned* accs;
  double* tr_vref;
  double* tr_pop;
  double* tr_dteff;
  int* tr_births;
  int* tr_deaths;
  long long tmax;
  pvd::Piece* pend;
  pvd::Piece* fin;
  int* klist;
  unsigned long long* killtag;
  double* rstage;
  pvd::Cand* cand;
  unsigned char* rsg;  // global-memory replay scratch (continuous weighting)
  unsigned* ncnt;      // candidate sub-list counts (continuous weighting)
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
  // surrogate steps replayed as graphs of 2^k steps (k = 1..6), without / with descendant
  // tracking; captured on cap_stream and dropped whenever the network description changes
  hipGraphExec_t sgraph[2][7];
  hipStream_t cap_stream;
  // surrogate potential (external-potential simulations only)
  int surr_on;
  pvd_mlp surr;             // network description; its device parameters belong to the caller
  pvd_pes_desc surr_fb;     // analytic fallback and judge potential
  int judge_every, judge_stride;
  double* vsur;             // surrogate V of the current walkers (cap)
  double* jlog;             // judge records, 4 doubles each
  int jcap, jn;
};

namespace pvd {

inline size_t esize(const pvd_sim* s) { return s->prec == PVD_FP64 ? 8 : 4; }

// buffer holding the current walkers (discrete: by step parity; continuous: in place)
inline int cur_buf(const pvd_sim* s) { return s->cont ? 0 : (int)(s->next_step & 1); }

inline void drop_surrogate_graphs(pvd_sim* s) {
  for (int a = 0; a < 2; ++a)
    for (int k = 0; k < 7; ++k)
      if (s->sgraph[a][k]) {
        (void)hipGraphExecDestroy(s->sgraph[a][k]);
        s->sgraph[a][k] = nullptr;
      }
}

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
  A.rsg = s->rsg;
  A.ncnt = s->ncnt;
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
  if (s->htrial) {
    A.xc_in = s->xc[cb];
    A.xc_out = s->cont ? s->xc[0] : s->xc[cb ^ 1];
    A.nx = s->nx;
    A.htrial = 1;
  }
  return A;
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

// Per-precision operations of the simulation handle (pvd_sim_step.hpp).
template <typename R>
int run_steps(pvd_sim* s, int n_steps, const double* disp, const double* um, const double* ub, hipStream_t st);
template <typename R>
int run_steps_timed(pvd_sim* s, int n_steps, hipStream_t st, float* avg_ms, int32_t* n_kernels);
template <typename R>
int do_move(pvd_sim* s, const double* disp, const double* um, hipStream_t st);
template <typename R>
int do_branch_ext(pvd_sim* s, const double* v, const double* ub, hipStream_t st);
template <typename R>
int do_start(pvd_sim* s, const double* v, hipStream_t st);
template <typename R>
int do_is_cache(pvd_sim* s, const double* x, const double* psi, const double* f, const double* sd, hipStream_t st);
template <typename R>
int do_is_propose(pvd_sim* s, const double* disp, double* y, hipStream_t st);
template <typename R>
int do_is_accept(pvd_sim* s, const double* y, const double* psi, const double* f, const double* sd,
                 const double* um, hipStream_t st);
template <typename R>
int do_pes(pvd_sim* s, double* v, hipStream_t st);

#define PVD_SIM_OPS(R)                                                                                   \
  template int run_steps<R>(pvd_sim*, int, const double*, const double*, const double*, hipStream_t);  \
  template int run_steps_timed<R>(pvd_sim*, int, hipStream_t, float*, int32_t*);                       \
  template int do_move<R>(pvd_sim*, const double*, const double*, hipStream_t);                         \
  template int do_branch_ext<R>(pvd_sim*, const double*, const double*, hipStream_t);                   \
  template int do_start<R>(pvd_sim*, const double*, hipStream_t);                                      \
  template int do_is_cache<R>(pvd_sim*, const double*, const double*, const double*, const double*,     \
                              hipStream_t);                                                              \
  template int do_is_propose<R>(pvd_sim*, const double*, double*, hipStream_t);                         \
  template int do_is_accept<R>(pvd_sim*, const double*, const double*, const double*, const double*,    \
                               const double*, hipStream_t);                                              \
  template int do_pes<R>(pvd_sim*, double*, hipStream_t);

}  // namespace pvd
