// C ABI of the DMC propagation library: simulation handle, launch dispatch and the
// auxiliary kernels (layout transposes, first-step vref, descendant weights, stand-alone
// potential / trial evaluation).  Declarations and conventions: include/pyvibdmc_amd.h.
#include <initializer_list>

#include "pvd_sim.hpp"

using namespace pvd;

static_assert(sizeof(pvd_pes_desc) == sizeof(PesParams), "pes layout");
static_assert(sizeof(pvd_trial_desc) == sizeof(TrialParams), "trial layout");
static_assert(PVD_MAX_COMP == MAXNP, "component bound");

namespace {
thread_local std::string g_err;
}  // namespace

int pvd::fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

namespace {

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

// The compiled model specialisation for a configuration: MS_PES_WATER for the [H, H, O]
// Morse+bend surface in 3 dimensions, plus MS_TRIAL_WATER when importance sampling uses the
// matching two-bond / one-angle Gaussian trial; MS_PES_CLUSTER / MS_TRIAL_CLUSTER for the
// [O, H, H, O, H, H] cluster surface and its per-monomer trial.  Other layouts keep the
// runtime parameters.  The single-precision trial specialisations weight the Laplacian and
// the drift with one 1/m for every hydrogen and one for every oxygen, so they are chosen
// only when the masses of those atoms are equal (an isotopologue such as HOD, or a mixed
// H2O / D2O cluster, takes the per-atom path).
bool same_mass(const pvd_config& c, std::initializer_list<int> atoms) {
  const double m0 = c.inv_mass[3 * *atoms.begin()];
  for (int a : atoms)
    for (int k = 0; k < 3; ++k)
      if (c.inv_mass[3 * a + k] != m0) return false;
  return true;
}

int model_spec(const pvd_config& c, int np) {
  const pvd_pes_desc& p = c.pes;
  const pvd_trial_desc& t = c.trial;
  if (np == 18 && c.n_atoms == 6 && c.n_dim == 3) {
    const bool pc = p.kind == PES_CLUSTER6;
    const bool tc = t.kind == TRIAL_GAUSS && t.nb == 4 && t.na == 2 && t.bi[0] == 1 && t.bj[0] == 0 &&
                    t.bi[1] == 2 && t.bj[1] == 0 && t.bi[2] == 4 && t.bj[2] == 3 && t.bi[3] == 5 && t.bj[3] == 3 &&
                    t.ai[0] == 1 && t.aj[0] == 0 && t.ak[0] == 2 && t.ai[1] == 4 && t.aj[1] == 3 && t.ak[1] == 5 &&
                    same_mass(c, {1, 2, 4, 5}) && same_mass(c, {0, 3});
    return (pc ? MS_PES_CLUSTER : MS_ANY) | (tc ? MS_TRIAL_CLUSTER : MS_ANY);
  }
  if (np != 9 || c.n_atoms != 3 || c.n_dim != 3) return MS_ANY;
  const bool pw = p.kind == PES_MORSE_BEND && p.idx[0] == 0 && p.idx[1] == 1 && p.idx[2] == 2;
  const bool tw = t.kind == TRIAL_GAUSS && t.nb == 2 && t.na == 1 && t.bi[0] == 0 && t.bj[0] == 2 &&
                  t.bi[1] == 1 && t.bj[1] == 2 && t.ai[0] == 0 && t.aj[0] == 2 && t.ak[0] == 1 &&
                  same_mass(c, {0, 1}) && same_mass(c, {2});
  return (pw ? MS_PES_WATER : MS_ANY) | (tw ? MS_TRIAL_WATER : MS_ANY);
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
  drop_surrogate_graphs(s);
  if (s->cap_stream) (void)hipStreamDestroy(s->cap_stream);
  void* ptrs[] = {s->coords[0], s->coords[1], s->who[0], s->who[1], s->w, s->E, s->ke, s->kmask,
                  s->ctrl, s->flags, s->gflags, s->rec, s->drec, s->accs, s->tr_vref, s->tr_pop, s->tr_dteff, s->tr_births,
                  s->tr_deaths, s->pend, s->fin, s->klist, s->killtag, s->rstage, s->cand, s->rsg, s->ncnt, s->asg, s->cpy, s->stage, s->parent, s->parent_w, s->desc,
                  s->cnt, s->dacc, s->vsur, s->jlog, s->xc[0], s->xc[1]};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  delete s;
}

// Streaming copy of np SoA planes (component p of walker i at planes[p*cap + i]): each
// thread moves 4 consecutive walkers of every plane with one 16-byte load and store (the
// tail walkers one by one).  The read-then-write ceiling of the walker layout.
__global__ __launch_bounds__(256) void k_plane_copy(const float* __restrict__ in, float* __restrict__ out, int np,
                                                    long long cap, long long n) {
  const long long i = ((long long)blockIdx.x * 256 + threadIdx.x) * 4;
  if (i >= n) return;
  if (i + 3 < n) {
    for (int p = 0; p < np; ++p)
      *(float4*)(out + p * cap + i) = *(const float4*)(in + p * cap + i);
  } else {
    for (int p = 0; p < np; ++p)
      for (long long k = i; k < n; ++k) out[p * cap + k] = in[p * cap + k];
  }
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
  if (cfg->trial.kind != TRIAL_NONE && cfg->trial.kind != TRIAL_HOST && np > PVD_FUSED_MAX_COMP)
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
  s->htrial = cfg->trial.kind == TRIAL_HOST;
  s->nx = s->htrial ? 2 * np + 1 : 0;
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
    ALLOC(s->rsg, RSG_BYTES);
    ALLOC(s->ncnt, CNT_BYTES);
    ALLOC(s->asg, sizeof(int) * KILL_CAP);
    ALLOC(s->cpy, sizeof(int) * 2 * KILL_CAP);
    ALLOC(s->killtag, sizeof(unsigned long long) * cap);
    ALLOC(s->rstage, sizeof(double) * (size_t)RP_FCAP * (np + 2 + (cfg->trial.kind == TRIAL_HOST ? 2 * np + 1 : 0)));
  }
  ALLOC(s->stage, sizeof(double) * np * cap);
  if (s->htrial) {
    ALLOC(s->xc[0], sizeof(double) * s->nx * cap);
    if (!s->cont) ALLOC(s->xc[1], sizeof(double) * s->nx * cap);
  }
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
  if (e == hipSuccess && s->ncnt) e = hipMemset(s->ncnt, 0, CNT_BYTES);
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

int pvd_bench_plane_copy(const float* in, float* out, int32_t np, int64_t cap, int64_t n, int32_t reps, void* stream,
                         float* avg_ms) {
  if (!in || !out || !avg_ms) return fail(E_BAD, "null argument");
  if (np < 1 || n < 1 || cap < n || (cap & 3) || reps < 1) return fail(E_BAD, "bad plane shape");
  if ((((uintptr_t)in) | ((uintptr_t)out)) & 15) return fail(E_BAD, "planes must be 16-byte aligned");
  hipStream_t st = (hipStream_t)stream;
  const unsigned grid = (unsigned)((n + 1023) / 1024);
  hipEvent_t a = nullptr, b = nullptr;
  HIPCHK(hipEventCreate(&a));
  if (hipEventCreate(&b) != hipSuccess) {
    (void)hipEventDestroy(a);
    return fail(E_HIP, "hipEventCreate failed");
  }
  hipLaunchKernelGGL(k_plane_copy, dim3(grid), dim3(256), 0, st, in, out, (int)np, (long long)cap, (long long)n);
  hipError_t e = hipEventRecord(a, st);
  for (int r = 0; r < reps && e == hipSuccess; ++r) {
    hipLaunchKernelGGL(k_plane_copy, dim3(grid), dim3(256), 0, st, in, out, (int)np, (long long)cap, (long long)n);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipEventRecord(b, st);
  if (e == hipSuccess) e = hipEventSynchronize(b);
  float ms = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  if (e != hipSuccess) return fail(E_HIP, std::string("plane copy: ") + hipGetErrorString(e));
  *avg_ms = ms / (float)reps;
  return 0;
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

int pvd_sim_is_cache(pvd_sim* s, const double* x, const double* psi, const double* dpsi, const double* d2psi,
                     void* stream) {
  if (!s || !psi || !dpsi || !d2psi) return fail(E_BAD, "null argument");
  return s->prec == PVD_FP64 ? do_is_cache<double>(s, x, psi, dpsi, d2psi, (hipStream_t)stream)
                             : do_is_cache<float>(s, x, psi, dpsi, d2psi, (hipStream_t)stream);
}

int pvd_sim_is_propose(pvd_sim* s, const double* disp, double* y, void* stream) {
  if (!s || !y) return fail(E_BAD, "null argument");
  return s->prec == PVD_FP64 ? do_is_propose<double>(s, disp, y, (hipStream_t)stream)
                             : do_is_propose<float>(s, disp, y, (hipStream_t)stream);
}

int pvd_sim_is_accept(pvd_sim* s, const double* y, const double* psi_y, const double* dpsi_y, const double* d2psi_y,
                      const double* u_metro, void* stream) {
  if (!s || !y || !psi_y || !dpsi_y || !d2psi_y) return fail(E_BAD, "null argument");
  return s->prec == PVD_FP64 ? do_is_accept<double>(s, y, psi_y, dpsi_y, d2psi_y, u_metro, (hipStream_t)stream)
                             : do_is_accept<float>(s, y, psi_y, dpsi_y, d2psi_y, u_metro, (hipStream_t)stream);
}

int pvd_sim_pes(pvd_sim* s, double* v, void* stream) {
  if (!s || !v) return fail(E_BAD, "null argument");
  return s->prec == PVD_FP64 ? do_pes<double>(s, v, (hipStream_t)stream) : do_pes<float>(s, v, (hipStream_t)stream);
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
  if (net->n_atoms != s->M) return fail(E_BAD, "the network's descriptor is for another number of atoms");
  if (judge_every < 0 || (judge_every > 0 && judge_stride < 1)) return fail(E_BAD, "bad judge schedule");
  if (!s->vsur) HIPCHK(hipMalloc((void**)&s->vsur, sizeof(double) * s->cap));
  if (judge_every > 0 && !s->jlog) {
    s->jcap = (int)std::min<long long>(1 << 20, s->tmax / judge_every + 2);
    HIPCHK(hipMalloc((void**)&s->jlog, sizeof(double) * 4 * s->jcap));
  }
  drop_surrogate_graphs(s);  // they carry the previous description's scalars
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
  if (pes->kind == PES_PS_WATER) return fail(E_BAD, "this potential has no analytic gradient");
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
