/*
 * pyvibdmc_amd — C ABI of the MI355X (gfx950) DMC propagation library.
 *
 * The reference (rjdirisio/pyvibdmc) is pure Python; its hot path is the per-step loop of
 * DMC_Sim (pyvibdmc/pyvibdmc.py:701-876) and the plugin calls it makes:
 *   potential.getpot(cds)          (simulation_utilities/potential_manager.py:71-99)
 *   imp_samp.drift(cds)            (simulation_utilities/imp_samp.py:21-26)
 * The entry points below are what a binding of that path needs: a simulation handle that
 * owns the walkers in HBM and advances them, a split path for energies computed outside
 * the library, and stand-alone potential / trial-wave-function kernels.
 *
 * Conventions
 *  - Every function returns an int status: 0 ok, 1 population outside [0.5,1.5]*N0,
 *    3 non-finite or huge branching weight, 4 capacity, 5 unsupported continuous split,
 *    6 device wait expired, 7 bad argument, 8 HIP runtime error.  pvd_last_error() gives
 *    a thread-local message for the last failure.
 *  - Codes 1 and 3 are the reference's ValueError("Massive walker birth or death
 *    event!!!!!!! Dying...") (pyvibdmc.py:397-400, 409-413, 714-717).
 *  - Step functions are asynchronous on `stream` (a hipStream_t, NULL = default stream);
 *    device-side failures are reported by pvd_sim_status().
 *  - Pointers documented "device" must be device memory of the current HIP device;
 *    "host" pointers are ordinary host memory.  Inputs are borrowed for the call only.
 *  - One simulation handle must not be used from two threads at once.
 */
#ifndef PYVIBDMC_AMD_H
#define PYVIBDMC_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PVD_MAX_COMP 96 /* components per walker (atoms * dims) */
#define PVD_FUSED_MAX_COMP 18

enum pvd_precision { PVD_FP64 = 0, PVD_FP32 = 1 };
enum pvd_weighting { PVD_DISCRETE = 0, PVD_CONTINUOUS = 1 };
enum pvd_pes_kind {
  PVD_PES_EXTERNAL = 0,   /* energies supplied by the caller (split path)               */
  PVD_PES_HARMONIC_X = 1, /* V = p0 * x^2, x = first component of atom 0                */
  PVD_PES_MORSE_1D = 2,   /* V = p0 * (1 - exp(p1 * x))^2  (p1 = -alpha)                */
  PVD_PES_MORSE_BEND = 3, /* idx = (h1, h2, o); p = (De, a, re, kb, theta0)             */
  PVD_PES_CLUSTER6 = 4,   /* [O,H,H,O,H,H]; p = (De,a,re,kb,theta0, De_oo,a_oo,re_oo, qO,qH) */
  PVD_PES_PS_WATER = 5    /* Partridge-Schwenke water (h2opes_v2.f); idx = (h1, h2, o); no gradient */
};
enum pvd_trial_kind {
  PVD_TRIAL_NONE = 0,
  PVD_TRIAL_HARMONIC_1D = 1,
  PVD_TRIAL_GAUSS = 2,
  PVD_TRIAL_HOST = 3 /* psi_T evaluated by the caller: pvd_sim_is_propose / pvd_sim_is_accept */
};

/* Device potential.  Replaces the user function behind Potential*.getpot
 * (potential_manager.py:71-99, 144-174, 237-253). */
typedef struct pvd_pes_desc {
  int32_t kind;
  int32_t idx[4];
  double p[12];
} pvd_pes_desc;

/* Trial wave function psi_T = prod of Gaussians in bonds / angles, or the 1-D oscillator of
 * sample_potentials/PythonPots/harm_trial_wfn.py:6-40.  Replaces call_trial / call_derivs
 * (imp_samp_manager.py:92-139, 197-224). */
typedef struct pvd_trial_desc {
  int32_t kind, nb, na;
  int32_t bi[8], bj[8];
  int32_t ai[4], aj[4], ak[4];
  double balpha[8], br0[8];
  double aalpha[4], ath0[4];
  double c0, c1, c2; /* HARMONIC_1D: (alpha/pi)^0.25, alpha, alpha^2 | GAUSS: c0 = ln norm */
} pvd_trial_desc;

/* Simulation parameters: the DMC_Sim keyword arguments on the propagation path
 * (pyvibdmc.py:67-98, 132-297). */
typedef struct pvd_config {
  int32_t num_walkers;    /* N0                                                        */
  int32_t capacity;       /* walker slots; 0 = 1.5*N0 + 1 rounded up to 256            */
  int32_t n_atoms;        /* M                                                         */
  int32_t n_dim;          /* d                                                         */
  int32_t precision;      /* pvd_precision                                             */
  int32_t weighting;      /* pvd_weighting                                             */
  int32_t branch_every;   /* pyvibdmc.py:139                                           */
  int32_t num_timesteps;  /* length of the per-step record                             */
  double delta_t;         /* pyvibdmc.py:78                                            */
  double alpha;           /* 1/(2 dt) or DEBUG_alpha, pyvibdmc.py:200-203              */
  double thresh_lower;    /* continuous lower weight threshold, pyvibdmc.py:214-233    */
  double sigma[PVD_MAX_COMP];    /* sqrt(dt/m) per component, pyvibdmc.py:199          */
  double inv_mass[PVD_MAX_COMP]; /* 1/m per component, pyvibdmc.py:266, 274            */
  uint64_t seed;          /* Philox key for the production draws                       */
  pvd_pes_desc pes;
  pvd_trial_desc trial;
} pvd_config;

typedef struct pvd_sim pvd_sim;

const char* pvd_version(void);
const char* pvd_last_error(void);

/* --- simulation handle: DMC_Sim.__init__ / _initialize (pyvibdmc.py:67-297) ---------- */
int pvd_sim_create(const pvd_config* cfg, pvd_sim** out);
int pvd_sim_destroy(pvd_sim* sim);

/* Load walkers (host, (n, M, d) float64 C order) and continuous weights (host or NULL).
 * Starts the record at step `step0` (cur_timestep, pyvibdmc.py:85). */
int pvd_sim_set_walkers(pvd_sim* sim, const double* coords, int32_t n, const double* weights,
                        int64_t step0, void* stream);
/* Copy the walkers out (host, capacity >= n*M*d); weights NULL for discrete. Synchronous. */
int pvd_sim_get_walkers(pvd_sim* sim, double* coords, double* weights, int64_t max_walkers,
                        int32_t* n, void* stream);

/* First-step reference energy: V(x0) [+ local kinetic energy] then calc_vref
 * (pyvibdmc.py:760-769).  The _external form takes V(x0) from device memory; with a
 * surrogate set, pvd_sim_start evaluates it. */
int pvd_sim_start(pvd_sim* sim, void* stream);
int pvd_sim_start_external(pvd_sim* sim, const double* v_dev, void* stream);

/* n_steps fused steps with Philox draws: move, PES, weights / branching, vref
 * (pyvibdmc.py:772-848). */
int pvd_sim_run(pvd_sim* sim, int32_t n_steps, void* stream);
/* pvd_sim_run with one kernel launch at a time, each dispatch carrying a HIP start / stop
 * event pair on `stream` (hipExtLaunchKernel); avg_ms[k] (k < *n_kernels <= 8) is the mean
 * duration of the step's k-th kernel in launch order.  Synchronous.  No counterpart in the
 * reference (a measurement hook for the per-kernel timings a profiler also reports). */
int pvd_sim_run_timed(pvd_sim* sim, int32_t n_steps, void* stream, float* avg_ms, int32_t* n_kernels);
/* Streaming-copy ceiling of the walker layout: `reps` launches copying n walkers of np FP32
 * planes (in, out: device, stride cap, 16-byte aligned, cap a multiple of 4) with 16-byte
 * loads and stores; *avg_ms = mean launch time from events on `stream`.  Synchronous.  No
 * counterpart in the reference (a measurement hook: the bandwidth a one-pass read-then-write
 * kernel on these planes reaches). */
int pvd_bench_plane_copy(const float* in, float* out, int32_t np, int64_t cap, int64_t n, int32_t reps,
                         void* stream, float* avg_ms);

/* One step with the caller's draws (device, float64): displacements (n, M, d) as
 * np.random.normal(0, sigma, (n, d, M)).transpose(0, 2, 1) (pyvibdmc.py:544-546),
 * Metropolis uniforms (n) or NULL, branching uniforms (n) or NULL (pyvibdmc.py:392, 601). */
int pvd_sim_step_injected(pvd_sim* sim, const double* disp, const double* u_metro,
                          const double* u_branch, void* stream);

/* Split path for energies computed outside the library (a host Python PES or a device
 * surrogate): move the walkers in place, let the caller evaluate V on the moved walkers,
 * then branch with those energies.  disp / u_metro / u_branch: device or NULL (Philox). */
int pvd_sim_move(pvd_sim* sim, const double* disp, const double* u_metro, void* stream);
int pvd_sim_branch_external(pvd_sim* sim, const double* v_dev, const double* u_branch,
                            void* stream);
/* Importance sampling with a trial wave function evaluated by the caller (trial.kind =
 * PVD_TRIAL_HOST): imp_move_randomly (pyvibdmc.py:549-612) split around its plugin calls
 * self.impsamp.drift(cds) -> ImpSampManager.call_trial / call_derivs (imp_samp.py:21-26,
 * imp_samp_manager.py:92-139, 197-224).  The library keeps the move's caches (psi_1, f_x,
 * psi_sec_der) per walker and carries them through branching (pyvibdmc.py:424-427, 353-356).
 * All arrays are device float64 rows of the current population n: (n, M, d) or (n).
 *  - pvd_sim_is_cache: caches from the caller's drift(x) of the current walkers; x (or NULL)
 *    replaces the walkers (a finite-difference call perturbs its input in place); stores the
 *    local kinetic energy of d2psi for the first-step energies (pyvibdmc.py:760-769).
 *  - pvd_sim_is_propose: y = x + disp + (1/m) f_x dt (pyvibdmc.py:556-564); disp as in
 *    pvd_sim_step_injected, or NULL for Philox draws.
 *  - pvd_sim_is_accept: Metropolis of imp_samp.py:28-47 against the caller's psi / dpsi /
 *    d2psi at y (y as the caller left it), acceptance, caches and local kinetic energy
 *    (pyvibdmc.py:566-609, 807-809); u_metro device or NULL (Philox).  Then a potential and
 *    pvd_sim_branch_external complete the step. */
int pvd_sim_is_cache(pvd_sim* sim, const double* x, const double* psi, const double* dpsi,
                     const double* d2psi, void* stream);
int pvd_sim_is_propose(pvd_sim* sim, const double* disp, double* y, void* stream);
int pvd_sim_is_accept(pvd_sim* sim, const double* y, const double* psi_y, const double* dpsi_y,
                      const double* d2psi_y, const double* u_metro, void* stream);
/* The simulation's device potential at its current walkers (potential.getpot(
 * self._walker_coords), pyvibdmc.py:786-793): v (device, float64, n). */
int pvd_sim_pes(pvd_sim* sim, double* v, void* stream);

/* Current walkers as SoA planes: component p of walker i at planes[p*stride + i];
 * element type float64 (PVD_FP64) or float32 (PVD_FP32). */
int pvd_sim_planes(pvd_sim* sim, void** planes, int64_t* stride, int32_t* n, void* stream);

/* Synchronises `stream`; returns the device status as the return value and fills the
 * step the failure happened at, the population and the current vref. */
int pvd_sim_status(pvd_sim* sim, int32_t* status, int64_t* fail_step, int32_t* n, double* vref,
                   int64_t* next_step, void* stream);
/* Per-step record (host arrays of `count`): vref_vs_tau, pop_vs_tau, births (discrete) or
 * number branched (continuous), deaths (discrete) or the split replay that ran (continuous:
 * 0 none; from the candidate lists: 1 list A whole, 5 list A with a cut inside, 6 lists A
 * and B; 2-4 the general interval replay because the lists held too few / too many walkers /
 * other), cumulative effective
 * time (IS) (pyvibdmc.py:674-682, 372-378). Any pointer may be NULL. */
int pvd_sim_trajectory(pvd_sim* sim, int64_t t0, int32_t count, double* vref, double* pop,
                       int32_t* births, int32_t* deaths, double* eff_t);

/* Descendant weighting window (pyvibdmc.py:739-747, 663-672, 856-872).  end() writes
 * desc_wts (host, n_parent), parent coords (host, (n_parent, M, d)) and parent weights
 * (host or NULL). */
int pvd_sim_desc_begin(pvd_sim* sim, void* stream);
int pvd_sim_desc_end(pvd_sim* sim, double* desc_wts, double* parent, double* parent_wts,
                     int32_t* n_parent, void* stream);

/* --- stand-alone kernels on AoS walkers (n, M, d), element type by `precision` ------- */
/* V(x): the device form of getpot. */
int pvd_pes_eval(const pvd_pes_desc* pes, int32_t precision, int32_t n_atoms, int32_t n_dim,
                 const void* coords, int64_t n, void* v_out, void* stream);
/* grad_out[i] * dV/dx at each walker (backward of V; forces are -dV/dx). */
int pvd_pes_vjp(const pvd_pes_desc* pes, int32_t precision, int32_t n_atoms, int32_t n_dim,
                const void* coords, int64_t n, const void* grad_out, void* grad_coords,
                void* stream);
/* psi, ln psi, dpsi/psi and d2psi/psi (diagonal) at each walker (ImpSamp.drift). NULL
 * outputs are skipped. */
int pvd_trial_eval(const pvd_trial_desc* trial, int32_t precision, int32_t n_atoms,
                   int32_t n_dim, const void* coords, int64_t n, void* psi, void* lnpsi,
                   void* dlnpsi, void* d2psi_over_psi, void* stream);
/* grad_out[i] * d ln psi / dx (backward of ln psi; recomputes the forward). */
int pvd_trial_vjp(const pvd_trial_desc* trial, int32_t precision, int32_t n_atoms,
                  int32_t n_dim, const void* coords, int64_t n, const void* grad_out,
                  void* grad_coords, void* stream);

/* --- surrogate potential: descriptor + MLP on BF16 MFMA / FP32 accumulation ---------
 * Learned replacement of a user potential (the NN_Potential convention,
 * potential_manager.py:177-214 and sample_potentials/TensorflowPots/call_sample_model.py:4-9),
 * with the descriptor of simulation_utilities/tensorflow_descriptors/distance_descriptors.py
 * (inverse pair distances sorted inside like-atom groups, :102-152).  Network:
 * D -> H -> H -> H -> 1 with SiLU, H in {128, 256}, D <= 32; flat FP32 parameters in the
 * order W0 (H x D), b0, W1 (H x H), b1, W2, b2, w3 (H), b3. */
typedef struct pvd_mlp {
  int32_t D, H;
  /* Descriptor (simulation_utilities/tensorflow_descriptors/distance_descriptors.py:6-214,
   * DistIt with method 'coulomb'): the inverse distances of all atom pairs after permuting like
   * atoms, position pairs (0,1), (0,2), ..., (1,2), ...  n_atoms = 0: a network trained on
   * given features only; otherwise 2..6 atoms and D = n_atoms (n_atoms - 1) / 2. */
  int32_t n_atoms;
  int32_t z[6];                /* nuclear charges; the sort key is the Coulomb-matrix row norm */
  int32_t sorted_atoms[6];     /* DistIt sorted_atoms: sublists concatenated, covering every atom */
  int32_t sorted_atoms_len[6]; /* their lengths; n_sorted_atoms = 0: no atom sorting */
  int32_t n_sorted_atoms;
  int32_t sorted_groups[6];    /* DistIt sorted_groups: n_groups x group_size positions */
  int32_t n_groups, group_size;  /* n_groups < 2: no group sorting */
  int32_t use_range;       /* 1: walkers outside [lo, hi] use the fallback potential */
  float xmean[32], xinv[32], lo[32], hi[32];
  float ymean, ystd;
  /* FP32 quadratic baseline in the normalised features x: V = ystd * y + ymean + q(x),
   * q(x) = quad[0] + sum_i quad[1 + i] x_i + sum_{i <= j} quad[..] x_i x_j (pairs in row
   * order i = 0..D-1, j = i..D-1); all zero for a plain network. */
  float quad[136];
  const float* params;     /* device, pvd_mlp_param_count(D, H) floats */
  void* packs;             /* device, pvd_mlp_pack_bytes(H) bytes, written by pvd_mlp_pack */
} pvd_mlp;

const char* pvd_mlp_last_error(void);
int64_t pvd_mlp_param_count(int32_t D, int32_t H);
int64_t pvd_mlp_pack_bytes(int32_t H);
int64_t pvd_mlp_workspace_bytes(int32_t D, int32_t H, int64_t batch);
/* FP32 master parameters -> BF16 MFMA operand packs (forward and transposed). */
int pvd_mlp_pack(const pvd_mlp* m, void* stream);
/* Descriptors of AoS walkers (device, (n, n_atoms, 3)) into features_t (device, D x n),
 * optionally normalised with xmean / xinv. */
int pvd_mlp_descriptors(const pvd_mlp* m, int32_t precision, int32_t n_atoms, const void* coords, int64_t n,
                        int32_t normalise, float* features_t, void* stream);
/* V (device, float64) of walkers in SoA planes (component p of walker i at
 * planes[p*stride + i]); the walker count is read from n_dev on the device when non-NULL,
 * else n_max.  `fallback` is the analytic potential used outside the training range. */
int pvd_mlp_energy(const pvd_mlp* m, int32_t precision, int32_t n_atoms, const void* planes, int64_t stride,
                   const int32_t* n_dev, int64_t n_max, double* v_out, const pvd_pes_desc* fallback, void* stream);
/* Error judge of a surrogate: the analytic potential on every `every`-th walker of the SoA
 * planes against v (device, float64).  No counterpart in the reference (its NN potential,
 * simulation_utilities/potential_manager.py:177-214, is never checked during a run);
 * record (device, 4 doubles) = {step, sum of squared errors, max |error|, samples}. */
int pvd_mlp_judge(const pvd_pes_desc* pes, int32_t precision, int32_t n_atoms, const void* planes, int64_t stride,
                  const int32_t* n_dev, int64_t n_max, const double* v, int32_t every, int64_t step, double* record,
                  void* stream);
/* Mean-squared-error gradients (device, flat parameter order) on normalised features
 * x_t (D x n) and targets (n); y_out (n) or NULL; loss_parts (ceil(n/256) doubles). */
int pvd_mlp_grads(const pvd_mlp* m, const float* x_t, const float* targets, int64_t n, float* grads, float* y_out,
                  double* loss_parts, void* workspace, void* stream);
/* pvd_mlp_grads without a workspace argument and without any device allocation.  No
 * counterpart in the reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, training loops that forbid hipMalloc).  The workspace is a static
 * device array of pvd_mlp_noalloc_scratch_bytes() (32 MiB) per device, held until the process
 * ends; batches whose workspace (pvd_mlp_workspace_bytes) exceeds it are rejected with code 7.
 * Every call on one device shares that device's array: calls that overlap on one GPU race.
 * Nothing in it needs clearing (each call writes what it reads). */
int64_t pvd_mlp_noalloc_scratch_bytes(void);
int pvd_mlp_grads_noalloc(const pvd_mlp* m, const float* x_t, const float* targets, int64_t n, float* grads,
                          float* y_out, double* loss_parts, void* stream);
/* One Adam step (Kingma & Ba) with bias correction at step t >= 1. */
int pvd_mlp_adam(float* params, float* m1, float* m2, const float* grads, int64_t size, float lr, float beta1,
                 float beta2, float eps, int64_t t, void* stream);
/* Device view of the current walkers without a host sync: SoA planes, stride, and a
 * device pointer to the current population. */
int pvd_sim_device_view(pvd_sim* sim, void** planes, int64_t* stride, int32_t** n_dev);
/* Use a surrogate network as the potential of a simulation created with an external
 * potential (the NN_Potential role, simulation_utilities/potential_manager.py:177-214):
 * pvd_sim_run then moves, evaluates the network on the device (walkers outside the training
 * range use `fallback`), judges every `judge_every` steps on every `judge_stride`-th walker
 * (0: no judge) and branches, with no host synchronisation.  The network's device
 * parameters must stay alive while the simulation uses them; net == NULL removes it. */
int pvd_sim_set_surrogate(pvd_sim* sim, const pvd_mlp* net, const pvd_pes_desc* fallback, int32_t judge_every,
                          int32_t judge_stride);
/* Judge records written so far (each {step, sum sq err, max |err|, samples}); copies up to
 * max_records into out (host) and sets *n. Synchronous. */
int pvd_sim_judge_log(pvd_sim* sim, double* out, int32_t max_records, int32_t* n, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* PYVIBDMC_AMD_H */
