// Step-kernel launches of the simulation handle, templated on the walker element type:
// the fused step table, HIP graphs of 64 steps, the timed launches, the split path (move /
// external-energy branching) and the first-step vref.  Included by pvd_step_f32.hip and
// pvd_step_f64.hip, each of which instantiates one precision.
#pragma once
#include "pvd_sim.hpp"

namespace pvd {

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
      if (A.htrial) E = E + A.ke[i];  // local kinetic energy from pvd_sim_is_cache
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

// The kernels of one fused step, in launch order, for the current parity.  Used both for
// direct launches and as the node list of a replayed graph.
struct Launch {
  const void* fn;
  dim3 grid, block;
};

// Resident workgroups of a discrete step kernel on this device (the tiling unit), cached.
inline int disc_slots(pvd_sim* s, const void* fn) {
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
inline unsigned disc_grid(const pvd_sim* s, int slots, int W) {
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

// A graph of g = 2^lg surrogate steps from an even step, captured on the handle's own stream:
// every launch of a step is parameter-identical to the one two steps later (step-dependent
// state lives in the device control block), so the graph replays at any even step.  The judge
// runs outside graphs, on its own steps.
template <typename R>
int build_surrogate_graph(pvd_sim* s, int lg, hipGraphExec_t* out) {
  if (!s->cap_stream) HIPCHK(hipStreamCreateWithFlags(&s->cap_stream, hipStreamNonBlocking));
  const long long saved = s->next_step, saved_e = s->epoch_next;
  const int saved_je = s->judge_every;
  s->judge_every = 0;
  HIPCHK(hipStreamBeginCapture(s->cap_stream, hipStreamCaptureModeRelaxed));
  int rc = 0;
  for (int k = 0; k < (1 << lg) && rc == 0; ++k) rc = surrogate_step<R>(s, s->cap_stream);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(s->cap_stream, &g);
  s->next_step = saved;
  s->epoch_next = saved_e;
  s->judge_every = saved_je;
  if (rc == 0 && e != hipSuccess) rc = fail(E_HIP, std::string("surrogate graph capture: ") + hipGetErrorString(e));
  if (rc == 0) {
    const hipError_t ei = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    if (ei != hipSuccess) rc = fail(E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
  }
  if (g) (void)hipGraphDestroy(g);
  return rc;
}

template <typename R>
int run_surrogate_steps(pvd_sim* s, int n_steps, hipStream_t st) {
  int k = 0;
  while (k < n_steps) {
    // steps before the next judge step (which runs eagerly)
    long long free_steps = n_steps - k;
    if (s->judge_every > 0 && s->jn < s->jcap) {
      const long long t = s->next_step, je = s->judge_every;
      free_steps = std::min(free_steps, ((t + je) / je) * je - 1 - t);
    }
    if (s->use_graphs && s->cfg.branch_every <= 1 && !(s->next_step & 1) && free_steps >= 2) {
      int lg = 6;
      while ((1ll << lg) > free_steps) --lg;
      hipGraphExec_t& ge = s->sgraph[s->who_on ? 1 : 0][lg];
      if (!ge) {
        const int rc = build_surrogate_graph<R>(s, lg, &ge);
        if (rc) return rc;
      }
      HIPCHK(hipGraphLaunch(ge, st));
      s->next_step += 1ll << lg;
      s->epoch_next += 1ll << lg;
      k += 1 << lg;
      continue;
    }
    const int rc = surrogate_step<R>(s, st);
    if (rc) return rc;
    ++k;
  }
  return 0;
}

template <typename R>
int run_steps(pvd_sim* s, int n_steps, const double* disp, const double* um, const double* ub,
              hipStream_t st) {
  if (s->htrial) return fail(E_BAD, "a caller-evaluated trial runs the split step (is_propose / is_accept)");
  if (s->surr_on) {
    if (disp) return fail(E_BAD, "injected draws with a surrogate potential: use move / branch_external");
    return run_surrogate_steps<R>(s, n_steps, st);
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
  if (s->htrial) return fail(E_BAD, "a caller-evaluated trial runs the split step (is_propose / is_accept)");
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
  if (s->htrial) return fail(E_BAD, "a caller-evaluated trial moves with is_propose / is_accept");
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
  if (s->htrial && !v) return fail(E_BAD, "a caller-evaluated trial needs the first-step energies (start_external)");
  if (v) {
    if (s->is && !s->htrial) {
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

template <typename R>
int do_is_cache(pvd_sim* s, const double* x, const double* psi, const double* f, const double* sd, hipStream_t st) {
  if (!s->htrial) return fail(E_BAD, "the simulation was not created with a caller-evaluated trial");
  StepArgs<R> A = make_args<R>(s);
  hipLaunchKernelGGL((k_his_cache<R>), dim3(nblk(s->cap, TILE)), dim3(TILE), 0, st, A, x, psi, f, sd);
  HIPCHK(hipGetLastError());
  return 0;
}

template <typename R>
int do_is_propose(pvd_sim* s, const double* disp, double* y, hipStream_t st) {
  if (!s->htrial) return fail(E_BAD, "the simulation was not created with a caller-evaluated trial");
  StepArgs<R> A = make_args<R>(s);
  A.inj_disp = disp;
  if (disp) hipLaunchKernelGGL((k_his_propose<R, true>), dim3(nblk(s->cap, TILE)), dim3(TILE), 0, st, A, y);
  else hipLaunchKernelGGL((k_his_propose<R, false>), dim3(nblk(s->cap, TILE)), dim3(TILE), 0, st, A, y);
  HIPCHK(hipGetLastError());
  return 0;
}

template <typename R>
int do_is_accept(pvd_sim* s, const double* y, const double* psi, const double* f, const double* sd,
                 const double* um, hipStream_t st) {
  if (!s->htrial) return fail(E_BAD, "the simulation was not created with a caller-evaluated trial");
  StepArgs<R> A = make_args<R>(s);
  A.inj_umet = um;
  if (um) hipLaunchKernelGGL((k_his_accept<R, true>), dim3(nblk(s->cap, TILE)), dim3(TILE), 0, st, A, y, psi, f, sd);
  else hipLaunchKernelGGL((k_his_accept<R, false>), dim3(nblk(s->cap, TILE)), dim3(TILE), 0, st, A, y, psi, f, sd);
  HIPCHK(hipGetLastError());
  return 0;
}

template <typename R>
int do_pes(pvd_sim* s, double* v, hipStream_t st) {
  if (s->cfg.pes.kind == PES_EXTERNAL) return fail(E_BAD, "the simulation has no device potential");
  StepArgs<R> A = make_args<R>(s);
  PVD_NP_SWITCH(s->np, hipLaunchKernelGGL((k_pes_planes<R, NP>), dim3(nblk(s->cap, TILE)), dim3(TILE), 0, st, A, v));
  HIPCHK(hipGetLastError());
  return 0;
}

}  // namespace pvd
