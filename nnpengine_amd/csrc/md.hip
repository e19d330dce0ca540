// md.hip -- the MD step driver (the omp-parallel loop of src/run.cpp:230-393 for classical and
// :1127-1353 for nnp_run) on one HIP stream, with each distinct step shape captured once into a
// hipGraph and replayed.
//
// Step (in the reference's order):
//   updateX -> [rebuild: refold, sort, ghosts, lists] -> [g(r) tally (+ S(q))] ->
//   forces: tabulated pair (writes f), bonded (+=), NNP descriptors -> MLP -> NNP forces (+=)
//   -> updateV -> [kinetic energy + energy/virial log record]
// Rebuilds are either every `rebuild_every` steps (device-side skin check counts any step that
// came too late), or decided by NEIGHLIST::isvalid each step (host read-back, no graphs).
// Graph safety: every enqueue helper launches kernels only (no host reads, no memsets, fixed
// buffers); energy records are appended by a kernel at a device-held cursor.
#include <algorithm>
#include <map>

#include "internal.h"

namespace {

enum StepBits { kRebuild = 1, kGr = 2, kGrFull = 4, kLog = 8 };
constexpr int kCopyProbeBytes = 64 << 20;  // each direction

struct MdGraphs {
  // one graph per step shape, and one per sequence of step shapes of a whole rebuild interval
  std::map<std::vector<int>, hipGraphExec_t> exec;
  std::map<std::vector<int>, hipGraph_t> graph;
};

// step shape of step s of a run with parameters P (first: the step that builds the first list)
int step_key(const nnpe_md_params& P, long long s, bool first) {
  int key = 0;
  if (P.gr_every > 0 && s % P.gr_every == 0) key |= kGr | ((P.gr_long_every > 0 && s % P.gr_long_every == 0) ? kGrFull : 0);
  if (P.log_every > 0 && s % P.log_every == 0) key |= kLog;
  if (P.rebuild_every > 0 && (first || s % P.rebuild_every == 0)) key |= kRebuild;
  return key;
}

__global__ void k_skin_check(int n, float dx2, const float4* xq, const float4* xprev, DevCounters* cnt) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  bool bad = false;
  if (i < n) {
    float4 a = xq[i], b = xprev[i];
    float dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
    bad = dx * dx + dy * dy + dz * dz > dx2;
  }
  if (__any(bad) && (threadIdx.x & 63) == 0) atomicOr(&cnt->invalid, 1);
}

// the RMDF slot is not cleared: it holds the residual of the last refinement step
__global__ void k_log_append(double* ev_sum, double* log, int cap, int nval, DevCounters* cnt) {
  const int slot = cnt->log_count;
  if (threadIdx.x < nval) {
    if (slot < cap) log[(size_t)slot * nval + threadIdx.x] = ev_sum[threadIdx.x];
    if (threadIdx.x / 2 != NNPE_TERM_RMDF) ev_sum[threadIdx.x] = 0.0;
  }
  __syncthreads();
  if (threadIdx.x == 0) cnt->log_count = slot + 1;
}

__global__ void k_md_reset(DevCounters* cnt) {
  cnt->log_count = 0;
  cnt->skin_violations = 0;
  cnt->invalid = 0;
}

__global__ void k_clear_force(int n, float4* f) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) f[i] = make_float4(0, 0, 0, 0);
}

// plain streaming copy of a known number of bytes (calibrates memory-traffic counters)
__global__ void __launch_bounds__(256) k_copy_probe(int n, const float4* __restrict__ a, float4* __restrict__ b) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) b[i] = a[i];
}

// MFMA issue-rate probe: every wave runs kMfmaProbeIters rounds of four independent
// v_mfma_f32_32x32x16_bf16 chains on register operands, then stores one value per lane so the
// chains stay live.  kMfmaProbeBlocks workgroups of 4 waves: two waves per SIMD on 256 CUs.
// Flops per launch: 4 kMfmaProbeBlocks x kMfmaProbeIters x 4 x 2 x 32 x 32 x 16.
constexpr int kMfmaProbeIters = 2048;
constexpr int kMfmaProbeBlocks = 512;
__global__ void __launch_bounds__(256) k_mfma_probe(float* __restrict__ out, float seed) {
  typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
  typedef float f32x16 __attribute__((ext_vector_type(16)));
  bf16x8 a, b;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    a[k] = (__bf16)(seed * float((threadIdx.x + k) & 15));
    b[k] = (__bf16)(seed * float((k + 3 * threadIdx.x) & 15));
  }
  f32x16 c0, c1, c2, c3;
#pragma unroll
  for (int j = 0; j < 16; ++j) c0[j] = c1[j] = c2[j] = c3[j] = 0.0f;
  for (int it = 0; it < kMfmaProbeIters; ++it) {
    c0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c0, 0, 0, 0);
    c1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(b, a, c1, 0, 0, 0);
    c2 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, a, c2, 0, 0, 0);
    c3 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(b, b, c3, 0, 0, 0);
  }
  float sum = 0.0f;
#pragma unroll
  for (int j = 0; j < 16; ++j) sum += (c0[j] + c1[j]) + (c2[j] + c3[j]);
  out[blockIdx.x * blockDim.x + threadIdx.x] = sum;
}

MdGraphs* graphs_of(nnpe_ctx* ctx) {
  if (!ctx->md) ctx->md = new MdGraphs;
  return static_cast<MdGraphs*>(ctx->md);
}

// fx: the position update fused into the NNP force pass, inside a captured rebuild interval of
// the FP64 form (0: not fused; 1: the interval's first step, whose update is the separate kernel;
// 2: a middle step; 3: the last step, whose force pass copies the positions instead of advancing
// them).  A fused force pass writes the next positions into ctx->xq_alt while other atoms still
// read this step's from ctx->xq, and the two pointers trade places after the step; over an
// interval of even length they end where they began, with the current positions in ctx->xq.
int enqueue_step(nnpe_ctx* ctx, const nnpe_md_params& P, int key, bool first, int fx = 0) {
  int rc;
  hipStream_t s = ctx->stream;
  // on a rebuild step the position update also makes the skin check (skin/2 since the last build)
  const bool skin = (key & kRebuild) && !first && P.rebuild_every > 0;
  const bool upd = !first && fx <= 1;
  if (upd && (rc = skin ? verlet_enqueue_x_skin(ctx) : verlet_enqueue_x(ctx))) return rc;
  // a full g(r) tally on a rebuild step rides on the rebuild (same positions, fresh cells)
  const bool gr_fused = (key & kRebuild) && (key & kGr) && (key & kGrFull) && gr_fusable(ctx);
  // The classical terms (they write f first) and the NNP descriptors, network and adjoints (they
  // touch no force) are independent: with an NNP term the classical ones go to stream2 and the
  // NNP force, which adds into f, waits for them.  Captured, the step graph gets two branches.
  // The branch opens before the descriptors, so the classical kernels overlap the descriptors,
  // the network and the adjoints, and the cross-queue start of the side branch overlaps the
  // descriptors instead of delaying everything after them (0.1508 -> 0.1418 ms per step in the
  // 24000-atom BF16 form against opening it after the descriptors; md_fork_early false).  On a
  // tally step the tally's own passes (the shell pass and fold of a tally folded into the
  // rebuild, or the whole tally, then S(q)) read only positions, cells and lists, so they go to
  // stream2 ahead of the classical terms; with RMDF, whose table rewrite the same step's forces
  // read, they stay in line.
  const bool fork = P.nnp != 0 && ctx->md_fork;
  const bool gr_side = fork && (key & kGr) && !P.rmdf;
  if (key & kRebuild) {
    if (skin && !upd)
      hipLaunchKernelGGL(k_skin_check, dim3(blocks_for(ctx->natom, 256)), dim3(256), 0, s, ctx->natom, ctx->dx2, ctx->xq,
                         ctx->xprev, ctx->counters);
    if ((rc = neigh_enqueue_build(ctx, gr_fused, skin, gr_side))) return rc;
  }
  int rg = NNPE_OK;
  if (gr_side) {
    NNPE_CHECK_HIP(ctx, hipEventRecord(ctx->ev_fork, s));
    NNPE_CHECK_HIP(ctx, hipStreamWaitEvent(ctx->stream2, ctx->ev_fork, 0));
    ctx->stream = ctx->stream2;
    rg = gr_fused ? gr_shell_finish_enqueue(ctx) : gr_enqueue(ctx, (key & kGrFull) != 0);
    if (!rg && P.sq) rg = sq_enqueue(ctx);
    ctx->stream = s;
  } else if (key & kGr) {
    if (!gr_fused && (rc = gr_enqueue(ctx, (key & kGrFull) != 0))) return rc;
    if (P.sq && (rc = sq_enqueue(ctx))) return rc;
    // RMDF compute_Qspace / compute_Rspace, and compute_Potential on logged steps when
    // gamma != 1 (src/run.cpp:294-298, :1206-1210)
    if (P.rmdf && (rc = rmdf_enqueue(ctx, (key & kLog) && rmdf_gamma(ctx) != 1.0f))) return rc;
  }
  const bool ev = (key & kLog) != 0;
  const bool early = fork && !gr_side && ctx->md_fork_early;
  if (early) {
    NNPE_CHECK_HIP(ctx, hipEventRecord(ctx->ev_fork, s));
    NNPE_CHECK_HIP(ctx, hipStreamWaitEvent(ctx->stream2, ctx->ev_fork, 0));
  }
  int rn = NNPE_OK;
  if (P.nnp == 1) rn = nnp_tbl_desc_enqueue(ctx);
  else if (P.nnp == 2) rn = sf_desc_enqueue(ctx);
  if (rn && !gr_side && !early) return rn;  // (with the branch already open, it is closed below first)
  if (fork) {
    if (!gr_side && !early) {
      NNPE_CHECK_HIP(ctx, hipEventRecord(ctx->ev_fork, s));
      NNPE_CHECK_HIP(ctx, hipStreamWaitEvent(ctx->stream2, ctx->ev_fork, 0));
    }
    ctx->stream = ctx->stream2;
  }
  rc = rg;
  if (rc) {
  } else if (P.pair) {
    rc = pair_enqueue(ctx, ev);
  } else {
    hipLaunchKernelGGL(k_clear_force, dim3(blocks_for(ctx->natom, 256)), dim3(256), 0, ctx->stream, ctx->natom, ctx->f);
    rc = NNPE_OK;
  }
  if (!rc && P.bonded) rc = bonded_enqueue(ctx, ev);
  if (fork) {
    ctx->stream = s;  // the join is recorded even after an error, so a capture always closes
    NNPE_CHECK_HIP(ctx, hipEventRecord(ctx->ev_join, ctx->stream2));
  }
  if (rn) {
  } else if (P.nnp == 1) {
    rn = mlp_enqueue(ctx, ev);
  } else if (P.nnp == 2) {
    rn = mlp_enqueue(ctx, ev);
    if (!rn) rn = sf_adjoint_enqueue(ctx);
  }
  if (fork) NNPE_CHECK_HIP(ctx, hipStreamWaitEvent(s, ctx->ev_join, 0));
  if (rc) return rc;
  if (rn) return rn;
  // the SF force pass is the last writer of each atom's force: the FP64 half-kick rides on it
  const bool kick = P.nnp == 2 && ctx->fp64;
  if (P.nnp == 1) {
    if ((rc = nnp_tbl_force_enqueue(ctx, ev))) return rc;
  } else if (P.nnp == 2) {
    if ((rc = sf_force_only_enqueue(ctx, ev, kick, fx == 0 ? 0 : (fx == 3 ? 2 : 1)))) return rc;
  }
  if (!kick && (rc = verlet_enqueue_v(ctx))) return rc;
  if (ev) {
    if ((rc = ke_enqueue(ctx))) return rc;
    hipLaunchKernelGGL(k_log_append, dim3(1), dim3(64), 0, s, ctx->ev_sum, ctx->md_log, ctx->md_log_cap, 2 * NNPE_NTERM,
                       ctx->counters);
  }
  NNPE_CHECK_LAUNCH(ctx);
  if (fx) std::swap(ctx->xq, ctx->xq_alt);
  return NNPE_OK;
}

}  // namespace

void md_state_free(nnpe_ctx* ctx) {
  if (ctx->md) {
    MdGraphs* G = static_cast<MdGraphs*>(ctx->md);
    for (auto& kv : G->exec) (void)hipGraphExecDestroy(kv.second);
    for (auto& kv : G->graph) (void)hipGraphDestroy(kv.second);
    delete G;
    ctx->md = nullptr;
  }
}

extern "C" int nnpe_md_run(nnpe_ctx* ctx, const nnpe_md_params* p, int nsteps, double* log_out, int log_cap, int* nlog) {
  if (!ctx || !p || nsteps < 0) return NNPE_ERR_ARG;
  const nnpe_md_params P = *p;
  int rc0;
  if (!ctx->have_grid) return ctx_fail(ctx, NNPE_ERR_STATE, "nnpe_md_run before nnpe_neigh_init");
  if (P.pair && !ctx->have_pair) return ctx_fail(ctx, NNPE_ERR_STATE, "pair term without a table");
  if (P.nnp && !ctx->mlp) return ctx_fail(ctx, NNPE_ERR_STATE, "NNP term without nnpe_mlp_set");
  if (P.nnp == 1 && (!ctx->nnp || !nnp_state_of(ctx) || ctx->mlp == nullptr)) return ctx_fail(ctx, NNPE_ERR_STATE, "NNP tables missing");
  if (P.nnp == 2 && !ctx->cap_short) return ctx_fail(ctx, NNPE_ERR_STATE, "symmetry functions need the short lists (nnpe_neigh_init short_cutoff)");
  if ((P.gr_every > 0) && !ctx->gr_count) return ctx_fail(ctx, NNPE_ERR_STATE, "g(r) requested without nnpe_gr_init");
  if (P.sq && !ctx->sq_ftmx) return ctx_fail(ctx, NNPE_ERR_STATE, "S(q) requested without nnpe_sq_init");
  if (P.rmdf && (!rmdf_active(ctx) || !P.sq || P.gr_every <= 0))
    return ctx_fail(ctx, NNPE_ERR_STATE, "RMDF requested without nnpe_rmdf_init, S(q) or g(r) tallies");
  if (P.gr_long_every > 0 && (P.gr_every <= 0 || P.gr_long_every % P.gr_every))
    return ctx_fail(ctx, NNPE_ERR_ARG, "gr_long_every must be a multiple of gr_every");
  if (ctx->dt != P.dt || ctx->fp64 != (P.fp64 != 0)) {
    md_state_free(ctx);
    ctx->dt = P.dt;
    if ((rc0 = verlet_switch_precision(ctx, P.fp64 != 0))) return rc0;
  }
  // log buffer
  const int nrec = P.log_every > 0 ? nsteps / P.log_every + 2 : 1;
  if (nrec > ctx->md_log_cap) {
    md_state_free(ctx);  // graphs hold the old log pointer
    NNPE_CHECK_HIP(ctx, dev_alloc(ctx->md_log, size_t(nrec) * 2 * NNPE_NTERM));
    ctx->md_log_cap = nrec;
  }
  hipLaunchKernelGGL(k_md_reset, dim3(1), dim3(1), 0, ctx->stream, ctx->counters);
  NNPE_CHECK_LAUNCH(ctx);
  MdGraphs* G = graphs_of(ctx);
  int rc;
  // a rebuild interval of the FP64 NNP form, of even length, fuses each step's position update
  // into the force pass of the step before it
  const bool fuse_x = ctx->md_fuse_x && P.nnp == 2 && ctx->fp64 && P.rebuild_every >= 2 && P.rebuild_every % 2 == 0;
  // captures the steps of `keys` (consecutive, starting at the current step) into one graph, or
  // replays the graph of that sequence; fused: the interval with fused position updates (its
  // graph is cached under the keys and a trailing -1)
  auto run_graph = [&](std::vector<int> keys, bool fused) -> int {
    const size_t nk = keys.size();
    if (fused) keys.push_back(-1);
    auto it = G->exec.find(keys);
    if (it == G->exec.end()) {
      hipGraph_t g = nullptr;
      float4* const x0 = ctx->xq;
      float4* const x1 = ctx->xq_alt;
      NNPE_CHECK_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
      int rcs = NNPE_OK;
      for (size_t t = 0; t < nk && !rcs; ++t)
        rcs = enqueue_step(ctx, P, keys[t], false, fused ? (t == 0 ? 1 : (t + 1 == nk ? 3 : 2)) : 0);
      hipError_t ce = hipStreamEndCapture(ctx->stream, &g);
      ctx->xq = x0;
      ctx->xq_alt = x1;
      if (rcs) return rcs;
      NNPE_CHECK_HIP(ctx, ce);
      hipGraphExec_t ex = nullptr;
      NNPE_CHECK_HIP(ctx, hipGraphInstantiate(&ex, g, nullptr, nullptr, 0));
      G->graph[keys] = g;
      it = G->exec.emplace(keys, ex).first;
    }
    NNPE_CHECK_HIP(ctx, hipGraphLaunch(it->second, ctx->stream));
    return NNPE_OK;
  };
  for (int k = 0; k < nsteps;) {
    const long long s = ctx->md_step;
    const bool first = s == 0 || !ctx->have_list;
    const int key = step_key(P, s, first);
    if (P.rebuild_every <= 0) {
      // NEIGHLIST::isvalid every step (src/run.cpp:236-246): host decision, eager launches
      if (!first && (rc = verlet_enqueue_x(ctx))) return rc;
      int valid = 0;
      if (!first && (rc = nnpe_neigh_isvalid(ctx, &valid))) return rc;
      if (!valid) {
        if ((rc = nnpe_neigh_build(ctx))) return rc;
      }
      // the rest of the step without the position update
      if ((rc = enqueue_step(ctx, P, key & ~kRebuild, true))) return rc;
    } else if (first) {
      // list first (with the capacity check and growth), then the step on the positions as read
      if ((rc = nnpe_neigh_build(ctx))) return rc;
      G = graphs_of(ctx);
      if ((rc = enqueue_step(ctx, P, key & ~kRebuild, true))) return rc;
    } else if (P.use_graph && s % P.rebuild_every == 0 && nsteps - k >= P.rebuild_every) {
      // a whole rebuild interval as one graph: no launch gaps between its steps
      std::vector<int> keys(P.rebuild_every);
      for (int t = 0; t < P.rebuild_every; ++t) keys[t] = step_key(P, s + t, false);
      if ((rc = run_graph(keys, fuse_x))) return rc;
      ctx->md_step += P.rebuild_every;
      k += P.rebuild_every;
      continue;
    } else if (P.use_graph) {
      if ((rc = run_graph(std::vector<int>{key}, false))) return rc;
    } else {
      if ((rc = enqueue_step(ctx, P, key, false))) return rc;
    }
    ctx->md_step++;
    ++k;
  }
  if ((rc = counters_read(ctx))) return rc;
  const DevCounters c = *ctx->h_counters;
  if (nlog) *nlog = std::min(c.log_count, ctx->md_log_cap);
  if (log_out && c.log_count > 0) {
    const int n = std::min(std::min(c.log_count, ctx->md_log_cap), log_cap);
    NNPE_CHECK_HIP(ctx, hipMemcpy(log_out, ctx->md_log, sizeof(double) * 2 * NNPE_NTERM * n, hipMemcpyDeviceToHost));
  }
  if (c.cell_range) return ctx_fail(ctx, NNPE_ERR_CELL_RANGE, "atoms left the cell grid during the run");
  if (c.ghost_overflow || c.overflow_long || c.overflow_short) {
    bool grown = false;
    rc = neigh_check_and_grow(ctx, &grown);
    md_state_free(ctx);
    return rc ? rc : ctx_fail(ctx, NNPE_ERR_NEIGH_OVERFLOW, "list or ghost capacity exceeded during the run (grown; rerun)");
  }
  return NNPE_OK;
}

extern "C" int nnpe_md_stats(nnpe_ctx* ctx, long long* out) {
  if (!ctx || !out) return NNPE_ERR_ARG;
  int rc = counters_read(ctx);
  if (rc) return rc;
  const DevCounters c = *ctx->h_counters;
  out[0] = ctx->md_step;
  out[1] = ctx->nbuild;
  out[2] = c.skin_violations;
  out[3] = c.max_long;
  out[4] = c.max_short;
  out[5] = c.nghost;
  out[6] = ctx->cap_long;
  out[7] = ctx->cap_short;
  return NNPE_OK;
}

// Average duration of one stage over `reps` back-to-back launches, HIP events on the context's
// stream (the stream the MD kernels run on).  The stage sees the current state; stages that
// accumulate (forces += ...) are timed as launched, their results are not meant to be used.
extern "C" int nnpe_profile_stage(nnpe_ctx* ctx, int stage, int reps, double* avg_ms) {
  if (!ctx || reps <= 0 || !avg_ms) return NNPE_ERR_ARG;
  if (!ctx->have_list && stage != NNPE_STAGE_COPY && stage != NNPE_STAGE_MFMA)
    return ctx_fail(ctx, NNPE_ERR_STATE, "profile needs a built list");
  float4* probe = nullptr;
  auto launch = [&](int st) -> int {
    switch (st) {
      case NNPE_STAGE_PAIR: return ctx->have_pair ? pair_enqueue(ctx, false) : NNPE_ERR_STATE;
      case NNPE_STAGE_BONDED: return bonded_enqueue(ctx, false);
      case NNPE_STAGE_SF_DESC: return sf_desc_enqueue(ctx);
      case NNPE_STAGE_MLP: return ctx->mlp ? mlp_enqueue(ctx, false) : NNPE_ERR_STATE;
      case NNPE_STAGE_SF_ADJOINT: return sf_adjoint_enqueue(ctx);
      case NNPE_STAGE_SF_FORCE: return sf_force_only_enqueue(ctx, false, false);
      case NNPE_STAGE_REBUILD: return neigh_enqueue_build(ctx);
      case NNPE_STAGE_REBUILD_GR: return neigh_enqueue_build(ctx, true);
      case NNPE_STAGE_GR_FULL: return ctx->gr_count ? gr_enqueue(ctx, true) : NNPE_ERR_STATE;
      case NNPE_STAGE_GR_LIST: return ctx->gr_count ? gr_enqueue(ctx, false) : NNPE_ERR_STATE;
      case NNPE_STAGE_VERLET: {
        int rc = verlet_enqueue_x(ctx);
        return rc ? rc : verlet_enqueue_v(ctx);
      }
      case NNPE_STAGE_SQ: return ctx->sq_ftmx ? sq_enqueue(ctx) : NNPE_ERR_STATE;
      case NNPE_STAGE_TBL_DESC: return nnp_tbl_desc_enqueue(ctx);
      case NNPE_STAGE_TBL_FORCE: return nnp_tbl_force_enqueue(ctx, false);
      case NNPE_STAGE_COPY: {
        const int n = kCopyProbeBytes / 16;
        if (!probe) {
          if (hipMalloc(reinterpret_cast<void**>(&probe), 2 * size_t(kCopyProbeBytes)) != hipSuccess) return NNPE_ERR_HIP;
          if (hipMemsetAsync(probe, 0, 2 * size_t(kCopyProbeBytes), ctx->stream) != hipSuccess) return NNPE_ERR_HIP;
        }
        hipLaunchKernelGGL(k_copy_probe, dim3(2048), dim3(256), 0, ctx->stream, n, probe, probe + n);
        return hipGetLastError() == hipSuccess ? NNPE_OK : NNPE_ERR_HIP;
      }
      case NNPE_STAGE_MFMA: {
        static_assert(size_t(kMfmaProbeBlocks) * 256 * sizeof(float) <= 2 * size_t(kCopyProbeBytes), "probe output");
        if (!probe && hipMalloc(reinterpret_cast<void**>(&probe), 2 * size_t(kCopyProbeBytes)) != hipSuccess) return NNPE_ERR_HIP;
        hipLaunchKernelGGL(k_mfma_probe, dim3(kMfmaProbeBlocks), dim3(256), 0, ctx->stream, reinterpret_cast<float*>(probe),
                           1.0f / 64.0f);
        return hipGetLastError() == hipSuccess ? NNPE_OK : NNPE_ERR_HIP;
      }
      default: return NNPE_ERR_ARG;
    }
  };
  int rc = launch(stage);  // warm-up (and argument check)
  if (rc) return rc == NNPE_ERR_STATE ? ctx_fail(ctx, rc, "stage not configured") : rc;
  NNPE_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
  // the reps launches are captured into one graph and timed as one replay, so the average is
  // the kernels' own duration (as a kernel trace reports it) without per-launch host gaps;
  // eager back-to-back launches where the stage cannot be captured
  hipGraph_t g = nullptr;
  hipGraphExec_t gx = nullptr;
  if (hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal) == hipSuccess) {
    int rcc = 0;
    for (int r = 0; r < reps && !rcc; ++r) rcc = launch(stage);
    const hipError_t ce = hipStreamEndCapture(ctx->stream, &g);
    if (rcc || ce != hipSuccess || hipGraphInstantiate(&gx, g, nullptr, nullptr, 0) != hipSuccess) gx = nullptr;
    (void)hipGetLastError();
  }
  if (gx) NNPE_CHECK_HIP(ctx, hipGraphLaunch(gx, ctx->stream));  // warm replay
  hipEvent_t e0, e1;
  NNPE_CHECK_HIP(ctx, hipEventCreate(&e0));
  NNPE_CHECK_HIP(ctx, hipEventCreate(&e1));
  NNPE_CHECK_HIP(ctx, hipEventRecord(e0, ctx->stream));
  if (gx) {
    NNPE_CHECK_HIP(ctx, hipGraphLaunch(gx, ctx->stream));
  } else {
    for (int r = 0; r < reps && !rc; ++r) rc = launch(stage);
  }
  NNPE_CHECK_HIP(ctx, hipEventRecord(e1, ctx->stream));
  NNPE_CHECK_HIP(ctx, hipEventSynchronize(e1));
  if (gx) (void)hipGraphExecDestroy(gx);
  if (g) (void)hipGraphDestroy(g);
  float ms = 0;
  NNPE_CHECK_HIP(ctx, hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (probe) (void)hipFree(probe);
  if (rc) return rc;
  *avg_ms = double(ms) / reps;
  return NNPE_OK;
}

extern "C" int nnpe_md_reset_step(nnpe_ctx* ctx) {
  if (!ctx) return NNPE_ERR_ARG;
  ctx->md_step = 0;
  return NNPE_OK;
}

extern "C" int nnpe_md_set_step(nnpe_ctx* ctx, long long step) {
  if (!ctx || step < 0) return NNPE_ERR_ARG;
  ctx->md_step = step;
  return NNPE_OK;
}
