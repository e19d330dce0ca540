// ctx.hip -- context lifecycle, system upload (ENVIRON::initialize), per-atom state access and
// the energy/virial accumulators.
#include <algorithm>
#include <cstring>

#include <cstdlib>

#include "internal.h"

int ctx_fail(nnpe_ctx* ctx, int code, const std::string& msg) {
  if (ctx) ctx->err = msg;
  return code;
}

extern "C" const char* nnpe_status_string(int s) {
  switch (s) {
    case NNPE_OK: return "ok";
    case NNPE_ERR_ARG: return "bad argument";
    case NNPE_ERR_HIP: return "HIP runtime error";
    case NNPE_ERR_STATE: return "call out of order";
    case NNPE_ERR_NEIGH_OVERFLOW: return "nei_list overflow";
    case NNPE_ERR_GHOST_OVERFLOW: return "ghost overflow";
    case NNPE_ERR_CELL_RANGE: return "atom outside the cell grid";
    case NNPE_ERR_NO_DEVICE: return "no HIP device";
    case NNPE_ERR_UNSUPPORTED: return "unsupported configuration";
    default: return "unknown status";
  }
}

extern "C" const char* nnpe_last_error(const nnpe_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

extern "C" int nnpe_create(nnpe_ctx** out, int device) {
  if (!out) return NNPE_ERR_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= device || device < 0) return NNPE_ERR_NO_DEVICE;
  nnpe_ctx* ctx = new nnpe_ctx;
  ctx->device = device;
  NNPE_CHECK_HIP(ctx, hipSetDevice(device));
  NNPE_CHECK_HIP(ctx, hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
  NNPE_CHECK_HIP(ctx, hipStreamCreateWithFlags(&ctx->stream2, hipStreamNonBlocking));
  if (const char* e = std::getenv("NNPE_MD_FORK")) ctx->md_fork = std::atoi(e) != 0;
  if (const char* e = std::getenv("NNPE_MD_FORK_EARLY")) ctx->md_fork_early = std::atoi(e) != 0;
  if (const char* e = std::getenv("NNPE_MD_FUSE_X")) ctx->md_fuse_x = std::atoi(e) != 0;
  NNPE_CHECK_HIP(ctx, hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming));
  NNPE_CHECK_HIP(ctx, hipEventCreateWithFlags(&ctx->ev_join, hipEventDisableTiming));
  NNPE_CHECK_HIP(ctx, hipMalloc(reinterpret_cast<void**>(&ctx->counters), sizeof(DevCounters)));
  NNPE_CHECK_HIP(ctx, hipHostMalloc(reinterpret_cast<void**>(&ctx->h_counters), sizeof(DevCounters), hipHostMallocDefault));
  NNPE_CHECK_HIP(ctx, hipMemset(ctx->counters, 0, sizeof(DevCounters)));
  NNPE_CHECK_HIP(ctx, dev_alloc(ctx->ev_sum, 2 * NNPE_NTERM));
  NNPE_CHECK_HIP(ctx, hipMemset(ctx->ev_sum, 0, sizeof(double) * 2 * NNPE_NTERM));
  NNPE_CHECK_HIP(ctx, hipHostMalloc(reinterpret_cast<void**>(&ctx->h_ev), sizeof(double) * 2 * NNPE_NTERM, hipHostMallocDefault));
  const int rcs = bonded_scratch_setup(ctx);  // the allocation-free bonded entry point's device array
  if (rcs) {
    nnpe_destroy(ctx);
    return rcs;
  }
  *out = ctx;
  return NNPE_OK;
}

static void free_all(nnpe_ctx* c) {
  dev_free(c->xq); dev_free(c->xq_alt); dev_free(c->v); dev_free(c->f); dev_free(c->fprev); dev_free(c->xprev);
  dev_free(c->xd); dev_free(c->vd); dev_free(c->ad);
  dev_free(c->rm); dev_free(c->typ); dev_free(c->info); dev_free(c->inv); dev_free(c->cell); dev_free(c->xmol_code);
  dev_free(c->t_xq); dev_free(c->t_v); dev_free(c->t_fprev); dev_free(c->t_xprev);
  dev_free(c->t_xd); dev_free(c->t_vd); dev_free(c->t_ad);
  dev_free(c->t_rm); dev_free(c->t_typ); dev_free(c->t_cell); dev_free(c->order); dev_free(c->t_info);
  dev_free(c->mol_first); dev_free(c->mt_natom); dev_free(c->mt_special_off); dev_free(c->mt_special);
  dev_free(c->rbox); dev_free(c->t_rbox); dev_free(c->d_box); dev_free(c->cell_box); dev_free(c->home_cells);
  dev_free(c->cell_start); dev_free(c->cell_cnt); dev_free(c->cell_fill);
  dev_free(c->gcell_start); dev_free(c->gcell_cnt); dev_free(c->gcell_fill);
  dev_free(c->g_src); dev_free(c->g_code); dev_free(c->g_cell); dev_free(c->gs_src); dev_free(c->gs_code);
  dev_free(c->stencil_pair); dev_free(c->stencil_gr); dev_free(c->stencil_gr_outer);
  dev_free(c->nbr_long); dev_free(c->nbr_short); dev_free(c->cnt_long); dev_free(c->cnt_short);
  dev_free(c->scan_tmp); dev_free(c->ev_partial);
  dev_free(c->pt_tab);
  for (int k = 0; k < 3; ++k) {
    dev_free(c->bd_coef[k]); dev_free(c->bd_type[k]); dev_free(c->bd_atoms[k]);
    dev_free(c->bd_start[k]); dev_free(c->bd_ref[k]);
    c->bd_start_len[k] = 0;
  }
  dev_free(c->gr_count); dev_free(c->gr_shell); dev_free(c->gr_h); dev_free(c->gr_norm); dev_free(c->sq_out); dev_free(c->sq_ftmx);
  md_state_free(c);
  rmdf_state_free(c);
  sqd_state_free(c);
  nnp_state_free(c);
  mlp_state_free(c);
  dev_free(c->G); dev_free(c->dG); dev_free(c->e_atom); dev_free(c->type_idx); dev_free(c->type_fill);
  c->D_cur = c->D_cap = 0;
}

extern "C" int nnpe_destroy(nnpe_ctx* ctx) {
  if (!ctx) return NNPE_OK;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  free_all(ctx);
  dev_free(ctx->counters);
  dev_free(ctx->ev_sum);
  if (ctx->h_counters) (void)hipHostFree(ctx->h_counters);
  if (ctx->h_ev) (void)hipHostFree(ctx->h_ev);
  if (ctx->stream2) (void)hipStreamSynchronize(ctx->stream2);
  if (ctx->ev_fork) (void)hipEventDestroy(ctx->ev_fork);
  if (ctx->ev_join) (void)hipEventDestroy(ctx->ev_join);
  if (ctx->stream2) (void)hipStreamDestroy(ctx->stream2);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return NNPE_OK;
}

extern "C" int nnpe_synchronize(nnpe_ctx* ctx) {
  if (!ctx) return NNPE_ERR_ARG;
  NNPE_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return NNPE_OK;
}

extern "C" int nnpe_system_set(nnpe_ctx* ctx, int natom, int ntype, float bl, const float* x, const float* v,
                               const float* q, const float* mass, const int* typ, const int* mol_id,
                               const int* intra_id, int nmol, const int* mol_type, int nmoltype,
                               const int* mt_natom, const int* mt_special_off, const int* mt_special) {
  if (!ctx || natom <= 0 || natom >= NNPE_MAX_ATOMS || ntype <= 0 || !(bl > 0) || !x || !q || !mass || !typ ||
      !mol_id || !intra_id || nmol <= 0 || !mol_type || nmoltype <= 0 || !mt_natom || !mt_special_off || !mt_special)
    return ctx_fail(ctx, NNPE_ERR_ARG, "nnpe_system_set: bad argument");
  NNPE_CHECK_HIP(ctx, hipSetDevice(ctx->device));
  NNPE_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
  free_all(ctx);
  ctx->have_grid = ctx->have_list = ctx->have_pair = false;
  ctx->fp64 = false;
  ctx->md_step = 0;
  for (int k = 0; k < 3; ++k) ctx->bd_nterm[k] = ctx->bd_ntyp[k] = 0;
  ctx->natom = natom; ctx->ntype = ntype; ctx->npair = ntype * (ntype + 1) / 2;
  ctx->nmol = nmol; ctx->nmoltype = nmoltype;
  ctx->stride = (natom + 63) / 64 * 64;
  // ENVIRON::set_bl (src/environ.h:105)
  ctx->bl = bl; ctx->r_bl = 1.0f / bl; ctx->h_bl = bl * .5f; ctx->m_bl = -bl;
  ctx->h_typecount.assign(ntype, 0);
  for (int i = 0; i < natom; ++i) {
    if (typ[i] < 0 || typ[i] >= ntype || mol_id[i] < 0 || mol_id[i] >= nmol) return ctx_fail(ctx, NNPE_ERR_ARG, "atom type / molecule id out of range");
    ++ctx->h_typecount[typ[i]];
  }
  const size_t N = natom;
  std::vector<float4> hx(N), hv(N), hz(N, make_float4(0, 0, 0, 0));
  std::vector<float> hrm(N);
  std::vector<int4> hinfo(N);
  std::vector<int> hinv(N), hmf(nmol, -1);
  for (size_t i = 0; i < N; ++i) {
    hx[i] = make_float4(x[3 * i], x[3 * i + 1], x[3 * i + 2], q[i]);
    hv[i] = v ? make_float4(v[3 * i], v[3 * i + 1], v[3 * i + 2], 0) : make_float4(0, 0, 0, 0);
    hrm[i] = 1.0f / mass[i];
    hinfo[i] = make_int4(int(i), mol_id[i], intra_id[i], mol_type[mol_id[i]]);
    hinv[i] = int(i);
    if (intra_id[i] == 0) hmf[mol_id[i]] = int(i);
  }
  for (int m = 0; m < nmol; ++m) if (hmf[m] < 0) return ctx_fail(ctx, NNPE_ERR_ARG, "molecule without an intra-0 atom");
  int sp_total = 0;
  ctx->h_mt_natom.assign(mt_natom, mt_natom + nmoltype);
  for (int t = 0; t < nmoltype; ++t) sp_total = std::max(sp_total, mt_special_off[t] + mt_natom[t] * mt_natom[t]);
  auto up = [&](auto*& d, const auto* h, size_t n) -> hipError_t {
    hipError_t e = dev_alloc(d, n);
    if (e != hipSuccess) return e;
    return hipMemcpy(d, h, n * sizeof(*h), hipMemcpyHostToDevice);
  };
  NNPE_CHECK_HIP(ctx, up(ctx->xq, hx.data(), N));
  NNPE_CHECK_HIP(ctx, up(ctx->xq_alt, hx.data(), N));
  NNPE_CHECK_HIP(ctx, up(ctx->v, hv.data(), N));
  NNPE_CHECK_HIP(ctx, up(ctx->f, hz.data(), N));
  NNPE_CHECK_HIP(ctx, up(ctx->fprev, hz.data(), N));
  NNPE_CHECK_HIP(ctx, up(ctx->xprev, hx.data(), N));
  NNPE_CHECK_HIP(ctx, up(ctx->rm, hrm.data(), N));
  NNPE_CHECK_HIP(ctx, up(ctx->typ, typ, N));
  NNPE_CHECK_HIP(ctx, up(ctx->info, hinfo.data(), N));
  NNPE_CHECK_HIP(ctx, up(ctx->inv, hinv.data(), N));
  NNPE_CHECK_HIP(ctx, up(ctx->mol_first, hmf.data(), size_t(nmol)));
  NNPE_CHECK_HIP(ctx, up(ctx->mt_natom, mt_natom, size_t(nmoltype)));
  NNPE_CHECK_HIP(ctx, up(ctx->mt_special_off, mt_special_off, size_t(nmoltype)));
  NNPE_CHECK_HIP(ctx, up(ctx->mt_special, mt_special, size_t(sp_total)));
  std::vector<int> hcode(N, 13);
  NNPE_CHECK_HIP(ctx, up(ctx->xmol_code, hcode.data(), N));
  NNPE_CHECK_HIP(ctx, dev_alloc(ctx->cell, N));
  NNPE_CHECK_HIP(ctx, hipMemset(ctx->cell, 0, N * sizeof(int)));
  NNPE_CHECK_HIP(ctx, dev_alloc(ctx->t_xq, N)); NNPE_CHECK_HIP(ctx, dev_alloc(ctx->t_v, N));
  NNPE_CHECK_HIP(ctx, dev_alloc(ctx->t_fprev, N)); NNPE_CHECK_HIP(ctx, dev_alloc(ctx->t_xprev, N));
  NNPE_CHECK_HIP(ctx, dev_alloc(ctx->t_rm, N)); NNPE_CHECK_HIP(ctx, dev_alloc(ctx->t_typ, N));
  NNPE_CHECK_HIP(ctx, dev_alloc(ctx->t_cell, N)); NNPE_CHECK_HIP(ctx, dev_alloc(ctx->order, N));
  NNPE_CHECK_HIP(ctx, dev_alloc(ctx->t_info, N));
  NNPE_CHECK_HIP(ctx, dev_alloc(ctx->rbox, N)); NNPE_CHECK_HIP(ctx, dev_alloc(ctx->t_rbox, N));
  NNPE_CHECK_HIP(ctx, hipMemset(ctx->rbox, 0, sizeof(int) * N));
  ctx->nbox = 1;
  // FP64 integration state starts from the FP32 values
  std::vector<double4> hxd(N), hvd(N), had(N, make_double4(0, 0, 0, 0));
  for (size_t i = 0; i < N; ++i) {
    hxd[i] = make_double4(hx[i].x, hx[i].y, hx[i].z, 0);
    hvd[i] = make_double4(hv[i].x, hv[i].y, hv[i].z, 0);
  }
  NNPE_CHECK_HIP(ctx, up(ctx->xd, hxd.data(), N)); NNPE_CHECK_HIP(ctx, up(ctx->vd, hvd.data(), N));
  NNPE_CHECK_HIP(ctx, up(ctx->ad, had.data(), N));
  NNPE_CHECK_HIP(ctx, dev_alloc(ctx->t_xd, N)); NNPE_CHECK_HIP(ctx, dev_alloc(ctx->t_vd, N));
  NNPE_CHECK_HIP(ctx, dev_alloc(ctx->t_ad, N));
  // partial sums: one per 256-thread block at up to 64 lanes per atom
  NNPE_CHECK_HIP(ctx, ev_reserve(ctx, blocks_for((long long)N * 64, 256) + 64));
  NNPE_CHECK_HIP(ctx, hipMemset(ctx->ev_sum, 0, sizeof(double) * 2 * NNPE_NTERM));
  ctx->have_system = true;
  return NNPE_OK;
}

// ---- energy / virial ---------------------------------------------------------------------------
hipError_t ev_reserve(nnpe_ctx* ctx, int nblocks) {
  if (nblocks <= ctx->ev_max_blocks) return hipSuccess;
  ctx->ev_max_blocks = nblocks;
  return dev_alloc(ctx->ev_partial, size_t(NNPE_NTERM) * 2 * nblocks);
}
double* ev_part(nnpe_ctx* ctx, int term, int which) {
  return ctx->ev_partial + (size_t(term) * 2 + which) * ctx->ev_max_blocks;
}

__global__ void k_ev_finish(const double* part_e, const double* part_w, int nblocks, double* sum) {
  // fixed-order reduction: one wave, strided partial sums then a tree
  double a = 0, b = 0;
  for (int k = threadIdx.x; k < nblocks; k += 64) { a += part_e[k]; b += part_w[k]; }
  for (int o = 32; o > 0; o >>= 1) { a += __shfl_down(a, o, 64); b += __shfl_down(b, o, 64); }
  if (threadIdx.x == 0) { sum[0] += a; sum[1] += b; }
}

int ev_finish(nnpe_ctx* ctx, int term, int nblocks) {
  return ev_finish_from(ctx, term, ev_part(ctx, term, 0), ev_part(ctx, term, 1), nblocks);
}

int ev_finish_from(nnpe_ctx* ctx, int term, const double* part_e, const double* part_w, int nblocks) {
  hipLaunchKernelGGL(k_ev_finish, dim3(1), dim3(64), 0, ctx->stream, part_e, part_w, nblocks, ctx->ev_sum + 2 * term);
  NNPE_CHECK_LAUNCH(ctx);
  return NNPE_OK;
}

extern "C" int nnpe_energy_get(nnpe_ctx* ctx, double* out) {
  if (!ctx || !out) return NNPE_ERR_ARG;
  NNPE_CHECK_HIP(ctx, hipMemcpyAsync(ctx->h_ev, ctx->ev_sum, sizeof(double) * 2 * NNPE_NTERM, hipMemcpyDeviceToHost, ctx->stream));
  NNPE_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
  std::memcpy(out, ctx->h_ev, sizeof(double) * 2 * NNPE_NTERM);
  return NNPE_OK;
}

__global__ void k_zero_f64(double* p, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

extern "C" int nnpe_energy_clear(nnpe_ctx* ctx) {
  if (!ctx) return NNPE_ERR_ARG;
  hipLaunchKernelGGL(k_zero_f64, dim3(1), dim3(64), 0, ctx->stream, ctx->ev_sum, 2 * NNPE_NTERM);
  NNPE_CHECK_LAUNCH(ctx);
  return NNPE_OK;
}

int counters_read(nnpe_ctx* ctx) {
  NNPE_CHECK_HIP(ctx, hipMemcpyAsync(ctx->h_counters, ctx->counters, sizeof(DevCounters), hipMemcpyDeviceToHost, ctx->stream));
  NNPE_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return NNPE_OK;
}

// ---- per-atom state in atom-id order --------------------------------------------------------
__global__ void k_state_get(int what, int n, bool fp64, const int4* info, const float4* xq, const float4* v,
                            const float4* f, const float4* fprev, const double4* xd, const double4* vd,
                            const double4* ad, double* out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int a = info[i].x;
  double x0, x1, x2;
  if (what == NNPE_X) {
    if (fp64) { double4 t = xd[i]; x0 = t.x; x1 = t.y; x2 = t.z; }
    else { float4 t = xq[i]; x0 = t.x; x1 = t.y; x2 = t.z; }
  } else if (what == NNPE_V) {
    if (fp64) { double4 t = vd[i]; x0 = t.x; x1 = t.y; x2 = t.z; }
    else { float4 t = v[i]; x0 = t.x; x1 = t.y; x2 = t.z; }
  } else if (what == NNPE_F) {
    float4 t = f[i]; x0 = t.x; x1 = t.y; x2 = t.z;
  } else {
    if (fp64) { double4 t = ad[i]; x0 = t.x; x1 = t.y; x2 = t.z; }
    else { float4 t = fprev[i]; x0 = t.x; x1 = t.y; x2 = t.z; }
  }
  out[3 * a] = x0; out[3 * a + 1] = x1; out[3 * a + 2] = x2;
}

__global__ void k_state_set(int what, int n, bool fp64, const int4* info, float4* xq, float4* v, float4* f,
                            float4* fprev, double4* xd, double4* vd, double4* ad, const double* in) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int a = info[i].x;
  double x0 = in[3 * a], x1 = in[3 * a + 1], x2 = in[3 * a + 2];
  if (what == NNPE_X) {
    xq[i].x = float(x0); xq[i].y = float(x1); xq[i].z = float(x2);
    xd[i] = make_double4(fp64 ? x0 : double(float(x0)), fp64 ? x1 : double(float(x1)), fp64 ? x2 : double(float(x2)), 0);
  } else if (what == NNPE_V) {
    v[i] = make_float4(float(x0), float(x1), float(x2), 0);
    vd[i] = make_double4(x0, x1, x2, 0);
  } else if (what == NNPE_F) {
    f[i] = make_float4(float(x0), float(x1), float(x2), 0);
  } else {
    fprev[i] = make_float4(float(x0), float(x1), float(x2), 0);
    ad[i] = make_double4(x0, x1, x2, 0);
  }
}

extern "C" int nnpe_state_get(nnpe_ctx* ctx, int what, double* out) {
  if (!ctx || !out || what < 0 || what > 3) return ctx_fail(ctx, NNPE_ERR_ARG, "nnpe_state_get: bad argument");
  if (!ctx->have_system) return ctx_fail(ctx, NNPE_ERR_STATE, "no system");
  double* d = nullptr;
  NNPE_CHECK_HIP(ctx, hipMallocAsync(reinterpret_cast<void**>(&d), sizeof(double) * 3 * ctx->natom, ctx->stream));
  hipLaunchKernelGGL(k_state_get, dim3(blocks_for(ctx->natom, 256)), dim3(256), 0, ctx->stream, what, ctx->natom,
                     ctx->fp64, ctx->info, ctx->xq, ctx->v, ctx->f, ctx->fprev, ctx->xd, ctx->vd, ctx->ad, d);
  NNPE_CHECK_LAUNCH(ctx);
  NNPE_CHECK_HIP(ctx, hipMemcpyAsync(out, d, sizeof(double) * 3 * ctx->natom, hipMemcpyDeviceToHost, ctx->stream));
  NNPE_CHECK_HIP(ctx, hipFreeAsync(d, ctx->stream));
  NNPE_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return NNPE_OK;
}

extern "C" int nnpe_state_set(nnpe_ctx* ctx, int what, const double* in) {
  if (!ctx || !in || what < 0 || what > 3) return ctx_fail(ctx, NNPE_ERR_ARG, "nnpe_state_set: bad argument");
  if (!ctx->have_system) return ctx_fail(ctx, NNPE_ERR_STATE, "no system");
  double* d = nullptr;
  NNPE_CHECK_HIP(ctx, hipMallocAsync(reinterpret_cast<void**>(&d), sizeof(double) * 3 * ctx->natom, ctx->stream));
  NNPE_CHECK_HIP(ctx, hipMemcpyAsync(d, in, sizeof(double) * 3 * ctx->natom, hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(k_state_set, dim3(blocks_for(ctx->natom, 256)), dim3(256), 0, ctx->stream, what, ctx->natom,
                     ctx->fp64, ctx->info, ctx->xq, ctx->v, ctx->f, ctx->fprev, ctx->xd, ctx->vd, ctx->ad, d);
  NNPE_CHECK_LAUNCH(ctx);
  NNPE_CHECK_HIP(ctx, hipFreeAsync(d, ctx->stream));
  NNPE_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return NNPE_OK;
}
