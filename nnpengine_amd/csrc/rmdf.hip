// rmdf.hip -- reverse-MD / FMIRL refinement of the pair tables against experimental S(q)
// (RMDF, src/grsq.cpp:899-1377), on the device after each g(r) -> S(q).
//
// Set-up (host, the reference's FP32 formulas): per data set n and q point, the weights that
// combine the partial S_ab(q) into the measured total (sqwt) and that map the residual back onto
// each pair (fwt = form-factor weight x strength x data weight), and the kernels
// ftmx[r][q] = (j0(qr) - cos(qr)) / r^2 (force table) and potmx[r][q] = j0(qr) (potential).
// Per refinement step:
//   k_rmdf_q   one thread per q: S_n(q) = sum_ab S_ab(q) sqwt, residual energy, the relaxed
//              residual deltasq += gamma (S_n - S_exp - deltasq), force_q[ab][q] = sum_n deltasq fwt
//   k_rmdf_tab one thread per (pair, r bin): table = base + sum_q force_q[ab][q] kernel[r][q]
//              (FP64 accumulation); for the NNP pair tables the update goes to special class 0
//              and, scaled by the 1-4 factor 0.5, to class 3 (src/grsq.cpp:1213-1297)
// The pair force kernel reads the refined table on the next step.
#include <cmath>

#include "internal.h"

namespace {

constexpr int kMaxPair = 10;  // four elements

// one block: S_n(q), the relaxed residual and force_q for every q; the residual energy
// erg * natom / 4 is stored (not accumulated) into its energy slot, so a log record carries
// the value of the last refinement step as the reference's erg_exp does
__global__ void __launch_bounds__(256) k_rmdf_q(int nsq, int nq, int npair, const double* __restrict__ psq,
                                                const float* __restrict__ sqwt, const float* __restrict__ fwt,
                                                const float* __restrict__ sqex, const float* __restrict__ wt, float gamma,
                                                double escale, float* __restrict__ sq, float* __restrict__ deltasq,
                                                float* __restrict__ force_q, double* __restrict__ e_slot) {
  double erg = 0;
  for (int i = threadIdx.x; i < nq; i += 256) {
    float fq[kMaxPair];
#pragma unroll
    for (int p = 0; p < kMaxPair; ++p) fq[p] = 0.0f;
    for (int n = 0; n < nsq; ++n) {
      float s = 0.0f;
      for (int p = 0; p < npair; ++p) s += float(psq[(size_t)p * nq + i]) * sqwt[((size_t)n * npair + p) * nq + i];
      sq[(size_t)n * nq + i] = s;
      const float d = s - sqex[(size_t)n * nq + i];
      erg += double(d) * d * wt[(size_t)n * nq + i];
      float ds = deltasq[(size_t)n * nq + i];
      ds += gamma * (d - ds);
      deltasq[(size_t)n * nq + i] = ds;
#pragma unroll
      for (int p = 0; p < kMaxPair; ++p)
        if (p < npair) fq[p] += ds * fwt[((size_t)n * npair + p) * nq + i];
    }
#pragma unroll
    for (int p = 0; p < kMaxPair; ++p)
      if (p < npair) force_q[(size_t)p * nq + i] = fq[p];
  }
  __shared__ double red[4];
  for (int o = 32; o > 0; o >>= 1) erg += __shfl_down(erg, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = erg;
  __syncthreads();
  if (threadIdx.x == 0) {
    e_slot[0] = (red[0] + red[1] + red[2] + red[3]) * escale;
    e_slot[1] = 0.0;
  }
}

// one wave per r bin, lanes over q (coalesced rows of the kernel matrix), all pairs at once:
// table[p][r] = base[p][r] + sum_q force_q[p][q] kern[r][q] (FP64 lane sums, fixed-order
// butterfly).  nnp = 0: rows [npair][nb1]; nnp = 1: rows [npair][4][nb1], special class 0
// takes the update and class 3 half of it.
__global__ void __launch_bounds__(256) k_rmdf_tab(int npair, int nr, int nq, int nb1, int nnp,
                                                  const float* __restrict__ force_q, const float* __restrict__ kern,
                                                  const float* __restrict__ base, float* __restrict__ tab) {
  const int r = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (r >= nr) return;  // whole wave
  double acc[kMaxPair];
#pragma unroll
  for (int p = 0; p < kMaxPair; ++p) acc[p] = 0.0;
  const float* kr = kern + (size_t)r * nq;
  for (int q = lane; q < nq; q += 64) {
    const double k = kr[q];
#pragma unroll
    for (int p = 0; p < kMaxPair; ++p)
      if (p < npair) acc[p] += double(force_q[(size_t)p * nq + q]) * k;
  }
#pragma unroll
  for (int p = 0; p < kMaxPair; ++p)
    for (int o = 32; o > 0; o >>= 1) acc[p] += __shfl_xor(acc[p], o, 64);
  if (lane == 0) {
#pragma unroll
    for (int p = 0; p < kMaxPair; ++p) {
      if (p >= npair) break;
      if (!nnp) {
        tab[(size_t)p * nb1 + r] = float(double(base[(size_t)p * nb1 + r]) + acc[p]);
      } else {
        tab[((size_t)p * 4 + 0) * nb1 + r] = float(acc[p] + base[((size_t)p * 4 + 0) * nb1 + r]);
        tab[((size_t)p * 4 + 3) * nb1 + r] = float(acc[p] * 0.5 + base[((size_t)p * 4 + 3) * nb1 + r]);
      }
    }
  }
}

}  // namespace

struct RmdfState {
  int nsq = 0, nq = 0, nr = 0, target = 0, nb1 = 0;
  float gamma = 1;
  float *sqwt = nullptr, *fwt = nullptr, *sqex = nullptr, *wt = nullptr, *ftmx = nullptr, *potmx = nullptr;
  float *sq = nullptr, *deltasq = nullptr, *force_q = nullptr, *fbase = nullptr, *ubase = nullptr;
  float *ftab = nullptr, *utab = nullptr;  // the tables being refined (owned by the pair / NNP state)
};

static RmdfState* rmdf_of(nnpe_ctx* ctx) { return ctx->rmdf; }

void rmdf_state_free(nnpe_ctx* ctx) {
  RmdfState* R = rmdf_of(ctx);
  if (!R) return;
  dev_free(R->sqwt); dev_free(R->fwt); dev_free(R->sqex); dev_free(R->wt); dev_free(R->ftmx); dev_free(R->potmx);
  dev_free(R->sq); dev_free(R->deltasq); dev_free(R->force_q); dev_free(R->fbase); dev_free(R->ubase);
  delete R;
  ctx->rmdf = nullptr;
}

extern "C" int nnpe_rmdf_init(nnpe_ctx* ctx, int nsq, const float* data, float strength, float gamma, int target) {
  if (ctx) md_state_free(ctx);
  if (!ctx || nsq <= 0 || nsq > 8 || !data || (target != 0 && target != 1)) return ctx_fail(ctx, NNPE_ERR_ARG, "nnpe_rmdf_init: bad argument (1-8 data sets)");
  if (ctx->npair > kMaxPair) return ctx_fail(ctx, NNPE_ERR_UNSUPPORTED, "RMDF supports up to four elements");
  if (!(gamma >= 0 && gamma <= 1)) return ctx_fail(ctx, NNPE_ERR_ARG, "invalid gamma");
  if (!ctx->sq_ftmx || ctx->sq_qmin != 0.0) return ctx_fail(ctx, NNPE_ERR_STATE, "RMDF needs nnpe_sq_init with q_min = 0");
  // the table grid must be the g(r) grid (src/grsq.cpp:924-946, "RMDF mismatch")
  float tab_dr;
  int tab_nbin;
  float *ftab, *utab;
  if (target == 0) {
    if (!ctx->have_pair) return ctx_fail(ctx, NNPE_ERR_STATE, "RMDF on the pair table needs nnpe_pair_table");
    tab_nbin = ctx->pt_nbin; tab_dr = ctx->pt_cut / ctx->pt_nbin;
    ftab = ctx->pt_tab; utab = ctx->pt_tab + size_t(ctx->npair) * (tab_nbin + 1);
  } else {
    if (nnp_tbl_tables(ctx, &ftab, &utab, &tab_nbin, &tab_dr)) return ctx_fail(ctx, NNPE_ERR_STATE, "RMDF on NNP tables needs nnpe_nnp_tbl_set");
  }
  if (tab_nbin != ctx->gr_nbin_inner || std::fabs(tab_dr - ctx->gr_dr) > 1e-5f) return ctx_fail(ctx, NNPE_ERR_ARG, "RMDF mismatch");
  rmdf_state_free(ctx);
  RmdfState* R = new RmdfState;
  ctx->rmdf = R;
  const int nq = ctx->sq_nq, nt = ctx->ntype, np = ctx->npair, nr = ctx->gr_nbin_inner;
  R->nsq = nsq; R->nq = nq; R->nr = nr; R->target = target; R->gamma = gamma; R->nb1 = tab_nbin + 1;
  R->ftab = ftab; R->utab = utab;
  std::vector<float> sqwt(size_t(nsq) * np * nq), fwt(sqwt.size()), sqex(size_t(nsq) * nq), wt(sqex.size());
  const float natsq = float(ctx->natom) * float(ctx->natom);
  std::vector<float> ff(nt);
  for (int n = 0; n < nsq; ++n)
    for (int i = 0; i < nq; ++i) {
      const float* row = data + (size_t(n) * nq + i) * (3 + nt);
      sqex[size_t(n) * nq + i] = row[1];
      const float wti = row[2] * strength;
      wt[size_t(n) * nq + i] = wti;
      float cifi2 = 0;
      for (int j = 0; j < nt; ++j) { ff[j] = row[3 + j]; cifi2 += ff[j] * float(ctx->h_typecount[j]); }
      cifi2 = cifi2 / float(ctx->natom);
      cifi2 = cifi2 * cifi2;
      int ipr = 0;
      for (int j = 0; j < nt; ++j) {
        const float fi = ff[j], ni = float(ctx->h_typecount[j]) / natsq;
        for (int k = j; k < nt; ++k, ++ipr) {
          const float fw = fi * ff[k] / cifi2;
          sqwt[(size_t(n) * np + ipr) * nq + i] = fw * float(ctx->h_typecount[k]) * ni * float(k == j ? 1 : 2);
          fwt[(size_t(n) * np + ipr) * nq + i] = fw * wti;
        }
      }
    }
  std::vector<float> ftmx(size_t(nr) * nq), potmx(ftmx.size());
  const float dq = float(ctx->sq_qmax) / float(nq);
  for (int r = 0; r < nr; ++r) {
    const float rri = (float(r) + 0.5f) * ctx->gr_dr, fi = 1.0f / (rri * rri);
    for (int j = 0; j < nq; ++j) {
      const float qi = (float(j) + .5f) * dq * rri;
      const float j0 = std::sin(qi) / qi;
      potmx[size_t(r) * nq + j] = j0;
      ftmx[size_t(r) * nq + j] = fi * (j0 - std::cos(qi));
    }
  }
  auto up = [&](float*& d, const std::vector<float>& h) -> hipError_t {
    hipError_t e = dev_alloc(d, h.size());
    return e != hipSuccess ? e : hipMemcpy(d, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice);
  };
  NNPE_CHECK_HIP(ctx, up(R->sqwt, sqwt)); NNPE_CHECK_HIP(ctx, up(R->fwt, fwt));
  NNPE_CHECK_HIP(ctx, up(R->sqex, sqex)); NNPE_CHECK_HIP(ctx, up(R->wt, wt));
  NNPE_CHECK_HIP(ctx, up(R->ftmx, ftmx)); NNPE_CHECK_HIP(ctx, up(R->potmx, potmx));
  std::vector<float> z(size_t(nsq) * nq, 0.0f);
  NNPE_CHECK_HIP(ctx, up(R->sq, z)); NNPE_CHECK_HIP(ctx, up(R->deltasq, z));
  NNPE_CHECK_HIP(ctx, dev_alloc(R->force_q, size_t(np) * nq));
  const size_t tab_n = size_t(np) * (target ? 4 : 1) * R->nb1;
  NNPE_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
  NNPE_CHECK_HIP(ctx, dev_alloc(R->fbase, tab_n)); NNPE_CHECK_HIP(ctx, dev_alloc(R->ubase, tab_n));
  NNPE_CHECK_HIP(ctx, hipMemcpy(R->fbase, ftab, tab_n * sizeof(float), hipMemcpyDeviceToDevice));
  NNPE_CHECK_HIP(ctx, hipMemcpy(R->ubase, utab, tab_n * sizeof(float), hipMemcpyDeviceToDevice));
  return NNPE_OK;
}

// RMDF::loadDeltaSQ (src/grsq.cpp:1419-1443): resume the relaxed residual [nsq][nq]
extern "C" int nnpe_rmdf_set_delta(nnpe_ctx* ctx, const float* deltasq) {
  if (!ctx || !deltasq) return NNPE_ERR_ARG;
  RmdfState* R = rmdf_of(ctx);
  if (!R) return ctx_fail(ctx, NNPE_ERR_STATE, "no RMDF");
  NNPE_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
  NNPE_CHECK_HIP(ctx, hipMemcpy(R->deltasq, deltasq, sizeof(float) * R->nsq * R->nq, hipMemcpyHostToDevice));
  return NNPE_OK;
}

int rmdf_enqueue(nnpe_ctx* ctx, bool potential) {
  RmdfState* R = rmdf_of(ctx);
  hipLaunchKernelGGL(k_rmdf_q, dim3(1), dim3(256), 0, ctx->stream, R->nsq, R->nq, ctx->npair, ctx->sq_out, R->sqwt, R->fwt,
                     R->sqex, R->wt, R->gamma, 0.25 * ctx->natom, R->sq, R->deltasq, R->force_q,
                     ctx->ev_sum + 2 * NNPE_TERM_RMDF);
  const int nt = blocks_for(R->nr, 4);
  hipLaunchKernelGGL(k_rmdf_tab, dim3(nt), dim3(256), 0, ctx->stream, ctx->npair, R->nr, R->nq, R->nb1, R->target, R->force_q,
                     R->ftmx, R->fbase, R->ftab);
  if (potential)
    hipLaunchKernelGGL(k_rmdf_tab, dim3(nt), dim3(256), 0, ctx->stream, ctx->npair, R->nr, R->nq, R->nb1, R->target, R->force_q,
                       R->potmx, R->ubase, R->utab);
  NNPE_CHECK_LAUNCH(ctx);
  return NNPE_OK;
}

// compute_Qspace + compute_Rspace (+ compute_Potential) on the current S(q)
extern "C" int nnpe_rmdf_step(nnpe_ctx* ctx, int potential) {
  if (!ctx) return NNPE_ERR_ARG;
  if (!rmdf_of(ctx)) return ctx_fail(ctx, NNPE_ERR_STATE, "no RMDF");
  return rmdf_enqueue(ctx, potential != 0);
}

// S_n(q), deltasq [nsq][nq] and force_q [npair][nq] of the last step (any pointer may be NULL)
extern "C" int nnpe_rmdf_get(nnpe_ctx* ctx, float* sq, float* deltasq, float* force_q) {
  if (!ctx) return NNPE_ERR_ARG;
  RmdfState* R = rmdf_of(ctx);
  if (!R) return ctx_fail(ctx, NNPE_ERR_STATE, "no RMDF");
  NNPE_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
  const size_t n = size_t(R->nsq) * R->nq;
  if (sq) NNPE_CHECK_HIP(ctx, hipMemcpy(sq, R->sq, n * sizeof(float), hipMemcpyDeviceToHost));
  if (deltasq) NNPE_CHECK_HIP(ctx, hipMemcpy(deltasq, R->deltasq, n * sizeof(float), hipMemcpyDeviceToHost));
  if (force_q) NNPE_CHECK_HIP(ctx, hipMemcpy(force_q, R->force_q, sizeof(float) * ctx->npair * R->nq, hipMemcpyDeviceToHost));
  return NNPE_OK;
}

// the refined tables: target 0 [npair][nb1] x2, target 1 [npair][4][nb1] x2
extern "C" int nnpe_rmdf_tables(nnpe_ctx* ctx, float* force, float* potential) {
  if (!ctx) return NNPE_ERR_ARG;
  RmdfState* R = rmdf_of(ctx);
  if (!R) return ctx_fail(ctx, NNPE_ERR_STATE, "no RMDF");
  NNPE_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
  const size_t n = size_t(ctx->npair) * (R->target ? 4 : 1) * R->nb1;
  if (force) NNPE_CHECK_HIP(ctx, hipMemcpy(force, R->ftab, n * sizeof(float), hipMemcpyDeviceToHost));
  if (potential) NNPE_CHECK_HIP(ctx, hipMemcpy(potential, R->utab, n * sizeof(float), hipMemcpyDeviceToHost));
  return NNPE_OK;
}

bool rmdf_active(nnpe_ctx* ctx) { return rmdf_of(ctx) != nullptr; }
float rmdf_gamma(nnpe_ctx* ctx) { return rmdf_of(ctx) ? rmdf_of(ctx)->gamma : 1.0f; }
