// forces.hip -- tabulated LJ + shifted-force Coulomb pair forces and the bonded terms.
//
// Pair (LJTableCoulDSF::compute, src/pair.cpp:289-419): a group of kPairG lanes per atom walks
// its full list row (G consecutive entries per step), and the group's butterfly sum is written
// once per atom: no atomics, deterministic.  The F/r table sits in LDS; each directed entry adds half of the
// pair energy and virial.  The arithmetic per pair is the reference's, in the same order:
//   fpair = lerp(F/r)*s_lj + qi*qj*K*(1/(r*rsq) + fc/r)*s_coul
// Bonded (src/bond.cpp:108-137, src/angle.cpp:87-136, src/dihedral.cpp:109-166): one thread
// per molecule evaluates all of the molecule's terms on unwrapped positions (x + image shift of
// the row's x_mol image), so every atom's bonded force has exactly one writer.
#include <algorithm>
#include <cmath>

#include "internal.h"

namespace {

// special-pair factors (src/molecule.h:11): class 0 full, classes 1 and 2 excluded, class 3 scaled
constexpr float kSpecialLj3 = 0.5f, kSpecialCoul3 = float(5 / 6.);
constexpr float kElectricK = 332.06371f;                                      // src/environ.h:77
constexpr int kPairG = 32;  // lanes per atom
constexpr int kPairU = 4;   // list entries per lane in flight at once

__device__ __forceinline__ int pair_index(int a, int b, int ntype) {
  const int i = min(a, b), j = max(a, b);  // small non-negative values: 24-bit multiplies
  return __mul24(i, ntype) - (__mul24(i, i - 1) >> 1) + (j - i);
}

template <bool EV, int G>
__global__ void __launch_bounds__(256) k_pair(int n, int cap, const uint32_t* __restrict__ nbr,
                                              const int* __restrict__ cnt, const float4* __restrict__ xq,
                                              const int* __restrict__ typ, const BoxPar* __restrict__ boxes,
                                              const int* __restrict__ rbox, int ntype, int npair,
                                              float cutsq, float cut, float rdr, float fc, float vc, int nb1,
                                              const float* __restrict__ tab, float4* __restrict__ f, double* pe,
                                              double* pw) {
#pragma clang fp contract(fast)
  extern __shared__ float s_tab[];  // [npair][nb1] F/r
  __shared__ float4 s_img[27];
  for (int k = threadIdx.x; k < npair * nb1; k += blockDim.x) s_tab[k] = tab[k];
  image_table_fill(s_img);
  __syncthreads();
  const int i = (xcd_block(gridDim.x) * 256 + (int)threadIdx.x) / G;
  const int lane = threadIdx.x & (G - 1);
  double e_acc = 0, w_acc = 0;
  if (i < n) {
    const float4 xi = xq[i];
    const int ti = typ[i];
    const BoxPar& B = boxes[rbox[i]];
    const float bl = B.bl;
    float fx = 0, fy = 0, fz = 0, e = 0, w = 0;
    const int m = cnt[i];
    const uint32_t* row = list_row(nbr, cap, i);
    // kPairU entries per lane and their gathers are in flight together; the entries of the
    // next batch are loaded before this batch is computed.  Each lane still sums its pairs in
    // list order.
    uint32_t en[kPairU];
#pragma unroll
    for (int u = 0; u < kPairU; ++u) {
      const int k = lane + u * G;
      en[u] = k < m ? row[k] : 0u;
    }
    for (int k0 = lane; k0 < m; k0 += kPairU * G) {
      uint32_t ec[kPairU];
      float4 xr[kPairU];
      int tr[kPairU];
#pragma unroll
      for (int u = 0; u < kPairU; ++u) {
        ec[u] = en[u];
        const int k = k0 + (kPairU + u) * G;
        en[u] = k < m ? row[k] : 0u;
      }
#pragma unroll
      for (int u = 0; u < kPairU; ++u) {
        const int j = k0 + u * G < m ? (int)NNPE_ENTRY_J(ec[u]) : i;
        xr[u] = xq[j];
        tr[u] = typ[j];
      }
#pragma unroll
      for (int u = 0; u < kPairU; ++u) {
        if (k0 + u * G >= m) break;
        const uint32_t ent = ec[u];
        const int code = NNPE_ENTRY_CODE(ent), sp = NNPE_ENTRY_SP(ent);
        const float4 xj = image_pos_t(xr[u], code, bl, s_img);
        float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
        const float rsq = dist2_rn(dx, dy, dz);  // the cut as the reference rounds it
        if (!(rsq < cutsq)) continue;
        const float ir = __builtin_amdgcn_rsqf(rsq), r = rsq * ir;
        const int p = pair_index(ti, tr[u], ntype);
        float ipw1 = r * rdr - 0.5f;
        const int t1 = int(ipw1), t2 = t1 + 1;
        const float ipw2 = ipw1 - float(t1);
        ipw1 = 1.0f - ipw2;
        const float* ft = s_tab + __mul24(p, nb1);
        // special-pair factors by select (a per-lane index into constant memory is a load)
        const float slj = sp == 0 ? 1.0f : (sp == 3 ? kSpecialLj3 : 0.0f);
        const float scl = sp == 0 ? 1.0f : (sp == 3 ? kSpecialCoul3 : 0.0f);
        float fpair = (ft[t1] * ipw1 + ft[t2] * ipw2) * slj;
        const float qq = xi.w * xj.w * kElectricK;
        fpair += qq * (ir * ir * ir + fc * ir) * scl;
        fx += (dx *= fpair); fy += (dy *= fpair); fz += (dz *= fpair);
        if (EV) {
          const float* ut = tab + __mul24(npair + p, nb1);
          // energies sum terms of both signs that cancel: correctly rounded r and 1/r here
          const float re = sqrtf(rsq);
          e += 0.5f * ((ut[t1] * ipw1 + ut[t2] * ipw2) * slj + qq * (1.0f / re + vc + fc * (cut - re)) * scl);
          w += 0.5f * fpair * rsq;
        }
      }
    }
    fx = group_sum<G>(fx); fy = group_sum<G>(fy); fz = group_sum<G>(fz);
    if (lane == 0) f[i] = make_float4(fx, fy, fz, 0.0f);
    e_acc = e; w_acc = w;
  }
  if (EV) block_sum2_store<256>(e_acc, w_acc, pe, pw);
}

__device__ __forceinline__ float4 xmol_pos(int aid, const int* inv, const float4* xq, const int* xcode,
                                           const BoxPar* boxes, const int* rbox) {
  const int r = inv[aid];
  const BoxPar& B = boxes[rbox[r]];
  return image_pos(xq[r], xcode[r], B.bl, B.m_bl);
}

template <bool EV>
__global__ void __launch_bounds__(256) k_bonded(int nmol, const int* __restrict__ inv, const float4* __restrict__ xq,
                                                const int* __restrict__ xcode, const BoxPar* __restrict__ boxes,
                                                const int* __restrict__ rbox,
                                                const int* __restrict__ bstart, const int* __restrict__ bterm,
                                                const float* __restrict__ bcoef, const int* __restrict__ astart,
                                                const int* __restrict__ aterm, const float* __restrict__ acoef,
                                                const int* __restrict__ dstart, const int* __restrict__ dterm,
                                                const float* __restrict__ dcoef, float4* __restrict__ f, double* pe_b,
                                                double* pw_b, double* pe_a, double* pw_a, double* pe_d, double* pw_d) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  double eb = 0, wb = 0, ea = 0, ed = 0;
  if (m < nmol) {
    // BOND::compute (src/bond.cpp:108-137)
    for (int t = bstart[m]; t < bstart[m + 1]; ++t) {
      const int* p = bterm + 3 * t;
      const float* icf = bcoef + 4 * p[0];
      const float4 xi = xmol_pos(p[1], inv, xq, xcode, boxes, rbox), xj = xmol_pos(p[2], inv, xq, xcode, boxes, rbox);
      float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
      const float rr = sqrtf(dx * dx + dy * dy + dz * dz);
      const float dr = rr - icf[0], bdr = icf[2] - dr;
      const float ff = dr * (icf[3] + icf[1] * bdr * (bdr - dr)) / rr;
      dx *= ff; dy *= ff; dz *= ff;
      float4* fi = f + inv[p[1]];
      float4* fj = f + inv[p[2]];
      fi->x -= dx; fi->y -= dy; fi->z -= dz;
      fj->x += dx; fj->y += dy; fj->z += dz;
      if (EV) { eb += 0.5f * dr * dr * (icf[1] * bdr * bdr + icf[3]); wb -= ff * rr * rr; }
    }
    // ANGLE::compute (src/angle.cpp:87-136)
    for (int t = astart[m]; t < astart[m + 1]; ++t) {
      const int* p = aterm + 4 * t;
      const float* icf = acoef + 4 * p[0];
      const float4 xi = xmol_pos(p[1], inv, xq, xcode, boxes, rbox), xj = xmol_pos(p[2], inv, xq, xcode, boxes, rbox),
                   xk = xmol_pos(p[3], inv, xq, xcode, boxes, rbox);
      float dx1 = xi.x - xj.x, dy1 = xi.y - xj.y, dz1 = xi.z - xj.z;
      float dx2 = xk.x - xj.x, dy2 = xk.y - xj.y, dz2 = xk.z - xj.z;
      const float r1 = 1.0f / sqrtf(dx1 * dx1 + dy1 * dy1 + dz1 * dz1);
      const float r2 = 1.0f / sqrtf(dx2 * dx2 + dy2 * dy2 + dz2 * dz2);
      dx1 *= r1; dy1 *= r1; dz1 *= r1; dx2 *= r2; dy2 *= r2; dz2 *= r2;
      const float cq = dx1 * dx2 + dy1 * dy2 + dz1 * dz2;
      const float dcq = cq - icf[0], b_dcq = icf[2] - dcq;
      const float fe = dcq * (icf[3] + icf[1] * b_dcq * (b_dcq - dcq));
      float ff = fe * r1;
      float fx = ff * (cq * dx1 - dx2), fy = ff * (cq * dy1 - dy2), fz = ff * (cq * dz1 - dz2);
      float4* fi = f + inv[p[1]];
      float4* fj = f + inv[p[2]];
      float4* fk = f + inv[p[3]];
      fi->x += fx; fi->y += fy; fi->z += fz; fj->x -= fx; fj->y -= fy; fj->z -= fz;
      ff = fe * r2;
      fx = ff * (dx1 - cq * dx2); fy = ff * (dy1 - cq * dy2); fz = ff * (dz1 - cq * dz2);
      fk->x -= fx; fk->y -= fy; fk->z -= fz; fj->x += fx; fj->y += fy; fj->z += fz;
      if (EV) ea += 0.5f * dcq * dcq * (icf[1] * b_dcq * b_dcq + icf[3]);
    }
    // DIHEDRAL::compute (src/dihedral.cpp:109-166)
    for (int t = dstart[m]; t < dstart[m + 1]; ++t) {
      const int* p = dterm + 5 * t;
      const float* icf = dcoef + 4 * p[0];
      const float4 xi = xmol_pos(p[1], inv, xq, xcode, boxes, rbox), xj = xmol_pos(p[2], inv, xq, xcode, boxes, rbox),
                   xk = xmol_pos(p[3], inv, xq, xcode, boxes, rbox), xl = xmol_pos(p[4], inv, xq, xcode, boxes, rbox);
      float dx1 = xi.x - xj.x, dy1 = xi.y - xj.y, dz1 = xi.z - xj.z;
      const float dx2 = xk.x - xj.x, dy2 = xk.y - xj.y, dz2 = xk.z - xj.z;
      float dx3 = xl.x - xk.x, dy3 = xl.y - xk.y, dz3 = xl.z - xk.z;
      const float r_rsq2 = 1.0f / (dx2 * dx2 + dy2 * dy2 + dz2 * dz2);
      const float dot12 = (dx1 * dx2 + dy1 * dy2 + dz1 * dz2) * r_rsq2;
      const float dot32 = (dx3 * dx2 + dy3 * dy2 + dz3 * dz2) * r_rsq2;
      dx1 -= dx2 * dot12; dy1 -= dy2 * dot12; dz1 -= dz2 * dot12;
      dx3 -= dx2 * dot32; dy3 -= dy2 * dot32; dz3 -= dz2 * dot32;
      float r1 = 1.0f / sqrtf(dx1 * dx1 + dy1 * dy1 + dz1 * dz1);
      float r3 = 1.0f / sqrtf(dx3 * dx3 + dy3 * dy3 + dz3 * dz3);
      dx1 *= r1; dy1 *= r1; dz1 *= r1; dx3 *= r3; dy3 *= r3; dz3 *= r3;
      const float cq = dx1 * dx3 + dy1 * dy3 + dz1 * dz3;
      const float ff = (icf[3] * cq + icf[2]) * cq + icf[1];
      r1 *= ff; r3 *= ff;
      const float f1x = r1 * (cq * dx1 - dx3), f1y = r1 * (cq * dy1 - dy3), f1z = r1 * (cq * dz1 - dz3);
      const float f3x = r3 * (cq * dx3 - dx1), f3y = r3 * (cq * dy3 - dy1), f3z = r3 * (cq * dz3 - dz1);
      const float f2x = f1x * dot12 + f3x * dot32, f2y = f1y * dot12 + f3y * dot32, f2z = f1z * dot12 + f3z * dot32;
      float4* fi = f + inv[p[1]];
      float4* fj = f + inv[p[2]];
      float4* fk = f + inv[p[3]];
      float4* fl = f + inv[p[4]];
      fi->x += f1x; fi->y += f1y; fi->z += f1z;
      fl->x += f3x; fl->y += f3y; fl->z += f3z;
      fj->x += f2x - f1x; fj->y += f2y - f1y; fj->z += f2z - f1z;
      fk->x -= f2x + f3x; fk->y -= f2y + f3y; fk->z -= f2z + f3z;
      if (EV) ed += ((icf[3] / 3.0f * cq + icf[2] / 2.0f) * cq + icf[1]) * cq + icf[0];
    }
  }
  if (EV) {
    block_sum2_store<256>(eb, wb, pe_b, pw_b);
    __syncthreads();
    block_sum2_store<256>(ea, 0.0, pe_a, pw_a);
    __syncthreads();
    block_sum2_store<256>(ed, 0.0, pe_d, pw_d);
  }
}

}  // namespace

extern "C" int nnpe_pair_table(nnpe_ctx* ctx, float cutoff, int nbin, const float* force, const float* potential) {
  if (ctx) md_state_free(ctx);  // captured step graphs hold the old parameters
  if (ctx) rmdf_state_free(ctx);  // a refinement is tied to the tables and grids it was set up on
  if (!ctx || !force || !potential || nbin <= 0 || !(cutoff > 0)) return ctx_fail(ctx, NNPE_ERR_ARG, "nnpe_pair_table: bad argument");
  if (!ctx->have_system) return ctx_fail(ctx, NNPE_ERR_STATE, "no system");
  const int nb1 = nbin + 1;
  if (size_t(ctx->npair) * nb1 * sizeof(float) > 60 * 1024)
    return ctx_fail(ctx, NNPE_ERR_UNSUPPORTED, "pair table larger than the 60 KiB LDS budget");
  // LJTableCoulDSF ctor (src/pair.cpp:171-195): dr = rc/nbin, fc = vc/rc, vc = -1/rc
  ctx->pt_cut = cutoff; ctx->pt_cutsq = cutoff * cutoff; ctx->pt_nbin = nbin;
  ctx->pt_rdr = float(nbin) / cutoff;
  ctx->pt_vc = -1.0f / cutoff; ctx->pt_fc = ctx->pt_vc / cutoff;
  std::vector<float> h(size_t(2) * ctx->npair * nb1);
  std::copy(force, force + size_t(ctx->npair) * nb1, h.begin());
  std::copy(potential, potential + size_t(ctx->npair) * nb1, h.begin() + size_t(ctx->npair) * nb1);
  NNPE_CHECK_HIP(ctx, dev_alloc(ctx->pt_tab, h.size()));
  NNPE_CHECK_HIP(ctx, hipMemcpy(ctx->pt_tab, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice));
  ctx->have_pair = true;
  return NNPE_OK;
}

int pair_enqueue(nnpe_ctx* ctx, bool ev) {
  const int N = ctx->natom, nb1 = ctx->pt_nbin + 1, nblk = blocks_for((long long)N * kPairG, 256);
  const size_t lds = size_t(ctx->npair) * nb1 * sizeof(float);
  double* pe = ev_part(ctx, NNPE_TERM_PAIR, 0);
  double* pw = ev_part(ctx, NNPE_TERM_PAIR, 1);
  if (ev)
    hipLaunchKernelGGL((k_pair<true, kPairG>), dim3(nblk), dim3(256), lds, ctx->stream, N, ctx->cap_long, ctx->nbr_long,
                       ctx->cnt_long, ctx->xq, ctx->typ, ctx->d_box, ctx->rbox, ctx->ntype, ctx->npair, ctx->pt_cutsq,
                       ctx->pt_cut, ctx->pt_rdr, ctx->pt_fc, ctx->pt_vc, nb1, ctx->pt_tab, ctx->f, pe, pw);
  else
    hipLaunchKernelGGL((k_pair<false, kPairG>), dim3(nblk), dim3(256), lds, ctx->stream, N, ctx->cap_long, ctx->nbr_long,
                       ctx->cnt_long, ctx->xq, ctx->typ, ctx->d_box, ctx->rbox, ctx->ntype, ctx->npair, ctx->pt_cutsq,
                       ctx->pt_cut, ctx->pt_rdr, ctx->pt_fc, ctx->pt_vc, nb1, ctx->pt_tab, ctx->f, pe, pw);
  NNPE_CHECK_LAUNCH(ctx);
  if (ev) return ev_finish(ctx, NNPE_TERM_PAIR, nblk);
  return NNPE_OK;
}

extern "C" int nnpe_pair_compute(nnpe_ctx* ctx, int ev) {
  if (!ctx) return NNPE_ERR_ARG;
  if (!ctx->have_list || !ctx->have_pair) return ctx_fail(ctx, NNPE_ERR_STATE, "pair compute needs a list and a table");
  if (ctx->pt_cut > ctx->pair_cutoff + ctx->skin) return ctx_fail(ctx, NNPE_ERR_ARG, "table cutoff beyond the list cutoff");
  return pair_enqueue(ctx, ev != 0);
}

extern "C" int nnpe_bonded_set(nnpe_ctx* ctx, int kind, int ntyp, const float* coef, int nterm, const int* terms) {
  if (ctx) md_state_free(ctx);  // captured step graphs hold the old parameters
  if (!ctx || kind < 0 || kind > 2 || ntyp < 0 || nterm < 0 || (nterm && (!coef || !terms)))
    return ctx_fail(ctx, NNPE_ERR_ARG, "nnpe_bonded_set: bad argument");
  if (!ctx->have_system) return ctx_fail(ctx, NNPE_ERR_STATE, "no system");
  const int w = 3 + kind;  // ints per term
  // per-molecule offsets; terms must be grouped by molecule (autoGenerate order,
  // src/bond.cpp:66-86); every atom of a term must belong to the same molecule
  std::vector<int> hmol(ctx->natom);
  {
    std::vector<int4> hinfo(ctx->natom);
    NNPE_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    NNPE_CHECK_HIP(ctx, hipMemcpy(hinfo.data(), ctx->info, sizeof(int4) * ctx->natom, hipMemcpyDeviceToHost));
    for (auto& in : hinfo) hmol[in.x] = in.y;
  }
  std::vector<int> start(ctx->nmol + 1, 0);
  int prev = -1;
  for (int t = 0; t < nterm; ++t) {
    const int* p = terms + w * t;
    if (p[0] < 0 || p[0] >= ntyp) return ctx_fail(ctx, NNPE_ERR_ARG, "auto generate error: term type out of range");
    for (int k = 1; k < w; ++k)
      if (p[k] < 0 || p[k] >= ctx->natom || hmol[p[k]] != hmol[p[1]]) return ctx_fail(ctx, NNPE_ERR_ARG, "term atoms outside one molecule");
    int m = hmol[p[1]];
    if (m < prev) return ctx_fail(ctx, NNPE_ERR_ARG, "terms not grouped by molecule");
    prev = m;
    start[m + 1]++;
  }
  for (int m = 0; m < ctx->nmol; ++m) start[m + 1] += start[m];
  int** pterm = kind == 0 ? &ctx->bterm : kind == 1 ? &ctx->aterm : &ctx->dterm;
  int** pstart = kind == 0 ? &ctx->bstart : kind == 1 ? &ctx->astart : &ctx->dstart;
  float** pcoef = kind == 0 ? &ctx->bcoef : kind == 1 ? &ctx->acoef : &ctx->dcoef;
  NNPE_CHECK_HIP(ctx, dev_alloc(*pterm, size_t(w) * std::max(nterm, 1)));
  if (nterm) NNPE_CHECK_HIP(ctx, hipMemcpy(*pterm, terms, sizeof(int) * w * nterm, hipMemcpyHostToDevice));
  NNPE_CHECK_HIP(ctx, dev_alloc(*pstart, start.size()));
  NNPE_CHECK_HIP(ctx, hipMemcpy(*pstart, start.data(), sizeof(int) * start.size(), hipMemcpyHostToDevice));
  ctx->start_len[kind] = int(start.size());
  NNPE_CHECK_HIP(ctx, dev_alloc(*pcoef, size_t(4) * std::max(ntyp, 1)));
  if (ntyp) NNPE_CHECK_HIP(ctx, hipMemcpy(*pcoef, coef, sizeof(float) * 4 * ntyp, hipMemcpyHostToDevice));
  if (kind == 0) { ctx->nbtyp = ntyp; ctx->nbond = nterm; }
  else if (kind == 1) { ctx->natyp = ntyp; ctx->nangle = nterm; }
  else { ctx->ndtyp = ntyp; ctx->ndih = nterm; }
  return NNPE_OK;
}

int bonded_enqueue(nnpe_ctx* ctx, bool ev) {
  // the kernel reads offsets [0, nmol]: kinds without terms get an all-zero table of that size
  // (again when the molecule count changed); a kind whose terms were set for another molecule
  // table is an error
  for (int kind = 0; kind < 3; ++kind) {
    int** pstart = kind == 0 ? &ctx->bstart : kind == 1 ? &ctx->astart : &ctx->dstart;
    const int nterm = kind == 0 ? ctx->nbond : kind == 1 ? ctx->nangle : ctx->ndih;
    if (*pstart && ctx->start_len[kind] == ctx->nmol + 1) continue;
    if (nterm > 0) return ctx_fail(ctx, NNPE_ERR_STATE, "bonded terms were set for another molecule count");
    std::vector<int> z(ctx->nmol + 1, 0);
    NNPE_CHECK_HIP(ctx, dev_alloc(*pstart, z.size()));
    NNPE_CHECK_HIP(ctx, hipMemcpy(*pstart, z.data(), sizeof(int) * z.size(), hipMemcpyHostToDevice));
    ctx->start_len[kind] = int(z.size());
  }
  if (ctx->nbond + ctx->nangle + ctx->ndih == 0) return NNPE_OK;
  const int nblk = blocks_for(ctx->nmol, 256);
  NNPE_CHECK_HIP(ctx, ev_reserve(ctx, nblk));
  auto P = [&](int term, int w) { return ev_part(ctx, term, w); };
  if (ev)
    hipLaunchKernelGGL(k_bonded<true>, dim3(nblk), dim3(256), 0, ctx->stream, ctx->nmol, ctx->inv, ctx->xq, ctx->xmol_code,
                       ctx->d_box, ctx->rbox, ctx->bstart, ctx->bterm, ctx->bcoef, ctx->astart, ctx->aterm, ctx->acoef,
                       ctx->dstart, ctx->dterm, ctx->dcoef, ctx->f, P(NNPE_TERM_BOND, 0), P(NNPE_TERM_BOND, 1),
                       P(NNPE_TERM_ANGLE, 0), P(NNPE_TERM_ANGLE, 1), P(NNPE_TERM_DIHEDRAL, 0), P(NNPE_TERM_DIHEDRAL, 1));
  else
    hipLaunchKernelGGL(k_bonded<false>, dim3(nblk), dim3(256), 0, ctx->stream, ctx->nmol, ctx->inv, ctx->xq, ctx->xmol_code,
                       ctx->d_box, ctx->rbox, ctx->bstart, ctx->bterm, ctx->bcoef, ctx->astart, ctx->aterm, ctx->acoef,
                       ctx->dstart, ctx->dterm, ctx->dcoef, ctx->f, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
  NNPE_CHECK_LAUNCH(ctx);
  if (ev) {
    int rc = ev_finish(ctx, NNPE_TERM_BOND, nblk);
    if (!rc) rc = ev_finish(ctx, NNPE_TERM_ANGLE, nblk);
    if (!rc) rc = ev_finish(ctx, NNPE_TERM_DIHEDRAL, nblk);
    return rc;
  }
  return NNPE_OK;
}

extern "C" int nnpe_bonded_compute(nnpe_ctx* ctx, int ev) {
  if (!ctx) return NNPE_ERR_ARG;
  if (!ctx->have_list) return ctx_fail(ctx, NNPE_ERR_STATE, "bonded compute needs a list build (unwrapped images)");
  return bonded_enqueue(ctx, ev != 0);
}
