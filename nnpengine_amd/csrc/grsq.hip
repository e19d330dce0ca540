// grsq.hip -- partial g(r) histograms and the S(q) transform.
//
// GR (src/grsq.cpp:28-219): bin = int(sqrt(rsq) * 1/dr) in FP32 exactly as the reference, so
// the integer counts are bit-identical.  Each workgroup keeps a private LDS histogram
// (npair x nbin ints, LDS atomics), flushed once per workgroup with global 64-bit atomics
// (integers: order-independent, deterministic).
//   inner tally  (GR::tally<true>, :94-149): every directed entry of the full pair list within
//                r_max_inner adds 1 (2 per unordered pair, as the reference's +2/+1 rule);
//                bins >= nbin_inner keep their previous counts, as the reference's do.
//   full tally   (GR::tally(bin_list), :151-205): cell stencil to r_max; real j > i add 2,
//                ghost images add 1; one workgroup per home cell, candidates staged in LDS.
// SQ (src/grsq.cpp:258-322): S_ab(q) = sum_r 4 pi rho dr r sin(q r)/q h_ab(r) in FP64, one
// workgroup per q point with a fixed-order reduction.
#include <algorithm>
#include <cmath>

#include "internal.h"

namespace {

constexpr int kHistLdsMax = 48 * 1024;
constexpr int kHistLdsFused = 16 * 1024;  // histogram added to the rebuild kernel's LDS
constexpr int kGrG = 16;  // lanes per atom, inner tally

__device__ __forceinline__ int npair_of(int ntype) { return ntype * (ntype + 1) / 2; }

__device__ __forceinline__ int pair_index(int a, int b, int ntype) {
  const int i = min(a, b), j = max(a, b);  // small non-negative values: 24-bit multiplies
  return __mul24(i, ntype) - (__mul24(i, i - 1) >> 1) + (j - i);
}

__device__ __forceinline__ int gr_bin(float rsq, float rdr) { return gr_bin_of(rsq, rdr); }

__global__ void k_gr_zero(unsigned long long* count, int npair, int nbin, int nzero) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= npair * nzero) return;
  count[(k / nzero) * nbin + (k % nzero)] = 0ull;
}

// inner tally: groups of G lanes per atom walk the list rows; each workgroup owns a contiguous
// range of atoms (so its LDS histogram is flushed once for many atoms)
template <bool kLds, int G>
__global__ void __launch_bounds__(256) k_gr_inner(int n, int cap, const uint32_t* __restrict__ nbr,
                                                  const int* __restrict__ cnt, const float4* __restrict__ xq,
                                                  const int* __restrict__ typ, int ntype, int npair,
                                                  const BoxPar* __restrict__ boxes, const int* __restrict__ rbox,
                                                  float rmaxsq, float rdr, int nbin_inner, int nbin,
                                                  unsigned long long* __restrict__ count) {
  extern __shared__ unsigned int s_hist[];
  if (kLds) {
    for (int k = threadIdx.x; k < npair * nbin_inner; k += blockDim.x) s_hist[k] = 0u;
    __syncthreads();
  }
  constexpr int kGroups = 256 / G;
  const int per = (n + gridDim.x - 1) / gridDim.x;
  const int a0 = xcd_block(gridDim.x) * per, a1 = min(n, a0 + per);
  const int lane = threadIdx.x & (G - 1);
  for (int i = a0 + (int)threadIdx.x / G; i < a1; i += kGroups) {
    const float4 xi = xq[i];
    const int ti = typ[i];
    const float bl = boxes[rbox[i]].bl, m_bl = boxes[rbox[i]].m_bl;
    const int m = cnt[i];
    const uint32_t* row = list_row(nbr, cap, i);
    for (int k = lane; k < m; k += G) {
      const uint32_t e = row[k];
      const int j = NNPE_ENTRY_J(e);
      const float4 xj = image_pos(xq[j], NNPE_ENTRY_CODE(e), bl, m_bl);
      const float rsq = dist2_rn(__fsub_rn(xi.x, xj.x), __fsub_rn(xi.y, xj.y), __fsub_rn(xi.z, xj.z));
      if (!(rsq < rmaxsq)) continue;
      const int b = gr_bin(rsq, rdr);
      if (b >= nbin_inner) continue;
      const int p = pair_index(ti, ntype <= NNPE_ENTRY_MAX_TYPES ? NNPE_ENTRY_T(e) : typ[j], ntype);
      if (kLds) atomicAdd(&s_hist[p * nbin_inner + b], 1u);
      else atomicAdd(&count[(size_t)p * nbin + b], 1ull);
    }
  }
  if (kLds) {
    __syncthreads();
    for (int k = threadIdx.x; k < npair * nbin_inner; k += blockDim.x) {
      unsigned int v = s_hist[k];
      if (v) atomicAdd(&count[(size_t)(k / nbin_inner) * nbin + (k % nbin_inner)], (unsigned long long)v);
    }
  }
}

// Full tally by home cell: persistent workgroups take the cells round-robin; for each home cell,
// the candidates of its stencil cells (real rows of the cells at or after it, then ghost
// images, per stencil cell) are
// staged in LDS in batches and each wave tests 64 of them per step against one home atom,
// adding 2 for a real j with a larger row (each unordered real pair once) and 1 for a ghost
// image, into the workgroup's LDS histogram (flushed once per workgroup).
constexpr int kGrBatch = 1024;
constexpr int kGrMaxRanges = 2 * 9 * 9 * 9;

template <bool kLds>
__global__ void __launch_bounds__(256) k_gr_cells(int ncell, const BoxPar* __restrict__ boxes,
                                                  const int* __restrict__ cell_box, const int4* __restrict__ stencil,
                                                  int nst, const float4* __restrict__ xq, const int* __restrict__ typ,
                                                  const int* __restrict__ cstart, const int* __restrict__ gstart,
                                                  const int* __restrict__ gs_src, const int* __restrict__ gs_code,
                                                  int ntype, int npair, float rmaxsq, float rdr,
                                                  int nbin, unsigned long long* __restrict__ count) {
  extern __shared__ unsigned int s_hist[];
  __shared__ int r_begin[kGrMaxRanges], r_pref[kGrMaxRanges + 1];
  __shared__ unsigned char r_ghost[kGrMaxRanges];
  __shared__ float4 c_pos[kGrBatch];
  __shared__ int c_row[kGrBatch];           // real row, or -1 for a ghost image
  __shared__ unsigned char c_typ[kGrBatch];
  __shared__ int s_wsum[256 / 64 + 1];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (kLds) {
    for (int k = tid; k < npair * nbin; k += blockDim.x) s_hist[k] = 0u;
  }
  const int nr2 = 2 * nst;
  // cells dealt round-robin: the cells holding atoms form a block in the middle of the grid
  // (the margin holds only ghosts), so contiguous runs would load some workgroups with many
  // occupied cells and leave others idle
  for (int c = xcd_block(gridDim.x); c < ncell; c += gridDim.x) {
    const int hb = cstart[c], he = cstart[c + 1];
    if (hb == he) continue;  // uniform across the workgroup
    const BoxPar B = boxes[cell_box[c]];
    const float bl = B.bl, m_bl = B.m_bl;
    const int num_gx = B.num_gx, lc = c - B.cell_base;
    const int cz = lc % num_gx, cy = (lc / num_gx) % num_gx, cx = lc / (num_gx * num_gx);
    __syncthreads();  // previous cell's ranges and batch are consumed
    for (int s = tid; s < nst; s += blockDim.x) {
      const int4 st = stencil[s];
      const int ax = cx + st.x, ay = cy + st.y, az = cz + st.z;
      int nr = 0, ng = 0, br = 0, bg = 0;
      if (ax >= 0 && ay >= 0 && az >= 0 && ax < num_gx && ay < num_gx && az < num_gx) {
        const int nc = B.cell_base + (ax * num_gx + ay) * num_gx + az;
        // real rows are sorted by cell, so a real partner with a larger row lies in this cell
        // or a later one: the earlier cells' real rows are not staged at all
        if (nc >= c) { br = cstart[nc]; nr = cstart[nc + 1] - br; }
        bg = gstart[nc]; ng = gstart[nc + 1] - bg;
      }
      r_begin[2 * s] = br; r_pref[2 * s] = nr; r_ghost[2 * s] = 0;
      r_begin[2 * s + 1] = bg; r_pref[2 * s + 1] = ng; r_ghost[2 * s + 1] = 1;
    }
    __syncthreads();
    const int total = block_scan_excl<256>(r_pref, nr2, s_wsum);
    for (int t0 = 0; t0 < total; t0 += kGrBatch) {
      const int tn = min(kGrBatch, total - t0);
      __syncthreads();
      for (int u = tid; u < tn; u += blockDim.x) {
        const int t = t0 + u;
        int lo = 0, hi = nr2 - 1;
        while (lo < hi) {
          const int mid = (lo + hi + 1) >> 1;
          if (r_pref[mid] <= t) lo = mid; else hi = mid - 1;
        }
        const int pos = r_begin[lo] + (t - r_pref[lo]);
        if (r_ghost[lo]) {
          const int src = gs_src[pos];
          c_pos[u] = image_pos(xq[src], gs_code[pos], bl, m_bl);
          c_row[u] = -1;
          c_typ[u] = (unsigned char)typ[src];
        } else {
          c_pos[u] = xq[pos];
          c_row[u] = pos;
          c_typ[u] = (unsigned char)typ[pos];
        }
      }
      __syncthreads();
      for (int h = hb + wave; h < he; h += blockDim.x / 64) {
        const float4 xi = xq[h];
        const int ti = typ[h];
        for (int u = lane; u < tn; u += 64) {
          const int jr = c_row[u];
          if (jr >= 0 && jr <= h) continue;
          const float4 xj = c_pos[u];
          const float rsq = dist2_rn(__fsub_rn(xi.x, xj.x), __fsub_rn(xi.y, xj.y), __fsub_rn(xi.z, xj.z));
          if (!(rsq < rmaxsq)) continue;
          const int b = gr_bin(rsq, rdr);
          if (b >= nbin) continue;
          const int p = pair_index(ti, c_typ[u], ntype);
          const unsigned int wgt = jr >= 0 ? 2u : 1u;
          if (kLds) atomicAdd(&s_hist[p * nbin + b], wgt);
          else atomicAdd(&count[(size_t)p * nbin + b], (unsigned long long)wgt);
        }
      }
    }
  }
  if (kLds) {
    __syncthreads();
    for (int k = tid; k < npair * nbin; k += blockDim.x) {
      const unsigned int v = s_hist[k];
      if (v) atomicAdd(&count[k], (unsigned long long)v);
    }
  }
}

// Shell pass of the full tally folded into a list rebuild: the rebuild counted every pair of its
// own candidates (the pair stencil); the pairs within r_max whose cells lie outside it are in
// the outer stencil cells (on the device only those a point of the home cell can reach).
// Lane per (atom, outer cell): cells whose box lies farther than r_max (plus a margin) from the
// atom, or that hold no candidate, are dropped; the wave then takes its remaining items eight at
// a time, 8 lanes per item walking the item cell's real rows above the atom (weight 2) and its
// ghost images (weight 1) -- 2 per unordered pair, the reference's rule.  Counts go to
// kGrShellCopies private copies (the few bins near r_max take every add), the same copies the
// rebuild kernel added its workgroup histograms to, folded into the tally by k_gr_fold_shell.
// Positions and cells are those of the rebuild that just ran, so the cell boxes bound the atoms.
__global__ void __launch_bounds__(256) k_gr_shell(int n, int nout, const int4* __restrict__ outer,
                                                  const BoxPar* __restrict__ boxes, const int* __restrict__ rbox,
                                                  const int* __restrict__ cell, const float4* __restrict__ xq,
                                                  const int* __restrict__ typ, const int* __restrict__ cstart,
                                                  const int* __restrict__ gstart, const int* __restrict__ gs_src,
                                                  const int* __restrict__ gs_code, int ntype, float rmaxsq,
                                                  float rcull_sq, float rdr, int nbin,
                                                  unsigned long long* __restrict__ copies) {
  unsigned long long* count = copies + (size_t)(blockIdx.x % kGrShellCopies) * (npair_of(ntype) * nbin);
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  bool keep = false;
  int i = 0, nc = 0;
  if (t < (long long)n * nout) {
    i = int(t / nout);
    const int s = int(t - (long long)i * nout);
    const BoxPar B = boxes[rbox[i]];
    const int num_gx = B.num_gx, lc = cell[i] - B.cell_base;
    const int cz = lc % num_gx, cy = (lc / num_gx) % num_gx, cx = lc / (num_gx * num_gx);
    const int4 st = outer[s];
    const int ax = cx + st.x, ay = cy + st.y, az = cz + st.z;
    if (ax >= 0 && ay >= 0 && az >= 0 && ax < num_gx && ay < num_gx && az < num_gx) {
      const float4 xi = xq[i];
      const float gdx = 1.0f / B.g_rdx;
      // distance from the atom to the cell box [k gdx - off, (k+1) gdx - off) along one axis
      auto gap = [&](int d, int k, float x) -> float {
        if (d > 0) return fmaxf(0.0f, float(k) * gdx - B.g_offset - x);
        if (d < 0) return fmaxf(0.0f, x - (float(k + 1) * gdx - B.g_offset));
        return 0.0f;
      };
      const float gx = gap(st.x, ax, xi.x), gy = gap(st.y, ay, xi.y), gz = gap(st.z, az, xi.z);
      nc = B.cell_base + (ax * num_gx + ay) * num_gx + az;
      keep = gx * gx + gy * gy + gz * gz <= rcull_sq &&
             (cstart[nc + 1] > max(cstart[nc], i + 1) || gstart[nc + 1] > gstart[nc]);
    }
  }
  // the wave takes its surviving items eight at a time, one per 8-lane part: an item is one
  // cell's rows above the atom and its ghost images, about a dozen candidates, so a whole wave
  // per item would leave most lanes idle (16 lanes per item measured 2.5 us slower per rebuild)
  unsigned long long bal = __ballot(keep);
  const int part = lane >> 3, ql = lane & 7;
  while (bal) {  // wave-uniform: every lane holds the same mask
    int src = -1;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int s = bal ? __ffsll((long long)bal) - 1 : -1;
      bal &= bal - 1ull;  // (0 stays 0)
      if (q == part) src = s;
    }
    const int ii = __shfl(i, max(src, 0), 64), cc = __shfl(nc, max(src, 0), 64);
    if (src < 0) continue;
    const float4 xi = xq[ii];
    const int ti = typ[ii];
    const BoxPar& B = boxes[rbox[ii]];
    const int j0 = max(cstart[cc], ii + 1), nr = max(0, cstart[cc + 1] - j0);
    const int g0 = gstart[cc], ng = gstart[cc + 1] - g0;
    for (int k = ql; k < nr + ng; k += 8) {
      float4 xj;
      int tj;
      unsigned long long wgt;
      if (k < nr) {
        xj = xq[j0 + k]; tj = typ[j0 + k]; wgt = 2ull;
      } else {
        const int g = g0 + (k - nr), sr = gs_src[g];
        xj = image_pos(xq[sr], gs_code[g], B.bl, B.m_bl); tj = typ[sr]; wgt = 1ull;
      }
      const float rsq = dist2_rn(__fsub_rn(xi.x, xj.x), __fsub_rn(xi.y, xj.y), __fsub_rn(xi.z, xj.z));
      if (!(rsq < rmaxsq)) continue;
      const int b = gr_bin(rsq, rdr);
      if (b < nbin) atomicAdd(&count[(size_t)pair_index(ti, tj, ntype) * nbin + b], wgt);
    }
  }
}

// count += the copies (integer sums: order-free), then h(r) as k_gr_reduce
__global__ void k_gr_fold_shell(int total, unsigned long long* count, const unsigned long long* copies, const double* norm,
                                double* h) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= total) return;
  unsigned long long c = count[k];
  for (int r = 0; r < kGrShellCopies; ++r) c += copies[(size_t)r * total + k];
  count[k] = c;
  h[k] = double(c) * norm[k] - 1.0;
}

__global__ void k_gr_reduce(int total, const unsigned long long* count, const double* norm, double* h) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < total) h[k] = double(count[k]) * norm[k] - 1.0;
}

// one workgroup per q: S_p(q) = sum_r ftmx[q][r] h_p[r]
__global__ void __launch_bounds__(256) k_sq(int nq, int nbin, int npair, const double* ftmx, const double* h, double* out) {
  const int q = blockIdx.x;
  __shared__ double red[256];
  for (int p = 0; p < npair; ++p) {
    double acc = 0;
    for (int r = threadIdx.x; r < nbin; r += 256) acc += ftmx[(size_t)q * nbin + r] * h[(size_t)p * nbin + r];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
      if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[(size_t)p * nq + q] = red[0];
    __syncthreads();
  }
}

}  // namespace

extern "C" int nnpe_gr_init(nnpe_ctx* ctx, int nbi, float r_max, float r_max_inner, int* nbin_total) {
  if (ctx) md_state_free(ctx);  // captured step graphs hold the old parameters
  if (ctx) rmdf_state_free(ctx);  // a refinement is tied to the tables and grids it was set up on
  if (!ctx || nbi <= 0) return NNPE_ERR_ARG;
  if (!ctx->have_grid) return ctx_fail(ctx, NNPE_ERR_STATE, "nnpe_gr_init before nnpe_neigh_init");
  if (ctx->nbox != 1) return ctx_fail(ctx, NNPE_ERR_UNSUPPORTED, "g(r) is defined for a single box");
  // GR ctor (src/grsq.cpp:28-47)
  if (r_max_inner <= 0) {
    r_max = r_max <= 0 ? ctx->h_bl : r_max;
    ctx->gr_nbin_inner = ctx->gr_nbin = nbi;
    ctx->gr_rmaxsq_inner = ctx->gr_rmaxsq = r_max * r_max;
    ctx->gr_dr = r_max / nbi; ctx->gr_rdr = nbi / r_max;
  } else {
    ctx->gr_nbin_inner = nbi;
    ctx->gr_dr = r_max_inner / nbi; ctx->gr_rdr = nbi / r_max_inner;
    ctx->gr_rmaxsq_inner = r_max_inner * r_max_inner;
    r_max = r_max <= 0 ? ctx->h_bl : r_max;
    if (r_max < r_max_inner) return ctx_fail(ctx, NNPE_ERR_ARG, "r_max_inner larger than r_max");
    ctx->gr_nbin = int(r_max * ctx->gr_rdr);
    const float t = ctx->gr_nbin * ctx->gr_dr;
    ctx->gr_rmaxsq = t * t;
  }
  if (r_max > ctx->gr_cutoff + ctx->skin + 1e-4f && r_max > ctx->pair_cutoff + ctx->skin + 1e-4f)
    return ctx_fail(ctx, NNPE_ERR_ARG, "g(r) range beyond the ghost shell");
  const int npair = ctx->npair, nb = ctx->gr_nbin;
  NNPE_CHECK_HIP(ctx, dev_alloc(ctx->gr_count, size_t(npair) * nb));
  NNPE_CHECK_HIP(ctx, hipMemset(ctx->gr_count, 0, sizeof(unsigned long long) * npair * nb));
  NNPE_CHECK_HIP(ctx, dev_alloc(ctx->gr_shell, size_t(kGrShellCopies) * npair * nb));
  NNPE_CHECK_HIP(ctx, dev_alloc(ctx->gr_h, size_t(npair) * nb));
  // normalisation (src/grsq.cpp:56-70) in FP64
  std::vector<double> norm(size_t(npair) * nb);
  const double pi4 = 4 * 3.141592653589793;
  const double rc_dr3 = std::pow(double(ctx->bl) * double(ctx->gr_rdr), 3) / pi4;
  int ip = 0;
  for (int i = 0; i < ctx->ntype; ++i)
    for (int j = i; j < ctx->ntype; ++j, ++ip) {
      const double ni = ctx->h_typecount[i], nj = ctx->h_typecount[j];
      const double normg = rc_dr3 / (ni * (i == j ? (nj - 1) : 2 * nj));
      for (int k = 0; k < nb; ++k) norm[size_t(ip) * nb + k] = normg / (double(k) * (k + 1) + 1.0 / 3);
    }
  NNPE_CHECK_HIP(ctx, dev_alloc(ctx->gr_norm, norm.size()));
  NNPE_CHECK_HIP(ctx, hipMemcpy(ctx->gr_norm, norm.data(), norm.size() * sizeof(double), hipMemcpyHostToDevice));
  if (nbin_total) *nbin_total = nb;
  return gr_outer_refresh(ctx);
}

int gr_enqueue(nnpe_ctx* ctx, bool full) {
  const int N = ctx->natom, npair = ctx->npair, nb = ctx->gr_nbin;
  hipStream_t s = ctx->stream;
  if (!full) {
    const int nz = ctx->gr_nbin_inner;
    hipLaunchKernelGGL(k_gr_zero, dim3(blocks_for(npair * nz, 256)), dim3(256), 0, s, ctx->gr_count, npair, nb, nz);
    const size_t lds = size_t(npair) * nz * sizeof(unsigned int);
    // about 64 atoms per workgroup, at most 2048 workgroups
    const int nblk = std::max(1, std::min(2048, blocks_for(N, 64)));
    if (lds <= kHistLdsMax)
      hipLaunchKernelGGL((k_gr_inner<true, kGrG>), dim3(nblk), dim3(256), lds, s, N, ctx->cap_long, ctx->nbr_long,
                         ctx->cnt_long, ctx->xq, ctx->typ, ctx->ntype, npair, ctx->d_box, ctx->rbox, ctx->gr_rmaxsq_inner,
                         ctx->gr_rdr, nz, nb, ctx->gr_count);
    else
      hipLaunchKernelGGL((k_gr_inner<false, kGrG>), dim3(nblk), dim3(256), 0, s, N, ctx->cap_long, ctx->nbr_long,
                         ctx->cnt_long, ctx->xq, ctx->typ, ctx->ntype, npair, ctx->d_box, ctx->rbox, ctx->gr_rmaxsq_inner,
                         ctx->gr_rdr, nz, nb, ctx->gr_count);
  } else {
    hipLaunchKernelGGL(k_gr_zero, dim3(blocks_for(npair * nb, 256)), dim3(256), 0, s, ctx->gr_count, npair, nb, nb);
    if (ctx->n_stencil_gr > kGrMaxRanges / 2) return ctx_fail(ctx, NNPE_ERR_UNSUPPORTED, "g(r) stencil too large");
    const size_t lds = size_t(npair) * nb * sizeof(unsigned int);
    const int G = ctx->num_grid;
    const int nblk = std::max(1, std::min(1024, G));  // persistent workgroups, cells round-robin
    if (lds <= kHistLdsMax)
      hipLaunchKernelGGL(k_gr_cells<true>, dim3(nblk), dim3(256), lds, s, G, ctx->d_box, ctx->cell_box, ctx->stencil_gr,
                         ctx->n_stencil_gr, ctx->xq, ctx->typ, ctx->cell_start, ctx->gcell_start, ctx->gs_src,
                         ctx->gs_code, ctx->ntype, npair, ctx->gr_rmaxsq, ctx->gr_rdr, nb, ctx->gr_count);
    else
      hipLaunchKernelGGL(k_gr_cells<false>, dim3(nblk), dim3(256), 0, s, G, ctx->d_box, ctx->cell_box, ctx->stencil_gr,
                         ctx->n_stencil_gr, ctx->xq, ctx->typ, ctx->cell_start, ctx->gcell_start, ctx->gs_src,
                         ctx->gs_code, ctx->ntype, npair, ctx->gr_rmaxsq, ctx->gr_rdr, nb, ctx->gr_count);
  }
  NNPE_CHECK_LAUNCH(ctx);
  hipLaunchKernelGGL(k_gr_reduce, dim3(blocks_for(npair * nb, 256)), dim3(256), 0, s, npair * nb, ctx->gr_count, ctx->gr_norm,
                     ctx->gr_h);
  NNPE_CHECK_LAUNCH(ctx);
  return NNPE_OK;
}

// the full tally can ride on a rebuild when its histogram fits the rebuild kernel's LDS
bool gr_fusable(nnpe_ctx* ctx) {
  return ctx->gr_count && ctx->nbox == 1 && size_t(ctx->npair) * ctx->gr_nbin * sizeof(unsigned int) <= kHistLdsFused;
}

int gr_shell_size(nnpe_ctx* ctx) { return kGrShellCopies * ctx->npair * ctx->gr_nbin; }

// The outer stencil cells that some point of the home cell can reach within the tally's range
// (r_max of nnpe_gr_init, else the g(r) cutoff + skin, plus a margin).  A cell at offset d along
// an axis is at least (|d| - 1) cell edges away.
int gr_outer_refresh(nnpe_ctx* ctx) {
  if (!ctx->have_grid) return NNPE_OK;
  const double edge = 1.0 / double(ctx->g_rdx);
  const double reach = (ctx->gr_count ? std::sqrt(double(ctx->gr_rmaxsq)) + 0.01 : double(ctx->gr_cutoff + ctx->skin)) + 0.05;
  std::vector<int4> keep;
  for (const int4& c : ctx->h_stencil_gr_outer) {
    auto g = [&](int d) { return d == 0 ? 0.0 : (std::abs(d) - 1) * edge; };
    const double gx = g(c.x), gy = g(c.y), gz = g(c.z);
    if (gx * gx + gy * gy + gz * gz < reach * reach) keep.push_back(c);
  }
  ctx->n_stencil_gr_outer = int(keep.size());
  NNPE_CHECK_HIP(ctx, dev_alloc(ctx->stencil_gr_outer, std::max<size_t>(1, keep.size())));
  if (!keep.empty())
    NNPE_CHECK_HIP(ctx, hipMemcpy(ctx->stencil_gr_outer, keep.data(), keep.size() * sizeof(int4), hipMemcpyHostToDevice));
  return NNPE_OK;
}

int gr_shell_finish_enqueue(nnpe_ctx* ctx) {
  const int npair = ctx->npair, nb = ctx->gr_nbin;
  if (ctx->n_stencil_gr_outer > 0) {
    const float rcull = std::sqrt(ctx->gr_rmaxsq) + 0.01f;
    const long long nt = (long long)ctx->natom * ctx->n_stencil_gr_outer;
    hipLaunchKernelGGL(k_gr_shell, dim3(blocks_for(nt, 256)), dim3(256), 0, ctx->stream, ctx->natom, ctx->n_stencil_gr_outer,
                       ctx->stencil_gr_outer, ctx->d_box, ctx->rbox, ctx->cell, ctx->xq, ctx->typ, ctx->cell_start,
                       ctx->gcell_start, ctx->gs_src, ctx->gs_code, ctx->ntype, ctx->gr_rmaxsq, rcull * rcull, ctx->gr_rdr, nb,
                       ctx->gr_shell);
    NNPE_CHECK_LAUNCH(ctx);
  }
  hipLaunchKernelGGL(k_gr_fold_shell, dim3(blocks_for(npair * nb, 256)), dim3(256), 0, ctx->stream, npair * nb, ctx->gr_count,
                     ctx->gr_shell, ctx->gr_norm, ctx->gr_h);
  NNPE_CHECK_LAUNCH(ctx);
  return NNPE_OK;
}

extern "C" int nnpe_gr_tally(nnpe_ctx* ctx, int full) {
  if (!ctx) return NNPE_ERR_ARG;
  if (!ctx->have_list || !ctx->gr_count) return ctx_fail(ctx, NNPE_ERR_STATE, "g(r) tally needs nnpe_gr_init and a list");
  if (!full && ctx->gr_rmaxsq_inner > ctx->cutsq_list) return ctx_fail(ctx, NNPE_ERR_ARG, "inner g(r) range beyond the list");
  return gr_enqueue(ctx, full != 0);
}

extern "C" int nnpe_gr_get(nnpe_ctx* ctx, long long* counts, double* h) {
  if (!ctx || !ctx->gr_count) return ctx_fail(ctx, NNPE_ERR_STATE, "no g(r)");
  const size_t n = size_t(ctx->npair) * ctx->gr_nbin;
  NNPE_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (counts) NNPE_CHECK_HIP(ctx, hipMemcpy(counts, ctx->gr_count, n * sizeof(long long), hipMemcpyDeviceToHost));
  if (h) NNPE_CHECK_HIP(ctx, hipMemcpy(h, ctx->gr_h, n * sizeof(double), hipMemcpyDeviceToHost));
  return NNPE_OK;
}

extern "C" int nnpe_sq_init(nnpe_ctx* ctx, int nq, double q_min, double q_max) {
  if (ctx) md_state_free(ctx);  // captured step graphs hold the old parameters
  if (ctx) rmdf_state_free(ctx);  // a refinement is tied to the tables and grids it was set up on
  if (!ctx || nq <= 0 || !(q_max > q_min) || q_min < 0) return ctx_fail(ctx, NNPE_ERR_ARG, "nnpe_sq_init: bad argument");
  if (!ctx->gr_count) return ctx_fail(ctx, NNPE_ERR_STATE, "nnpe_sq_init before nnpe_gr_init");
  ctx->sq_nq = nq; ctx->sq_qmin = q_min; ctx->sq_qmax = q_max;
  const int nb = ctx->gr_nbin;
  std::vector<double> ft(size_t(nq) * nb);
  const double dq = (q_max - q_min) / nq, dr = double(ctx->gr_dr);
  const double rho4pidr = 4 * 3.141592653589793 * ctx->natom / std::pow(double(ctx->bl), 3) * dr;
  for (int i = 0; i < nq; ++i) {
    const double q = q_min + (i + 0.5) * dq;
    for (int k = 0; k < nb; ++k) {
      const double r = (k + 0.5) * dr;
      ft[size_t(i) * nb + k] = std::sin(q * r) * rho4pidr * r / q;
    }
  }
  NNPE_CHECK_HIP(ctx, dev_alloc(ctx->sq_ftmx, ft.size()));
  NNPE_CHECK_HIP(ctx, hipMemcpy(ctx->sq_ftmx, ft.data(), ft.size() * sizeof(double), hipMemcpyHostToDevice));
  NNPE_CHECK_HIP(ctx, dev_alloc(ctx->sq_out, size_t(ctx->npair) * nq));
  return NNPE_OK;
}

int sq_enqueue(nnpe_ctx* ctx) {
  hipLaunchKernelGGL(k_sq, dim3(ctx->sq_nq), dim3(256), 0, ctx->stream, ctx->sq_nq, ctx->gr_nbin, ctx->npair, ctx->sq_ftmx,
                     ctx->gr_h, ctx->sq_out);
  NNPE_CHECK_LAUNCH(ctx);
  return NNPE_OK;
}

extern "C" int nnpe_sq_compute(nnpe_ctx* ctx, double* out) {
  if (!ctx) return NNPE_ERR_ARG;
  if (!ctx->sq_ftmx) return ctx_fail(ctx, NNPE_ERR_STATE, "nnpe_sq_compute before nnpe_sq_init");
  int rc = sq_enqueue(ctx);
  if (rc) return rc;
  if (out) {
    NNPE_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    NNPE_CHECK_HIP(ctx, hipMemcpy(out, ctx->sq_out, sizeof(double) * ctx->npair * ctx->sq_nq, hipMemcpyDeviceToHost));
  }
  return NNPE_OK;
}
