// nnp.hip -- NNP environment descriptors and NNP forces.
//
// Two descriptor forms share the MLP (mlp.hip) through the per-atom arrays G [N][D] (input of
// the network) and dG [N][D] (dE/dG, its output), both indexed by row:
//
// (1) NNPOTENTIAL_TBL, the reference's form (src/pair_nnp.cpp:433-793): 20 radial descriptors
//     per atom, G_i[k] = sum_j lerp(tbl_desc[pair][20*special + k]); the force of a pair is
//     fpair = lerp(tbl_f_pair[pair][special]) + sum_k (dEdG_i[k] + dEdG_j[k]) * lerp(tbl_f_desc)
//     with the reference's arithmetic.  Tables are re-laid [pair][special][bin][k] so the 20
//     values a lerp needs are two contiguous 80-byte runs.
//
// (2) Symmetry functions, radial + separable angular (the benchmark workloads):
//       G^R_{i,a,n}   = sum_{j in a} exp(-eta_r (r_ij - mu_n)^2) fc(r_ij)
//       G^A_{i,ab,m,zeta,lambda} = 2^(1-zeta) sum_{j in a, k in b, j != k}
//                                   (1 + lambda cos t_jik)^zeta w_m(r_ij) w_m(r_ik)
//     with w_m(r) = exp(-eta_m r^2) fc(r), fc(r) = (cos(pi r/rc) + 1)/2, zeta in {1,2},
//     lambda = +-1.  The angular sums are evaluated from per-element moment sums
//     A0 = sum w, A1 = sum w r^, A2 = sum w r^ r^, S = sum w^2 (one pass over the neighbours,
//     O(N * neighbours) instead of O(N * neighbours^2)).  Forces use the adjoint of those
//     moments: every atom turns dE/dG into dE/d(moments) once, then each pair's force is a
//     function of the pair vector and the two atoms' adjoints, gathered over the full list, so
//     each atom's force has one writer.
#include <algorithm>
#include <cmath>

#include "internal.h"

namespace {

constexpr int kND = 20;        // NNPOTENTIAL_TBL::n_descriptor (src/pair_nnp.h:65)
constexpr int kNRad = 16;
constexpr int kNEta = 4;
constexpr int kNMom = 11;      // A0, A1[3], A2[xx,yy,zz,xy,xz,yz], S
// adjoint rows: one 256-byte slice per element, [4 x 11 angular][16 radial][4 pad], so the slice
// a pair gathers from its neighbour is two aligned 128-byte lines
constexpr int kSlice = 64;
constexpr int kSliceRad = kNEta * kNMom;
constexpr float kPi = 3.14159265358979f;
constexpr int kSfU = 4;  // list entries per lane in flight at once (descriptor and force passes)
// in-range pairs staged per atom between the descriptor pass's scan and sums (one scan step).
// 64 rather than 128: the pass alone is slower (31.0 against 29.2 us), but its 16 KB smaller
// LDS footprint leaves room on the CUs for the pair kernel that runs beside it in the MD step
// (0.1361 -> 0.1341 ms per step)
constexpr int kSfBuf = 64;
constexpr int kTblG = 16;      // lanes per atom, tabulated form
constexpr int kSfG = 16;       // lanes per atom, symmetry functions

__device__ __forceinline__ int pair_index(int a, int b, int ntype) {
  const int i = min(a, b), j = max(a, b);  // small non-negative values: 24-bit multiplies
  return __mul24(i, ntype) - (__mul24(i, i - 1) >> 1) + (j - i);
}

// ---------------------------------------------------------------------------------------------
// (1) tabulated reference form

// NNPOTENTIAL_TBL::computeDescriptor over a full list: each atom sums its own neighbours (a
// group of G lanes per atom, butterfly-reduced)
template <int G>
__global__ void __launch_bounds__(256) k_tbl_desc(int n, int cap, const uint32_t* __restrict__ nbr,
                                                  const int* __restrict__ cnt, const float4* __restrict__ xq,
                                                  const int* __restrict__ typ, int ntype, const BoxPar* __restrict__ boxes,
                                                  const int* __restrict__ rbox,
                                                  float cutsq, float rdr, int nb1, const float* __restrict__ tdesc,
                                                  float* __restrict__ G_out) {
  const int i = (xcd_block(gridDim.x) * 256 + (int)threadIdx.x) / G;
  const int lane = threadIdx.x & (G - 1);
  if (i >= n) return;  // whole group
  const float4 xi = xq[i];
  const int ti = typ[i];
  const float bl = boxes[rbox[i]].bl, m_bl = boxes[rbox[i]].m_bl;
  float acc[kND];
#pragma unroll
  for (int k = 0; k < kND; ++k) acc[k] = 0.0f;
  const int m = cnt[i];
  const uint32_t* row = list_row(nbr, cap, i);
  for (int e = lane; e < m; e += G) {
    const uint32_t ent = row[e];
    const int j = NNPE_ENTRY_J(ent);
    const float4 xj = image_pos(xq[j], NNPE_ENTRY_CODE(ent), bl, m_bl);
    const float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
    const float rsq = dx * dx + dy * dy + dz * dz;
    if (!(rsq < cutsq)) continue;
    const float r = sqrtf(rsq);
    float ipw1 = r * rdr - 0.5f;
    const int t1 = int(ipw1);
    const float ipw2 = ipw1 - float(t1);
    ipw1 = 1.0f - ipw2;
    const int p = pair_index(ti, ntype <= NNPE_ENTRY_MAX_TYPES ? NNPE_ENTRY_T(ent) : typ[j], ntype);
    const float* t = tdesc + ((size_t)(p * 4 + NNPE_ENTRY_SP(ent)) * nb1 + t1) * kND;
#pragma unroll
    for (int k = 0; k < kND; ++k) acc[k] += t[k] * ipw1 + t[kND + k] * ipw2;
  }
#pragma unroll
  for (int k = 0; k < kND; ++k) {
    const float v = group_sum<G>(acc[k]);
    if (k % G == lane) G_out[(size_t)i * kND + k] = v;
  }
}

// NNPOTENTIAL_TBL::compute<ev> over a full list (each directed entry: half energy / virial)
template <bool EV, int G>
__global__ void __launch_bounds__(256) k_tbl_force(int n, int cap, const uint32_t* __restrict__ nbr,
                                                   const int* __restrict__ cnt, const float4* __restrict__ xq,
                                                   const int* __restrict__ typ, int ntype, const BoxPar* __restrict__ boxes,
                                                   const int* __restrict__ rbox,
                                                   float cutsq, float rdr, int nb1, const float* __restrict__ tfdesc,
                                                   const float* __restrict__ tpair, const float* __restrict__ tfpair,
                                                   const float* __restrict__ dG, float4* __restrict__ f, double* pe,
                                                   double* pw) {
  const int i = (xcd_block(gridDim.x) * 256 + (int)threadIdx.x) / G;
  const int lane = threadIdx.x & (G - 1);
  double ea = 0, wa = 0;
  if (i < n) {
    const float4 xi = xq[i];
    const int ti = typ[i];
    const float bl = boxes[rbox[i]].bl, m_bl = boxes[rbox[i]].m_bl;
    float gi[kND];
#pragma unroll
    for (int k = 0; k < kND; ++k) gi[k] = dG[(size_t)i * kND + k];
    float fx = 0, fy = 0, fz = 0, e = 0, w = 0;
    const int m = cnt[i];
    const uint32_t* row = list_row(nbr, cap, i);
    for (int q = lane; q < m; q += G) {
      const uint32_t ent = row[q];
      const int j = NNPE_ENTRY_J(ent), sp = NNPE_ENTRY_SP(ent);
      const float4 xj = image_pos(xq[j], NNPE_ENTRY_CODE(ent), bl, m_bl);
      float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
      const float rsq = dx * dx + dy * dy + dz * dz;
      if (!(rsq < cutsq)) continue;
      const float r = sqrtf(rsq);
      float ipw1 = r * rdr - 0.5f;
      const int t1 = int(ipw1), t2 = t1 + 1;
      const float ipw2 = ipw1 - float(t1);
      ipw1 = 1.0f - ipw2;
      const int p = pair_index(ti, ntype <= NNPE_ENTRY_MAX_TYPES ? NNPE_ENTRY_T(ent) : typ[j], ntype);
      const float* fp = tfpair + (size_t)(p * 4 + sp) * nb1;
      float fpair = fp[t1] * ipw1 + fp[t2] * ipw2;
      const float* t = tfdesc + ((size_t)(p * 4 + sp) * nb1 + t1) * kND;
      const float* gj = dG + (size_t)j * kND;
#pragma unroll
      for (int k = 0; k < kND; ++k) fpair += (gi[k] + gj[k]) * (t[k] * ipw1 + t[kND + k] * ipw2);
      fx += (dx *= fpair); fy += (dy *= fpair); fz += (dz *= fpair);
      if (EV) {
        const float* pu = tpair + (size_t)(p * 4 + sp) * nb1;
        e += 0.5f * (pu[t1] * ipw1 + pu[t2] * ipw2);
        w += 0.5f * fpair * rsq;
      }
    }
    fx = group_sum<G>(fx); fy = group_sum<G>(fy); fz = group_sum<G>(fz);
    if (lane == 0) {
      float4 fo = f[i];
      fo.x += fx; fo.y += fy; fo.z += fz;
      f[i] = fo;
    }
    ea = e; wa = w;
  }
  if (EV) block_sum2_store<256>(ea, wa, pe, pw);
}

// ---------------------------------------------------------------------------------------------
// (2) symmetry functions

constexpr int kRadS = 8;  // radial recurrence: the centre evaluated directly

struct SfParams {
  float rc, rcsq, eta_r, rc_inv;
  float eta_r_l2;         // -eta_r / ln 2: exp(-eta_r d^2) = 2^(eta_r_l2 d^2)
  float mu[kNRad];
  float eta[kNEta];
  float eta_l2[kNEta];    // -eta / ln 2
  // equally spaced centres mu_k = mu_s + (k - s) delta (s = kRadS):
  float rad_step_l2;      // 2 eta_r delta / ln 2
  float rad_rho[kRadS];   // exp(-eta_r delta^2 (2j + 1)), j = 0 .. kRadS-1
  float rad_q[kNRad];     // Q_k = exp(-eta_r delta^2 (k - s)^2): g_k = g_s R^(k-s) Q_k
  float rad_q_l2;         // -eta_r delta^2 / ln 2
  float rad_delta;        // delta
  // +1 and -1 as run-time values: a half-precision operand times one of them plus an FP32 value
  // is a single mixed-precision FMA (the compiler would turn a literal +-1 into an add after a
  // separate conversion)
  float one, m_one;
};

// g[k] = exp(-eta_r (r - mu_k)^2), k = 0 .. kNRad-1.  RR (equally spaced centres): g at the
// centre s and the step ratio R = exp(2 eta_r delta (r - mu_s)) from the hardware exp2, the rest
// from the exact identities g_{s+j+1} = g_{s+j} R rho_j and g_{s-j-1} = g_{s-j} rho_j / R,
// rho_j = exp(-eta_r delta^2 (2j + 1)): three exponentials instead of sixteen.  Walking outwards
// from the middle keeps every intermediate within [0, R] (no overflow); a product that
// underflows to 0 is followed only by smaller true values.
template <bool RR>
__device__ __forceinline__ void sf_gauss(float r, const SfParams& P, float* g) {
#pragma clang fp contract(fast)
  if constexpr (RR) {
    const float e = r - P.mu[kRadS];
    g[kRadS] = __builtin_amdgcn_exp2f(P.eta_r_l2 * e * e);
    const float l = P.rad_step_l2 * e;
    const float R = __builtin_amdgcn_exp2f(l), Ri = __builtin_amdgcn_exp2f(-l);
#pragma unroll
    for (int j = 0; kRadS + j + 1 < kNRad; ++j) g[kRadS + j + 1] = g[kRadS + j] * R * P.rad_rho[j];
#pragma unroll
    for (int j = 0; j < kRadS; ++j) g[kRadS - j - 1] = g[kRadS - j] * Ri * P.rad_rho[j];
  } else {
#pragma unroll
    for (int k = 0; k < kNRad; ++k) {
      const float d = r - P.mu[k];
      g[k] = __builtin_amdgcn_exp2f(P.eta_r_l2 * d * d);
    }
  }
}

constexpr int kHEta = kNEta / 2, kHRad = kNRad / 2;  // per half of a descriptor lane group

// The 8 radial Gaussians of half hg of the centres, in the order the half accumulates them.
// RR: from the middle centre s = 8 outwards, the same products as sf_gauss -- half 1 holds
// g_8, g_9, ..., g_15 (g_{8+t+1} = g_{8+t} R rho_t), half 0 holds g_7, g_6, ..., g_0
// (g_{8-t-1} = g_{8-t} R^-1 rho_t); without RR, centres 8 hg + t directly.
template <bool RR>
__device__ __forceinline__ void sf_gauss_half(float r, const SfParams& P, int hg, float* g) {
#pragma clang fp contract(fast)
  static_assert(kRadS == kHRad, "the recurrence starts at the first centre of the upper half");
  if constexpr (RR) {
    const float e = r - P.mu[kRadS];
    const float g8 = __builtin_amdgcn_exp2f(P.eta_r_l2 * e * e);
    const float l = P.rad_step_l2 * e;
    const float Rx = __builtin_amdgcn_exp2f(hg ? l : -l);
    float cur = hg ? g8 : g8 * Rx * P.rad_rho[0];
    g[0] = cur;
#pragma unroll
    for (int t = 1; t < kHRad; ++t) {
      cur = cur * Rx * (hg ? P.rad_rho[t - 1] : P.rad_rho[t]);
      g[t] = cur;
    }
  } else {
#pragma unroll
    for (int t = 0; t < kHRad; ++t) {
      const float d = r - (hg ? P.mu[kHRad + t] : P.mu[t]);
      g[t] = __builtin_amdgcn_exp2f(P.eta_r_l2 * d * d);
    }
  }
}

// radial centre of entry t of half hg (the order of sf_gauss_half)
template <bool RR>
__device__ __forceinline__ int sf_half_centre(int hg, int t) {
  if constexpr (RR) return hg ? kRadS + t : kRadS - 1 - t;
  else return kHRad * hg + t;
}

// fc = (cos(pi r/rc) + 1)/2 and dfc/dr for 0 <= r <= rc.  With y = pi (r/rc - 1/2), |y| <= pi/2:
// cos(pi r/rc) = -sin(y), sin(pi r/rc) = cos(y), both from their Taylor series about 0 (sin to
// y^11, cos to y^12; truncation below 6e-8 on the range), evaluated by Horner's rule.
__device__ __forceinline__ void sf_cut(float r, float rc_inv, float& fc, float& dfc) {
#pragma clang fp contract(fast)
  const float y = kPi * (r * rc_inv - 0.5f), y2 = y * y;
  const float s =
      y * (1.0f + y2 * (-1.0f / 6 + y2 * (1.0f / 120 + y2 * (-1.0f / 5040 + y2 * (1.0f / 362880 + y2 * (-1.0f / 39916800))))));
  const float c =
      1.0f + y2 * (-0.5f + y2 * (1.0f / 24 + y2 * (-1.0f / 720 + y2 * (1.0f / 40320 + y2 * (-1.0f / 3628800 + y2 * (1.0f / 479001600))))));
  fc = 0.5f * (1.0f - s);
  dfc = -0.5f * kPi * rc_inv * c;
}

// moments per atom: [NT][kNEta][kNMom] then radial [NT][kNRad].  A group of G lanes per atom,
// one pass per neighbour element a over the atom's row of that element's short list (so only
// that element's 60 accumulators are live and no neighbour type is looked up): each lane accumulates the moments of its share of the neighbours, the
// group butterfly-sums them, and the lanes split the stores (lane l writes the values whose
// index is l mod G) to the moment rows and to an LDS copy, from which the angular descriptors
// are formed at the end.
// five waves per SIMD (96 VGPRs, two spilled): beside the pair kernel in the MD step 0.1332 ->
// 0.1327 ms per step against the compiler's 101 VGPRs at four
template <int NT, int G, bool RR>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) k_sf_desc(int n, int stride, int cap, const uint32_t* __restrict__ nbr,
                                                 const int* __restrict__ cnt, const float4* __restrict__ xq,
                                                 const BoxPar* __restrict__ boxes, const int* __restrict__ rbox,
                                                 SfParams P, float* __restrict__ Mout,
                                                 float* __restrict__ G_out, uint32_t* __restrict__ act,
                                                 int* __restrict__ act_cnt) {
#pragma clang fp contract(fast)
  constexpr int NP = NT * (NT + 1) / 2;
  constexpr int NMT = NT * kNEta * kNMom + NT * kNRad;
  constexpr int RAD0 = NT * kNEta * kNMom;
  constexpr int D = NT * kNRad + NP * kNEta * 4;
  __shared__ __align__(16) float s_mom[256 / G][NMT];
  __shared__ float4 s_pair[256 / G][kSfBuf];
  __shared__ float4 s_img[27];
  image_table_fill(s_img);
  __syncthreads();
  static_assert(G == 16, "the moment sums split the 16 lanes of an atom into two halves");
  static_assert(NMT % 4 == 0, "moment rows are whole 16-byte chunks");
  const int blk = xcd_block(gridDim.x);
  const int i = (blk * 256 + (int)threadIdx.x) / G;
  const int lane = threadIdx.x & (G - 1), grp = threadIdx.x / G;
  const int hg = lane >> 3, hl = lane & 7;  // half of the group and lane within it
  const bool live = i < n;
  float* sm = s_mom[grp];
  if (live) {
    const float4 xi = xq[i];
    const float bl = boxes[rbox[i]].bl;
#pragma unroll
    for (int a = 0; a < NT; ++a) {
      const int cn = cnt[(size_t)a * stride + i];
      const uint32_t* row = nbr + ((size_t)a * stride + i) * cap;
      // the sums: lanes 0-7 of the group hold exponents 0-1 and the lower 8 radial centres,
      // lanes 8-15 exponents 2-3 and the upper 8 (half the accumulators live per lane)
      float A[kHEta][kNMom];
      float R[kHRad];
#pragma unroll
      for (int m = 0; m < kHEta; ++m)
#pragma unroll
        for (int c = 0; c < kNMom; ++c) A[m][c] = 0.0f;
#pragma unroll
      for (int k = 0; k < kHRad; ++k) R[k] = 0.0f;
      // in-range entries are also compacted, in list order, into this step's active list
      // (a group-masked ballot gives each lane its slot), so the force pass has no idle lanes
      uint32_t* arow = act + ((size_t)a * stride + i) * cap;
      const unsigned long long gmask = (G == 64 ? ~0ull : ((1ull << G) - 1ull)) << (threadIdx.x & 63 & ~(G - 1));
      const unsigned long long below = (1ull << (threadIdx.x & 63)) - 1ull;
      int na = 0;
      float4* buf = s_pair[grp];
      // Phase 1: all lanes scan the list (kSfU entries per lane and their position gathers in
      // flight together), and a group-masked ballot compacts the pairs within the cutoff, in
      // list order, into the active list and into the group's LDS buffer as (u, |u|^2).
      // Phase 2: all lanes accumulate the buffered pairs, so the exp-heavy moment arithmetic
      // runs on full lanes instead of under the scan's partial mask.
      static_assert(kSfBuf >= kSfU * G, "the scan must fit one step into the pair buffer");
      // entries of the next step are loaded while this step's positions are gathered and tested
      uint32_t en[kSfU];
#pragma unroll
      for (int u = 0; u < kSfU; ++u) {
        const int q = u * G + lane;
        en[u] = q < cn ? row[q] : 0u;
      }
      for (int q0 = 0; q0 < cn;) {
        int nb = 0;
        for (; q0 < cn && nb + kSfU * G <= kSfBuf; q0 += kSfU * G) {
          uint32_t ec[kSfU];
          float4 xr[kSfU];
#pragma unroll
          for (int u = 0; u < kSfU; ++u) ec[u] = en[u];
#pragma unroll
          for (int u = 0; u < kSfU; ++u) {
            const int q = q0 + u * G + lane;
            xr[u] = xq[q < cn ? (int)NNPE_ENTRY_J(ec[u]) : i];
          }
#pragma unroll
          for (int u = 0; u < kSfU; ++u) {
            const int q = q0 + (kSfU + u) * G + lane;
            en[u] = q < cn ? row[q] : 0u;
          }
#pragma unroll
          for (int u = 0; u < kSfU; ++u) {
            if (q0 + u * G >= cn) break;  // uniform across the group
            const int q = q0 + u * G + lane;
            // lanes past the count hold entry 0 and their own position: computed, never taken
            const float4 xj = image_pos_t(xr[u], NNPE_ENTRY_CODE(ec[u]), bl, s_img);
            const float ux = xj.x - xi.x, uy = xj.y - xi.y, uz = xj.z - xi.z;
            const float rsq = dist2_rn(ux, uy, uz);  // rounded as written: the force pass makes the same cut
            const bool in = q < cn && rsq < P.rcsq;
            const unsigned long long bal = __ballot(in) & gmask;
            if (in) {
              const int slot = __popcll(bal & below);
              arow[(unsigned)(na + slot)] = ec[u];
              buf[nb + slot] = make_float4(ux, uy, uz, rsq);
            }
            const int c = __popcll(bal);
            na += c;
            nb += c;
          }
        }
        __builtin_amdgcn_wave_barrier();  // a group lives in one wave: its LDS writes are ordered
        for (int k = hl; k < nb; k += G / 2) {
          const float4 pu = buf[k];
          const float rsq = pu.w;
          const float ir = __builtin_amdgcn_rsqf(rsq), r = rsq * ir;
          const float hx = pu.x * ir, hy = pu.y * ir, hz = pu.z * ir;
          float fc, dfc;
          sf_cut(r, P.rc_inv, fc, dfc);
          float gk[kHRad];
          sf_gauss_half<RR>(r, P, hg, gk);
#pragma unroll
          for (int k2 = 0; k2 < kHRad; ++k2) R[k2] += gk[k2] * fc;
#pragma unroll
          for (int m = 0; m < kHEta; ++m) {
            const float el2 = hg ? P.eta_l2[kHEta + m] : P.eta_l2[m];
            const float w = __builtin_amdgcn_exp2f(el2 * rsq) * fc, wx = w * hx, wy = w * hy, wz = w * hz;
            A[m][0] += w;
            A[m][1] += wx; A[m][2] += wy; A[m][3] += wz;
            A[m][4] += wx * hx; A[m][5] += wy * hy; A[m][6] += wz * hz;
            A[m][7] += wx * hy; A[m][8] += wx * hz; A[m][9] += wy * hz;
            A[m][10] += w * w;
          }
        }
        __builtin_amdgcn_wave_barrier();  // buffer consumed before the next scan overwrites it
      }
      if (lane == 0) act_cnt[(size_t)a * stride + i] = na;
      // The half's 30 sums (angular m*11 + c, then the 8 radial centres; 2 zero pads) are
      // reduce-scattered over its 8 lanes in three exchange steps (half-row mirror, then
      // distance 2 and 1 in the quad): at each step a lane keeps one half of its block and adds
      // the partner's copy of it, so lane hl ends with the full sums of values 4 hl .. 4 hl + 3
      // and writes just those to the LDS moment row (45 DPP adds and 4 stores per lane instead of
      // a 120-add butterfly and a conditional store per value)
      constexpr int NV = kHEta * kNMom + kHRad;
      static_assert(NV <= 32 && G == 16, "a half's sums fit 8 lanes x 4 values");
      float v[32];
#pragma unroll
      for (int m = 0; m < kHEta; ++m)
#pragma unroll
        for (int c = 0; c < kNMom; ++c) v[m * kNMom + c] = A[m][c];
#pragma unroll
      for (int k = 0; k < kHRad; ++k) v[kHEta * kNMom + k] = R[k];
#pragma unroll
      for (int k = NV; k < 32; ++k) v[k] = 0.0f;
      const bool up4 = (hl & 4) != 0, up2 = (hl & 2) != 0, up1 = (hl & 1) != 0;
      float w[16], x[8], y[4];
#pragma unroll
      for (int t = 0; t < 16; ++t) w[t] = (up4 ? v[16 + t] : v[t]) + dpp_f32<0x141>(up4 ? v[t] : v[16 + t]);
#pragma unroll
      for (int t = 0; t < 8; ++t) x[t] = (up2 ? w[8 + t] : w[t]) + dpp_f32<0x4E>(up2 ? w[t] : w[8 + t]);
#pragma unroll
      for (int t = 0; t < 4; ++t) y[t] = (up1 ? x[4 + t] : x[t]) + dpp_f32<0xB1>(up1 ? x[t] : x[4 + t]);
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int vv = 4 * hl + t;
        if (vv < kHEta * kNMom) sm[(a * kNEta + kHEta * hg) * kNMom + vv] = y[t];
        else if (vv < NV) sm[RAD0 + a * kNRad + sf_half_centre<RR>(hg, vv - kHEta * kNMom)] = y[t];
      }
    }
  }
  __syncthreads();
  // the workgroup's moment rows, contiguous in LDS and in Mout, copied out whole
  {
    const int i0 = blk * (256 / G), nl = min(256 / G, n - i0);
    const float4* src = reinterpret_cast<const float4*>(&s_mom[0][0]);
    float4* dst = reinterpret_cast<float4*>(Mout + (size_t)i0 * NMT);
    for (int q = threadIdx.x; q < nl * (NMT / 4); q += 256) dst[q] = src[q];
  }
  if (!live) return;
  float* g = G_out + (size_t)i * D;
  for (int k = lane; k < NT * kNRad; k += G) g[k] = sm[RAD0 + k];
  // angular: pair (a, b), exponent m -> 4 values; entry t of the NP * kNEta list per lane
  for (int t = lane; t < NP * kNEta; t += G) {
    const int pp = t / kNEta, m = t - pp * kNEta;
    int a = 0, b = 0, rem = pp;
    while (rem >= NT - a) { rem -= NT - a; ++a; }
    b = a + rem;
    const float* X = sm + (a * kNEta + m) * kNMom;
    const float* Y = sm + (b * kNEta + m) * kNMom;
    float P0 = X[0] * Y[0], P1 = X[1] * Y[1] + X[2] * Y[2] + X[3] * Y[3];
    float P2 = X[4] * Y[4] + X[5] * Y[5] + X[6] * Y[6] + 2.0f * (X[7] * Y[7] + X[8] * Y[8] + X[9] * Y[9]);
    if (a == b) { P0 -= X[10]; P1 -= X[10]; P2 -= X[10]; }
    float* o = g + NT * kNRad + t * 4;
    o[0] = P0 + P1;
    o[1] = P0 - P1;
    o[2] = 0.5f * (P0 + 2.0f * P1 + P2);
    o[3] = 0.5f * (P0 - 2.0f * P1 + P2);
  }
}

// dE/d(moments) of each atom from dE/dG and its moments.  A workgroup takes 32 atoms; lane
// (atom, element a, exponent m) forms the 11 adjoints of moment group (a, m) -- summing the
// partner elements b in order, each pair's terms in the order of the pair loop (x <= y) of the
// descriptor definition, the self pair twice -- and copies four of element a's sixteen radial
// dE/dG into the slice.  The slices are assembled in LDS and stored chunk by chunk with the atom
// index fastest across the lanes, so each store instruction writes whole runs of the
// structure-of-arrays chunk rows (instead of 4-byte pieces of 16-byte chunks).
template <int NT, bool HALF>
__global__ void __launch_bounds__(32 * kNEta * NT) k_sf_adjoint(int n, int stride, const float* __restrict__ M,
                                                               const float* __restrict__ dG, float* __restrict__ adj,
                                                               uint4* __restrict__ adjh) {
#pragma clang fp contract(fast)
  constexpr int NMT = NT * kNEta * kNMom + NT * kNRad;
  constexpr int D = NT * kNRad + NT * (NT + 1) / 2 * kNEta * 4;
  constexpr int NTH = 32 * kNEta * NT;
  constexpr int kRow = kSlice + 4;  // padded LDS row: the store phase's lanes read different banks
  __shared__ __align__(16) float s32[32 * NT][kRow];
  __shared__ __align__(16) _Float16 s16[HALF ? 32 * NT : 1][kSlice + 8];
  const int tid = threadIdx.x;
  const int m = tid % kNEta, a = (tid / kNEta) % NT, il = tid / (kNEta * NT);
  const int i0 = blockIdx.x * 32, i = i0 + il;
  static_assert(kSliceRad % 4 == 0 && kNRad == 4 * kNEta, "one radial chunk per exponent lane");
  if (i < n) {
    const float* mi = M + (size_t)i * NMT;
    const float* g = dG + (size_t)i * D;
    float dA[kNMom];
#pragma unroll
    for (int c = 0; c < kNMom; ++c) dA[c] = 0.0f;
#pragma unroll
    for (int b = 0; b < NT; ++b) {
      const int x = min(a, b), y = max(a, b);
      const int pp = x * NT - (x * (x - 1)) / 2 + (y - x);  // index of the pair (x, y), x <= y
      const float* gg = g + NT * kNRad + (pp * kNEta + m) * 4;
      const float u0 = gg[0] + gg[1] + 0.5f * (gg[2] + gg[3]);
      const float u1 = gg[0] - gg[1] + gg[2] - gg[3];
      const float u2 = 0.5f * (gg[2] + gg[3]);
      const float* Z = mi + (b * kNEta + m) * kNMom;  // the partner's moments
      for (int rep = 0; rep < (a == b ? 2 : 1); ++rep) {
        dA[0] += u0 * Z[0];
#pragma unroll
        for (int c = 1; c <= 3; ++c) dA[c] += u1 * Z[c];
#pragma unroll
        for (int c = 4; c <= 6; ++c) dA[c] += u2 * Z[c];
#pragma unroll
        for (int c = 7; c <= 9; ++c) dA[c] += 2.0f * u2 * Z[c];
      }
      if (a == b) dA[10] -= u0 + u1 + u2;
    }
    // The slice stores the off-diagonal second-moment adjoints (components 7..9, each the
    // adjoint of a product that appears twice in the symmetric matrix) halved, so the force pass
    // forms M h with plain FMAs.
#pragma unroll
    for (int c = 7; c <= 9; ++c) dA[c] *= 0.5f;
    const float* gr = g + a * kNRad + 4 * m;
    float* sl = s32[il * NT + a];
#pragma unroll
    for (int c = 0; c < kNMom; ++c) sl[m * kNMom + c] = dA[c];
#pragma unroll
    for (int k = 0; k < 4; ++k) sl[kSliceRad + 4 * m + k] = gr[k];
    if (m == 0)
#pragma unroll
      for (int k = kSliceRad + kNRad; k < kSlice; ++k) sl[k] = 0.0f;
    if constexpr (HALF) {
      // saturate instead of rounding to infinity (|adjoint| beyond 65504 would be a broken network)
      auto h16 = [](float x) { return (_Float16)fminf(fmaxf(x, -65504.0f), 65504.0f); };
      _Float16* sh = s16[il * NT + a];
#pragma unroll
      for (int c = 0; c < kNMom; ++c) sh[m * kNMom + c] = h16(dA[c]);
#pragma unroll
      for (int k = 0; k < 4; ++k) sh[kSliceRad + 4 * m + k] = h16(gr[k]);
      if (m == 0)
#pragma unroll
        for (int k = kSliceRad + kNRad; k < kSlice; ++k) sh[k] = (_Float16)0.0f;
    }
  }
  __syncthreads();
  // chunk q of element e of row j at adj4[(e * 16 + q) * stride + j] and (FP16, 8 values per
  // chunk) adjh[(e * 8 + q) * stride + j]
  const int nat = min(32, n - i0);
  float4* adj4 = reinterpret_cast<float4*>(adj);
  for (int idx = tid; idx < 32 * NT * (kSlice / 4); idx += NTH) {
    const int j = idx & 31, rest = idx >> 5, q = rest % (kSlice / 4), e = rest / (kSlice / 4);
    if (j < nat) adj4[(size_t)(e * (kSlice / 4) + q) * stride + i0 + j] = reinterpret_cast<const float4*>(s32[j * NT + e])[q];
  }
  if constexpr (HALF) {
    for (int idx = tid; idx < 32 * NT * (kSlice / 8); idx += NTH) {
      const int j = idx & 31, rest = idx >> 5, q = rest % (kSlice / 8), e = rest / (kSlice / 8);
      if (j < nat) adjh[(size_t)(e * (kSlice / 8) + q) * stride + i0 + j] = reinterpret_cast<const uint4*>(s16[j * NT + e])[q];
    }
  }
}

// dE_owner/du for one neighbour at unit vector h (u = r h) of the given element slice of an
// adjoint row (off-diagonal components 7..9 stored halved): the angular part (vector) and the
// radial part's scalar folded into `along`
__device__ __forceinline__ void sf_grad_ang(const float* __restrict__ d, const float* wv, const float* dwv, float ir,
                                            float hx, float hy, float hz, float& along, float& gx, float& gy, float& gz) {
#pragma clang fp contract(fast)
#pragma unroll
  for (int m = 0; m < kNEta; ++m) {
    const float* e = d + m * kNMom;
    const float w = wv[m], dw = dwv[m], wr = w * ir;
    const float p1 = e[1] * hx + e[2] * hy + e[3] * hz;
    const float mrx = e[4] * hx + e[7] * hy + e[8] * hz;
    const float mry = e[5] * hy + e[7] * hx + e[9] * hz;
    const float mrz = e[6] * hz + e[8] * hx + e[9] * hy;
    const float qq = mrx * hx + mry * hy + mrz * hz;
    along += dw * (e[0] + 2.0f * w * e[10] + p1 + qq) - wr * (p1 + 2.0f * qq);
    gx += wr * (e[1] + 2.0f * mrx);
    gy += wr * (e[2] + 2.0f * mry);
    gz += wr * (e[3] + 2.0f * mrz);
  }
}

// A group of G lanes per atom over the step's active lists (the pairs within the cutoff, one
// list per neighbour element, compacted by the descriptor pass); the atom's own adjoint row is
// staged in LDS (one row per group), each lane gathers the neighbour's element slice of its
// pairs from global memory.  Own and neighbour terms come from one evaluation: the force of the
// pair on atom i is dE_i/du - dE_j/du' (u = x_j - x_i, u' = -u), and since the angular term is
// linear in the adjoint and the neighbour's is taken at -h, where the moment components of odd
// degree in h (1..3) change sign, one evaluation on own + neighbour for the even components and
// own - neighbour for the odd ones gives it.
// RR (equally spaced radial centres, within the range the host checked): the radial force is
// a pair of polynomials in R = exp(2 eta_r delta (r - mu_s)) and 1/R.  With
// g_{s+-j} = g_s R^(+-j) Q_j, the sum over centres of c_k d/dr[g_k fc] with c_k the two atoms'
// radial adjoints is
//   S0 (dfc - 2 eta_r fc (r - mu_s)) + 2 eta_r delta fc T,
//   S0 = sum_k c_k g_k = g_s (U(R) + X L(X)),  T = sum_k (k - s) c_k g_k = g_s (R U'(R) - X (L + X L')),
// X = 1/R, U and L polynomials with coefficients c_k Q_k, evaluated with their derivatives by
// Horner's rule (two FMAs per centre).  The own adjoints are staged pre-multiplied by Q_k.
// VV: this pass finishes each atom's force (the classical terms were added before), so it also
// applies the FP64 velocity half-kick (vv_kick64) and stores the acceleration, in place of a
// separate Verlet pass.
// HALF (16-bit network forms): the neighbour's slice is gathered from its FP16 copy, 8 values
// per 16-byte load (8 loads per pair instead of 15); the atom's own slice stays FP32, and each
// half value enters through a mixed-precision FMA with an FP32 operand.
// EV: the virial of the pair, -(F_i . u), halved (each pair appears in both atoms' lists).
template <int NT, bool EV, int G, bool RR, bool VV, bool HALF>
// three waves per SIMD (at most 168 VGPRs): measured faster than the compiler's own 178-198
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) k_sf_force(int n, int stride, int cap, const uint32_t* __restrict__ nbr,
                                                  const int* __restrict__ cnt, const float4* __restrict__ xq,
                                                  const int* __restrict__ typ, const BoxPar* __restrict__ boxes,
                                                  const int* __restrict__ rbox, SfParams P,
                                                  const float* __restrict__ adj, const uint4* __restrict__ adjh,
                                                  float4* __restrict__ f, double* pe, double* pw, VvArgs V) {
#pragma clang fp contract(fast)
  constexpr int NMT = NT * kSlice;
  constexpr int QR0 = kSliceRad / 4, QR1 = (kSliceRad + kNRad) / 4;  // radial chunks of a slice
  __shared__ __align__(16) float s_adj[256 / G][NMT];
  __shared__ float4 s_img[27];
  image_table_fill(s_img);
  const int i = (xcd_block(gridDim.x) * 256 + (int)threadIdx.x) / G;
  const int lane = threadIdx.x & (G - 1), grp = threadIdx.x / G;
  const float4* __restrict__ adj4 = reinterpret_cast<const float4*>(adj);
  if (i < n)
    for (int q = lane; q < NMT / 4; q += G) {
      float4 v = adj4[(size_t)q * stride + i];
      const int qs = q % (kSlice / 4);
      if (RR && qs >= QR0 && qs < QR1) {
        // own radial adjoints times Q_k, Q_k = 2^(rad_q_l2 (k - s)^2)
        const float d0 = float(4 * (qs - QR0) - kRadS);
        v.x *= __builtin_amdgcn_exp2f(P.rad_q_l2 * d0 * d0);
        v.y *= __builtin_amdgcn_exp2f(P.rad_q_l2 * (d0 + 1.0f) * (d0 + 1.0f));
        v.z *= __builtin_amdgcn_exp2f(P.rad_q_l2 * (d0 + 2.0f) * (d0 + 2.0f));
        v.w *= __builtin_amdgcn_exp2f(P.rad_q_l2 * (d0 + 3.0f) * (d0 + 3.0f));
      }
      reinterpret_cast<float4*>(s_adj[grp])[q] = v;
    }
  __syncthreads();
  double wacc = 0;
  if (i < n) {
    const float4 xi = xq[i];
    const int ti = typ[i];
    const float bl = boxes[rbox[i]].bl;
    const float* ai = s_adj[grp];
    float fx = 0, fy = 0, fz = 0, vir = 0;
#pragma unroll
    for (int tj = 0; tj < NT; ++tj) {
    const int cn = cnt[(size_t)tj * stride + i];
    const uint32_t* row = nbr + ((size_t)tj * stride + i) * cap;
    // kSfU pairs per lane: entries and positions gathered together, then the pairs in order
    for (int q0 = lane; q0 < cn; q0 += kSfU * G) {
    uint32_t ec[kSfU];
    float4 xr[kSfU];
#pragma unroll
    for (int u = 0; u < kSfU; ++u) {
      const int q = q0 + u * G;
      ec[u] = q < cn ? row[q] : 0u;
    }
#pragma unroll
    for (int u = 0; u < kSfU; ++u) xr[u] = xq[q0 + u * G < cn ? (int)NNPE_ENTRY_J(ec[u]) : i];
#pragma unroll
    for (int u = 0; u < kSfU; ++u) {
      if (q0 + u * G >= cn) break;
      const uint32_t ent = ec[u];
      const int j = NNPE_ENTRY_J(ent);
      const float4 xj = image_pos_t(xr[u], NNPE_ENTRY_CODE(ent), bl, s_img);
      const float ux = xj.x - xi.x, uy = xj.y - xi.y, uz = xj.z - xi.z;
      const float rsq = dist2_rn(ux, uy, uz);
      if (!(rsq < P.rcsq)) continue;
      const float ir = __builtin_amdgcn_rsqf(rsq), r = rsq * ir;
      const float hx = ux * ir, hy = uy * ir, hz = uz * ir;
      float fc, dfc;
      sf_cut(r, P.rc_inv, fc, dfc);
      float wv[kNEta], dwv[kNEta];
#pragma unroll
      for (int m = 0; m < kNEta; ++m) {
        const float ex = __builtin_amdgcn_exp2f(P.eta_l2[m] * rsq);
        wv[m] = ex * fc;
        dwv[m] = ex * (dfc - 2.0f * P.eta[m] * r * fc);
      }
      // the neighbour's slice for this atom's element, 16-byte chunk q at adj4[(ti*16 + q) * stride
      // + j]: the lanes' neighbours are rows of a few cells, so a wave's chunk loads share lines
      const float* as = ai + tj * kSlice;                      // own slice for the neighbour's element
      // c[k] = own + neighbour (odd angular components own - neighbour); radial (RR) times Q_k
      float e[kSliceRad + kNRad];
      if constexpr (HALF) {
        _Float16 hv[kSlice];
#pragma unroll
        for (int q = 0; q < kSlice / 8; ++q) {
          const uint4 v = adjh[(size_t)(ti * (kSlice / 8) + q) * stride + j];
          __builtin_memcpy(hv + 8 * q, &v, 16);
        }
#pragma unroll
        for (int k = 0; k < kSliceRad + kNRad; ++k) {
          const int c = k % kNMom;
          const float mul = k >= kSliceRad ? (RR ? P.rad_q[k - kSliceRad] : P.one)
                                           : ((c >= 1 && c <= 3) ? P.m_one : P.one);
          e[k] = fmaf((float)hv[k], mul, as[k]);
        }
      } else {
#pragma unroll
        for (int q = 0; q < (kSliceRad + kNRad) / 4; ++q) {
          const float4 v = adj4[(size_t)(ti * (kSlice / 4) + q) * stride + j];
          const float vv[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
          for (int t = 0; t < 4; ++t) {
            const int k = 4 * q + t, c = k % kNMom;
            if (k >= kSliceRad) e[k] = RR ? fmaf(vv[t], P.rad_q[k - kSliceRad], as[k]) : as[k] + vv[t];
            else e[k] = (c >= 1 && c <= 3) ? as[k] - vv[t] : as[k] + vv[t];
          }
        }
      }
      const float* ce = e + kSliceRad;  // radial coefficients
      const float m2efc = -2.0f * P.eta_r * fc;
      float rad;
      if constexpr (RR) {
        const float d = r - P.mu[kRadS];
        const float gs = __builtin_amdgcn_exp2f(P.eta_r_l2 * d * d);
        const float l = P.rad_step_l2 * d;
        const float R = __builtin_amdgcn_exp2f(l), X = __builtin_amdgcn_exp2f(-l);
        // U(R) = sum_{j=0}^{7} c_{s+j} R^j and U'(R)
        float up = ce[kNRad - 1], ud = 0.0f;
#pragma unroll
        for (int k = kNRad - 2; k >= kRadS; --k) {
          ud = fmaf(ud, R, up);
          up = fmaf(up, R, ce[k]);
        }
        // L(X) = sum_{i=0}^{7} c_{s-1-i} X^i and L'(X)
        float lp = ce[0], ld = 0.0f;
#pragma unroll
        for (int k = 1; k < kRadS; ++k) {
          ld = fmaf(ld, X, lp);
          lp = fmaf(lp, X, ce[k]);
        }
        const float S0 = gs * fmaf(X, lp, up);
        const float T = gs * (R * ud - X * fmaf(X, ld, lp));
        rad = S0 * fmaf(m2efc, d, dfc) - m2efc * P.rad_delta * T;
      } else {
        // sum_n g_n'(r) (dR_i[tj][n] + dR_j[ti][n])
        float gk[kNRad];
        sf_gauss<false>(r, P, gk);
        rad = 0.0f;
#pragma unroll
        for (int k = 0; k < kNRad; ++k) rad += ce[k] * (gk[k] * (m2efc * (r - P.mu[k]) + dfc));
      }
      float al = 0, gx = 0, gy = 0, gz = 0;
      sf_grad_ang(e, wv, dwv, ir, hx, hy, hz, al, gx, gy, gz);
      const float px = gx + (al + rad) * hx, py = gy + (al + rad) * hy, pz = gz + (al + rad) * hz;
      fx += px; fy += py; fz += pz;
      if (EV) vir -= 0.5f * (px * ux + py * uy + pz * uz);
    }
    }
    }
    fx = group_sum<G>(fx); fy = group_sum<G>(fy); fz = group_sum<G>(fz);
    if (lane == 0) {
      float4 fo = f[i];
      fo.x += fx; fo.y += fy; fo.z += fz;
      if constexpr (VV) fo = V.adv ? vv_kick64_adv(i, V, fo, xi) : vv_kick64(i, V.dt, V.rm, V.vd, V.ad, fo);
      f[i] = fo;
    }
    wacc = vir;
  }
  if (EV) block_sum2_store<256>(0.0, wacc, pe, pw);
}

}  // namespace

// ---------------------------------------------------------------------------------------------
// host side

struct NnpState {
  // tabulated form
  bool tbl = false;
  float tbl_cut = 0, tbl_cutsq = 0, tbl_rdr = 0;
  int tbl_nbin = 0;
  float *tdesc = nullptr, *tfdesc = nullptr, *tpair = nullptr, *tfpair = nullptr;
  // symmetry functions
  bool sf = false;
  bool rad_recur = false;  // equally spaced radial centres: Gaussians by recurrence
  bool rad_horner = false; // ... and in the range where the force pass sums them as polynomials
  SfParams P{};
  float* M = nullptr;    // moments [N][NMT]
  float* adj = nullptr;  // [N][ntype][kSlice]
  // FP16 copy of the slices for the neighbour gathers of the 16-bit network forms: chunk q (8
  // halves, slice values 8q .. 8q+7) of element e of row j at adjh[(e * 8 + q) * stride + j]
  uint4* adjh = nullptr;
  uint32_t* act = nullptr;  // active lists: pairs within the cutoff, [ntype][stride][act_cap]
  int* act_cnt = nullptr;   // [ntype][stride]
  int act_cap = 0;
  int nmt = 0;
};

NnpState* nnp_state_of(nnpe_ctx* ctx) {
  if (!ctx->nnp) ctx->nnp = new NnpState;
  return ctx->nnp;
}

void nnp_state_free(nnpe_ctx* ctx) {
  NnpState* S = ctx->nnp;
  if (!S) return;
  dev_free(S->tdesc); dev_free(S->tfdesc); dev_free(S->tpair); dev_free(S->tfpair);
  dev_free(S->M); dev_free(S->adj); dev_free(S->adjh); dev_free(S->act); dev_free(S->act_cnt);
  delete S;
  ctx->nnp = nullptr;
}

// descriptor / gradient buffers sized for the largest descriptor in use
int nnp_desc_reserve(nnpe_ctx* ctx, int D) {
  if (D > ctx->D_cap) {
    NNPE_CHECK_HIP(ctx, dev_alloc(ctx->G, size_t(ctx->natom) * D));
    NNPE_CHECK_HIP(ctx, dev_alloc(ctx->dG, size_t(ctx->natom) * D));
    NNPE_CHECK_HIP(ctx, hipMemset(ctx->dG, 0, sizeof(float) * size_t(ctx->natom) * D));
    ctx->D_cap = D;
    ctx->D_cur = 0;
  }
  if (!ctx->e_atom) NNPE_CHECK_HIP(ctx, dev_alloc(ctx->e_atom, ctx->natom));
  return NNPE_OK;
}

static NnpState& nnp_state(nnpe_ctx* ctx) { return *nnp_state_of(ctx); }

int nnp_tbl_tables(nnpe_ctx* ctx, float** ftab, float** utab, int* nbin, float* dr) {
  if (!ctx->nnp || !ctx->nnp->tbl) return 1;
  *ftab = ctx->nnp->tfpair; *utab = ctx->nnp->tpair;
  *nbin = ctx->nnp->tbl_nbin; *dr = ctx->nnp->tbl_cut / float(ctx->nnp->tbl_nbin);
  return 0;
}
static float* nnp_G(nnpe_ctx* ctx) { return ctx->G; }
static float* nnp_dG(nnpe_ctx* ctx) { return ctx->dG; }
static void nnp_set_current_D(nnpe_ctx* ctx, int D) { ctx->D_cur = D; }
static int nnp_current_D(nnpe_ctx* ctx) { return ctx->D_cur; }

// NNPOTENTIAL_TBL ctor table construction (src/pair_nnp.cpp:521-590), FP32 like the reference
extern "C" int nnpe_nnp_tbl_set(nnpe_ctx* ctx, float cutoff, int nbin, const float* coef_p, const float* coef_w,
                                const float* coef_d, const float* coef_r) {
  if (ctx) md_state_free(ctx);  // captured step graphs hold the old parameters
  if (ctx) rmdf_state_free(ctx);  // a refinement holds the old tables as its base
  if (!ctx || nbin <= 0 || !(cutoff > 0) || !coef_p || !coef_w || !coef_d || !coef_r)
    return ctx_fail(ctx, NNPE_ERR_ARG, "nnpe_nnp_tbl_set: bad argument");
  if (!ctx->have_system) return ctx_fail(ctx, NNPE_ERR_STATE, "no system");
  NnpState& S = nnp_state(ctx);
  const int npair = ctx->npair, nb1 = nbin + 1, szd = kND * 4;
  std::vector<float> tdesc(size_t(npair) * 4 * nb1 * kND, 0.0f), tfdesc(tdesc.size(), 0.0f);
  std::vector<float> tpair(size_t(npair) * 4 * nb1, 0.0f), tfpair(tpair.size(), 0.0f);
  const float dr = cutoff / float(nbin);
  std::vector<float> cp(coef_p, coef_p + size_t(npair) * 24);
  auto sq = [](float v) { return v * v; };
  for (int p = 0; p < npair; ++p) {
    for (int k = 0; k < 4; ++k) {
      float* c = &cp[size_t(p) * 24 + k * 6];
      float src2 = sq(c[1] / cutoff), src6 = src2 * src2 * src2;
      float elim = -2 * c[5] / cutoff - c[0] * src6 * (-7 + src2 * (9 * c[2] + src2 * (11 * c[3] + src2 * 13 * c[4])));
      float flim = (c[5] / cutoff + c[0] * src6 * (-6 + src2 * ((c[2] *= 8) + src2 * ((c[3] *= 10) + src2 * (c[4] *= 12))))) / cutoff;
      c[1] *= c[1];
      for (int i = 0; i < nbin; ++i) {
        float rr = float(double(dr) * (i + 0.5));
        float rsq = sq(rr);
        float s2 = c[1] / rsq, s6 = s2 * s2 * s2;
        tfpair[(size_t(p) * 4 + k) * nb1 + i] = (c[5] / rr + c[0] * s6 * (-6 + s2 * (c[2] + s2 * (c[3] + s2 * c[4])))) / rsq - flim / rr;
        tpair[(size_t(p) * 4 + k) * nb1 + i] =
            c[5] / rr + c[0] * s6 * (-1 + s2 * (c[2] * .125F + s2 * (c[3] * .1F + s2 * c[4] * 0.08333333333333333333F))) + flim * rr + elim;
      }
    }
    for (int i = 0; i < nbin; ++i) {
      float rr = float(double(dr) * (i + 0.5));
      float cf2 = rr - cutoff, rrcsq_1 = 1 + sq(cf2);
      float cff = 1 - 1 / rrcsq_1;
      float cfr = (-2 * cff) / rr;
      cf2 /= (-.5f * rr * sq(rrcsq_1));
      for (int k = 0; k < szd; ++k) {
        float rr0 = rr - coef_r[size_t(p) * szd + k];
        float drr0 = coef_d[size_t(p) * szd + k] * rr0;
        float wexp = std::exp(drr0 * rr0) * coef_w[size_t(p) * szd + k];
        // [pair][special][bin][k]: special = k / 20, descriptor = k % 20
        size_t o = ((size_t(p) * 4 + k / kND) * nb1 + i) * kND + (k % kND);
        tfdesc[o] = wexp * (cf2 + drr0 * cfr);
        tdesc[o] = wexp * cff;
      }
    }
  }
  auto up = [&](float*& d, const std::vector<float>& h) -> hipError_t {
    hipError_t e = dev_alloc(d, h.size());
    if (e != hipSuccess) return e;
    return hipMemcpy(d, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice);
  };
  NNPE_CHECK_HIP(ctx, up(S.tdesc, tdesc)); NNPE_CHECK_HIP(ctx, up(S.tfdesc, tfdesc));
  NNPE_CHECK_HIP(ctx, up(S.tpair, tpair)); NNPE_CHECK_HIP(ctx, up(S.tfpair, tfpair));
  S.tbl = true;
  S.tbl_cut = cutoff; S.tbl_cutsq = cutoff * cutoff; S.tbl_nbin = nbin;
  S.tbl_rdr = float(nbin) / cutoff;
  return nnp_desc_reserve(ctx, kND);
}

int nnp_tbl_desc_enqueue(nnpe_ctx* ctx) {
  NnpState& S = nnp_state(ctx);
  hipLaunchKernelGGL(k_tbl_desc<kTblG>, dim3(blocks_for((long long)ctx->natom * kTblG, 256)), dim3(256), 0, ctx->stream,
                     ctx->natom, ctx->cap_long, ctx->nbr_long, ctx->cnt_long, ctx->xq, ctx->typ, ctx->ntype, ctx->d_box, ctx->rbox, S.tbl_cutsq,
                     S.tbl_rdr, S.tbl_nbin + 1, S.tdesc, nnp_G(ctx));
  NNPE_CHECK_LAUNCH(ctx);
  nnp_set_current_D(ctx, kND);
  return NNPE_OK;
}

int nnp_tbl_force_enqueue(nnpe_ctx* ctx, bool ev) {
  NnpState& S = nnp_state(ctx);
  const int nblk = blocks_for((long long)ctx->natom * kTblG, 256);
  double* pe = ev_part(ctx, NNPE_TERM_NNP_PAIR, 0);
  double* pw = ev_part(ctx, NNPE_TERM_NNP_PAIR, 1);
  if (ev)
    hipLaunchKernelGGL((k_tbl_force<true, kTblG>), dim3(nblk), dim3(256), 0, ctx->stream, ctx->natom, ctx->cap_long, ctx->nbr_long,
                       ctx->cnt_long, ctx->xq, ctx->typ, ctx->ntype, ctx->d_box, ctx->rbox, S.tbl_cutsq, S.tbl_rdr,
                       S.tbl_nbin + 1, S.tfdesc, S.tpair, S.tfpair, nnp_dG(ctx), ctx->f, pe, pw);
  else
    hipLaunchKernelGGL((k_tbl_force<false, kTblG>), dim3(nblk), dim3(256), 0, ctx->stream, ctx->natom, ctx->cap_long, ctx->nbr_long,
                       ctx->cnt_long, ctx->xq, ctx->typ, ctx->ntype, ctx->d_box, ctx->rbox, S.tbl_cutsq, S.tbl_rdr,
                       S.tbl_nbin + 1, S.tfdesc, S.tpair, S.tfpair, nnp_dG(ctx), ctx->f, pe, pw);
  NNPE_CHECK_LAUNCH(ctx);
  if (ev) return ev_finish(ctx, NNPE_TERM_NNP_PAIR, nblk);
  return NNPE_OK;
}

extern "C" int nnpe_nnp_tbl_descriptor(nnpe_ctx* ctx) {
  if (!ctx) return NNPE_ERR_ARG;
  if (!nnp_state(ctx).tbl || !ctx->have_list) return ctx_fail(ctx, NNPE_ERR_STATE, "needs nnpe_nnp_tbl_set and a list");
  if (nnp_state(ctx).tbl_cut > ctx->pair_cutoff + ctx->skin) return ctx_fail(ctx, NNPE_ERR_ARG, "NNP cutoff beyond the list");
  return nnp_tbl_desc_enqueue(ctx);
}

extern "C" int nnpe_nnp_tbl_force(nnpe_ctx* ctx, int ev) {
  if (!ctx) return NNPE_ERR_ARG;
  if (!nnp_state(ctx).tbl || !ctx->have_list) return ctx_fail(ctx, NNPE_ERR_STATE, "needs nnpe_nnp_tbl_set and a list");
  if (nnp_current_D(ctx) != kND) return ctx_fail(ctx, NNPE_ERR_STATE, "dE/dG is not of the tabulated form");
  return nnp_tbl_force_enqueue(ctx, ev != 0);
}

extern "C" int nnpe_sf_set(nnpe_ctx* ctx, float rc, int nrad, const float* mu, float eta_r, int neta, const float* eta) {
  if (ctx) md_state_free(ctx);  // captured step graphs hold the old parameters
  if (!ctx || nrad != kNRad || neta != kNEta || !mu || !eta || !(rc > 0))
    return ctx_fail(ctx, NNPE_ERR_ARG, "nnpe_sf_set: the library supports 16 radial centres and 4 angular exponents");
  if (!ctx->have_system) return ctx_fail(ctx, NNPE_ERR_STATE, "no system");
  if (ctx->ntype < 1 || ctx->ntype > 4) return ctx_fail(ctx, NNPE_ERR_UNSUPPORTED, "symmetry functions support 1-4 elements");
  NnpState& S = nnp_state(ctx);
  S.P.rc = rc; S.P.rcsq = rc * rc; S.P.eta_r = eta_r; S.P.rc_inv = 1.0f / rc;
  const double l2e = 1.4426950408889634;  // 1 / ln 2
  S.P.eta_r_l2 = float(-double(eta_r) * l2e);
  for (int k = 0; k < kNRad; ++k) S.P.mu[k] = mu[k];
  for (int k = 0; k < kNEta; ++k) { S.P.eta[k] = eta[k]; S.P.eta_l2[k] = float(-double(eta[k]) * l2e); }
  // equally spaced centres (to 1e-6 of their range): the radial Gaussians by recurrence
  const double delta = (double(mu[kNRad - 1]) - double(mu[0])) / (kNRad - 1);
  double dev = 0;
  for (int k = 0; k < kNRad; ++k) dev = std::max(dev, std::fabs(double(mu[k]) - (double(mu[0]) + k * delta)));
  S.rad_recur = delta > 0 && dev <= 1e-6 * std::max(1.0, std::fabs(double(mu[kNRad - 1])));
  S.P.rad_step_l2 = float(2.0 * double(eta_r) * delta * l2e);
  for (int j = 0; j < kRadS; ++j) S.P.rad_rho[j] = float(std::exp(-double(eta_r) * delta * delta * (2 * j + 1)));
  // the force pass's radial polynomials in R and 1/R (coefficients c_k Q_k): Q_k must stay a
  // normal float, and the largest power term times its Q, exp(a z^2) with a = eta_r delta^2 and
  // z the farthest distance (in centre spacings) of any r in [0, rc] from the middle centre,
  // must stay far from the float range
  const double a_q = double(eta_r) * delta * delta;
  const double zmax = std::max(std::fabs(double(mu[kRadS])), std::fabs(double(rc) - double(mu[kRadS]))) / std::max(delta, 1e-30);
  S.rad_horner = S.rad_recur && a_q <= 1.0 && a_q * zmax * zmax <= 75.0;
  for (int k = 0; k < kNRad; ++k) S.P.rad_q[k] = float(std::exp(-a_q * double(k - kRadS) * double(k - kRadS)));
  S.P.rad_q_l2 = float(-a_q * l2e);
  S.P.rad_delta = float(delta);
  S.P.one = 1.0f;
  S.P.m_one = -1.0f;
  S.nmt = ctx->ntype * kNEta * kNMom + ctx->ntype * kNRad;
  NNPE_CHECK_HIP(ctx, dev_alloc(S.M, size_t(ctx->natom) * S.nmt));
  NNPE_CHECK_HIP(ctx, dev_alloc(S.adj, size_t(ctx->stride) * ctx->ntype * kSlice));
  NNPE_CHECK_HIP(ctx, dev_alloc(S.adjh, size_t(ctx->stride) * ctx->ntype * (kSlice / 8)));
  S.sf = true;
  S.act_cap = 0;
  int rc0 = nnp_act_reserve(ctx);
  if (rc0) return rc0;
  return nnp_desc_reserve(ctx, sf_dim(ctx->ntype));
}

// active lists sized like the short lists (called whenever their capacity changes; never inside
// a captured step)
int nnp_act_reserve(nnpe_ctx* ctx) {
  NnpState* S = ctx->nnp;
  if (!S || !S->sf || !ctx->cap_short || S->act_cap == ctx->cap_short) return NNPE_OK;
  NNPE_CHECK_HIP(ctx, dev_alloc(S->act, size_t(ctx->ntype) * ctx->cap_short * ctx->stride));
  NNPE_CHECK_HIP(ctx, dev_alloc(S->act_cnt, size_t(ctx->ntype) * ctx->stride));
  NNPE_CHECK_HIP(ctx, hipMemset(S->act_cnt, 0, sizeof(int) * ctx->ntype * ctx->stride));
  S->act_cap = ctx->cap_short;
  return NNPE_OK;
}

int sf_dim(int ntype) { return ntype * kNRad + ntype * (ntype + 1) / 2 * kNEta * 4; }

int sf_desc_enqueue(nnpe_ctx* ctx) {
  NnpState& S = nnp_state(ctx);
  const int nblk = blocks_for((long long)ctx->natom * kSfG, 256);
#define NNPE_SF_DESC1(NT, RR)                                                                                      \
  hipLaunchKernelGGL((k_sf_desc<NT, kSfG, RR>), dim3(nblk), dim3(256), 0, ctx->stream, ctx->natom, ctx->stride,     \
                     ctx->cap_short, ctx->nbr_short, ctx->cnt_short, ctx->xq, ctx->d_box, ctx->rbox, S.P, S.M, nnp_G(ctx), \
                     S.act, S.act_cnt)
#define NNPE_SF_DESC(NT) if (S.rad_recur) NNPE_SF_DESC1(NT, true); else NNPE_SF_DESC1(NT, false)
  switch (ctx->ntype) {
    case 1: NNPE_SF_DESC(1); break;
    case 2: NNPE_SF_DESC(2); break;
    case 3: NNPE_SF_DESC(3); break;
    default: NNPE_SF_DESC(4); break;
  }
#undef NNPE_SF_DESC
#undef NNPE_SF_DESC1
  NNPE_CHECK_LAUNCH(ctx);
  nnp_set_current_D(ctx, sf_dim(ctx->ntype));
  return NNPE_OK;
}

// dE/dG from a 16-bit network form: the neighbours' adjoints are gathered in FP16 (dE/dG of the
// FP32 form or given by the caller keeps FP32 slices)
static bool sf_half_gather(nnpe_ctx* ctx) { return ctx->dG_kind != 0; }

int sf_adjoint_enqueue(nnpe_ctx* ctx) {
  NnpState& S = nnp_state(ctx);
  const int nblk = blocks_for(ctx->natom, 32);  // 32 atoms per workgroup
  const bool half = sf_half_gather(ctx);
#define NNPE_SF_ADJ1(NT, H)                                                                                         \
  hipLaunchKernelGGL((k_sf_adjoint<NT, H>), dim3(nblk), dim3(32 * kNEta * NT), 0, ctx->stream, ctx->natom, ctx->stride, \
                     S.M, nnp_dG(ctx), S.adj, S.adjh)
#define NNPE_SF_ADJ(NT) if (half) NNPE_SF_ADJ1(NT, true); else NNPE_SF_ADJ1(NT, false)
  switch (ctx->ntype) {
    case 1: NNPE_SF_ADJ(1); break;
    case 2: NNPE_SF_ADJ(2); break;
    case 3: NNPE_SF_ADJ(3); break;
    default: NNPE_SF_ADJ(4); break;
  }
#undef NNPE_SF_ADJ
#undef NNPE_SF_ADJ1
  NNPE_CHECK_LAUNCH(ctx);
  return NNPE_OK;
}

// the pair-force kernel alone (uses the adjoints of the last sf_adjoint_enqueue)
int sf_force_only_enqueue(nnpe_ctx* ctx, bool ev, bool kick, int adv) {
  NnpState& S = nnp_state(ctx);
  const int nblk = blocks_for((long long)ctx->natom * kSfG, 256);
  double* pe = ev_part(ctx, NNPE_TERM_SF, 0);
  double* pw = ev_part(ctx, NNPE_TERM_SF, 1);
  if (kick && !ctx->fp64) return ctx_fail(ctx, NNPE_ERR_STATE, "fused half-kick needs the FP64 integrator");
  if (adv && (!kick || !ctx->xq_alt)) return ctx_fail(ctx, NNPE_ERR_STATE, "fused position update needs the fused half-kick");
  const VvArgs V{ctx->dt, ctx->rm, ctx->vd, ctx->ad, ctx->xd, ctx->xq_alt, adv};
  const bool half = sf_half_gather(ctx);
#define NNPE_SF_FORCE3(NT, EV, RR, VV, H)                                                                                \
  hipLaunchKernelGGL((k_sf_force<NT, EV, kSfG, RR, VV, H>), dim3(nblk), dim3(256), 0, ctx->stream, ctx->natom, ctx->stride, \
                     S.act_cap, S.act, S.act_cnt, ctx->xq, ctx->typ, ctx->d_box, ctx->rbox, S.P, S.adj, S.adjh, ctx->f, pe, pw, V)
#define NNPE_SF_FORCE2(NT, EV, RR, VV) if (half) NNPE_SF_FORCE3(NT, EV, RR, VV, true); else NNPE_SF_FORCE3(NT, EV, RR, VV, false)
#define NNPE_SF_FORCE1(NT, EV, RR) if (kick) NNPE_SF_FORCE2(NT, EV, RR, true); else NNPE_SF_FORCE2(NT, EV, RR, false)
#define NNPE_SF_FORCE(NT, EV) if (S.rad_horner) NNPE_SF_FORCE1(NT, EV, true); else NNPE_SF_FORCE1(NT, EV, false)
  switch (ctx->ntype) {
    case 1: if (ev) NNPE_SF_FORCE(1, true); else NNPE_SF_FORCE(1, false); break;
    case 2: if (ev) NNPE_SF_FORCE(2, true); else NNPE_SF_FORCE(2, false); break;
    case 3: if (ev) NNPE_SF_FORCE(3, true); else NNPE_SF_FORCE(3, false); break;
    default: if (ev) NNPE_SF_FORCE(4, true); else NNPE_SF_FORCE(4, false); break;
  }
#undef NNPE_SF_FORCE
#undef NNPE_SF_FORCE1
#undef NNPE_SF_FORCE2
#undef NNPE_SF_FORCE3
  NNPE_CHECK_LAUNCH(ctx);
  if (ev) return ev_finish(ctx, NNPE_TERM_SF, nblk);
  return NNPE_OK;
}

int sf_force_enqueue(nnpe_ctx* ctx, bool ev) {
  int rc = sf_adjoint_enqueue(ctx);
  if (rc) return rc;
  return sf_force_only_enqueue(ctx, ev, false);
}

extern "C" int nnpe_sf_descriptor(nnpe_ctx* ctx) {
  if (!ctx) return NNPE_ERR_ARG;
  if (!nnp_state(ctx).sf || !ctx->have_list) return ctx_fail(ctx, NNPE_ERR_STATE, "needs nnpe_sf_set and a list");
  if (!ctx->cap_short) return ctx_fail(ctx, NNPE_ERR_STATE, "symmetry functions need the short lists (nnpe_neigh_init short_cutoff)");
  if (nnp_state(ctx).P.rc > ctx->short_cutoff + ctx->skin) return ctx_fail(ctx, NNPE_ERR_ARG, "symmetry-function cutoff beyond the short list");
  return sf_desc_enqueue(ctx);
}

extern "C" int nnpe_sf_force(nnpe_ctx* ctx, int ev) {
  if (!ctx) return NNPE_ERR_ARG;
  if (!nnp_state(ctx).sf || !ctx->have_list) return ctx_fail(ctx, NNPE_ERR_STATE, "needs nnpe_sf_set and a list");
  if (!ctx->cap_short) return ctx_fail(ctx, NNPE_ERR_STATE, "symmetry functions need the short lists (nnpe_neigh_init short_cutoff)");
  if (nnp_current_D(ctx) != sf_dim(ctx->ntype)) return ctx_fail(ctx, NNPE_ERR_STATE, "dE/dG is not of the symmetry-function form");
  return sf_force_enqueue(ctx, ev != 0);
}
