// sqd.hip -- direct k-space S(q) (SQD, src/grsq.cpp:362-606): for every reciprocal vector G of
// the periodic box in a half space with |G| <= q_max, the per-element structure factors
// rho_a(G) = sum_{i in a} exp(i G.x_i), and per |G| bin the average of
// Re(rho_a rho_b*) = S_a S_b + C_a C_b, normalised as natom / (N_a N_b) * avg - [a == b] natom / N_a.
//
// Phases are exact.  Every atom's fractional coordinates s = x / L (FP64 division) become 32-bit
// fixed-point integers S = round(s * 2^32) mod 2^32, so the phase of G = 2 pi (h, k, l) / L is
//   G.x / (2 pi) mod 1 = (h Sx + k Sy + l Sz) mod 2^32  (in units of 2^-32 revolutions),
// an integer sum that wraps exactly; no rounding of a large argument such as 1000 rad enters.
//
// Work split: a lane owns a run of kRun consecutive G vectors along one axis (G, G + e, ...,
// e a unit reciprocal step) and walks every atom: exp(i phi) of the run's first vector and the
// step exp(i theta), theta = 2 pi S_axis / 2^32, come from the hardware sin/cos on the exact
// phase (input in revolutions); the rest of the run follows by complex rotation (FMA), so a
// (G, atom) pair costs one rotation and two adds.  Atom records are read by uniform (scalar)
// loads; the element of the atom is uniform across the wave, so its accumulator set is chosen
// by a uniform branch.  Per-atom terms are summed in FP32 over chunks of kChunk atoms and the
// chunk sums are added in FP64.  Each G's pair products go to a per-vector FP64 buffer and a
// second kernel reduces each |G| bin over its vectors in a fixed order: results are
// deterministic run to run.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "internal.h"

namespace {

constexpr int kChunk = 64;  // atoms per FP32 partial sum
constexpr int kMaxType = 4;

// fixed-point fractional coordinates (Sx, Sy, Sz, element) of every row; rows n .. npad-1 pad
// the last chunk with element kMaxType (matched by no accumulator set)
__global__ void __launch_bounds__(256) k_sqd_prep(int n, int npad, const float4* __restrict__ xq, const int* __restrict__ typ,
                                                  double bl, uint4* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npad) return;
  if (i >= n) {
    out[i] = make_uint4(0u, 0u, 0u, (uint32_t)kMaxType);
    return;
  }
  const float4 p = xq[i];
  auto fix = [bl](float x) -> uint32_t {
    const double t = __builtin_rint(((double)x / bl) * 4294967296.0);  // |t| < 2^40: exact in int64
    return (uint32_t)(unsigned long long)(long long)t;                  // mod 2^32
  };
  out[i] = make_uint4(fix(p.x), fix(p.y), fix(p.z), (uint32_t)typ[i]);
}

// (m * S) mod 2^32 (unsigned arithmetic wraps; a negative m is its two's complement)
__device__ __forceinline__ uint32_t phase_mul(int m, uint32_t S) { return (uint32_t)m * S; }

// task: x = packed (h + 512) | (k + 512) << 10 | (l + 512) << 20, y = axis | len << 2,
// z = index of the run's first vector
template <int NT, int RUN>
__global__ void __launch_bounds__(256) k_sqd(int ntask, const int4* __restrict__ task, int natom,
                                             const uint4* __restrict__ atoms, double* __restrict__ val) {
#pragma clang fp contract(fast)
  constexpr int NP = NT * (NT + 1) / 2;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = t < ntask;
  const int4 tk = task[live ? t : 0];
  const int h = (tk.x & 1023) - 512, k = ((tk.x >> 10) & 1023) - 512, l = ((tk.x >> 20) & 1023) - 512;
  const int axis = tk.y & 3, len = live ? (tk.y >> 2) : 0;
  double Cd[NT][RUN], Sd[NT][RUN];
#pragma unroll
  for (int e = 0; e < NT; ++e)
#pragma unroll
    for (int j = 0; j < RUN; ++j) Cd[e][j] = Sd[e][j] = 0.0;
  constexpr float kRev = 2.3283064365386963e-10f;  // 2^-32
  for (int i0 = 0; i0 < natom; i0 += kChunk) {  // natom: a multiple of kChunk (padded)
    float Cf[NT][RUN], Sf[NT][RUN];
#pragma unroll
    for (int e = 0; e < NT; ++e)
#pragma unroll
      for (int j = 0; j < RUN; ++j) Cf[e][j] = Sf[e][j] = 0.0f;
    for (int i = i0; i < i0 + kChunk; i += 4) {
      uint4 a4[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) a4[u] = atoms[i + u];  // uniform addresses: scalar loads, issued together
#pragma unroll
      for (int u = 0; u < 4; ++u) {
      const uint4 a = a4[u];
      const uint32_t ph = phase_mul(h, a.x) + phase_mul(k, a.y) + phase_mul(l, a.z);
      const uint32_t st = axis == 0 ? a.x : (axis == 1 ? a.y : a.z);
      const float p0 = (float)(int)ph * kRev, ps = (float)(int)st * kRev;  // revolutions in [-1/2, 1/2)
      float c = __builtin_amdgcn_cosf(p0), s = __builtin_amdgcn_sinf(p0);
      const float cw = __builtin_amdgcn_cosf(ps), sw = __builtin_amdgcn_sinf(ps);
      const int ty = (int)a.w;
#pragma unroll
      for (int e = 0; e < NT; ++e) {
        if (ty == e) {
#pragma unroll
          for (int j = 0; j < RUN; ++j) {
            Cf[e][j] += c;
            Sf[e][j] += s;
            const float cn = c * cw - s * sw;
            s = s * cw + c * sw;
            c = cn;
          }
        }
      }
      }
    }
#pragma unroll
    for (int e = 0; e < NT; ++e)
#pragma unroll
      for (int j = 0; j < RUN; ++j) {
        Cd[e][j] += (double)Cf[e][j];
        Sd[e][j] += (double)Sf[e][j];
      }
  }
  if (!live) return;
  const int g0 = tk.z;
#pragma unroll
  for (int j = 0; j < RUN; ++j) {
    if (j >= len) break;
    double* o = val + (size_t)(g0 + j) * NP;
    int p = 0;
#pragma unroll
    for (int a = 0; a < NT; ++a)
#pragma unroll
      for (int b = a; b < NT; ++b, ++p) o[p] = Sd[a][j] * Sd[b][j] + Cd[a][j] * Cd[b][j];
  }
}

// per bin: the average over its vectors (fixed lane assignment and butterfly: deterministic)
__global__ void __launch_bounds__(64) k_sqd_reduce(int npair, const int* __restrict__ bstart, const int* __restrict__ perm,
                                                   const double* __restrict__ val, const double* __restrict__ factor,
                                                   const double* __restrict__ base, double* __restrict__ sq) {
  const int b = blockIdx.x, lane = threadIdx.x;
  const int m0 = bstart[b], m1 = bstart[b + 1];
  for (int p = 0; p < npair; ++p) {
    double acc = 0.0;
    for (int m = m0 + lane; m < m1; m += 64) acc += val[(size_t)perm[m] * npair + p];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if (lane == 0) sq[(size_t)b * npair + p] = m1 > m0 ? acc / (m1 - m0) * factor[p] - base[p] : 0.0;
  }
}

}  // namespace

static int sqd_padded(int natom) { return (natom + kChunk - 1) / kChunk * kChunk; }

struct SqdState {
  int nq = 0, run = 8, ntask = 0;
  long long ngv = 0;
  std::vector<double> qq;          // mean |G| of each bin
  std::vector<long long> count;    // vectors per bin
  int4* d_task = nullptr;
  int* d_perm = nullptr;           // vector indices grouped by bin
  int* d_bstart = nullptr;         // [nq + 1]
  double* d_val = nullptr;         // [ngv][npair]
  double* d_fb = nullptr;          // factor [npair], base [npair]
  double* d_sq = nullptr;          // [nq][npair]
  uint4* d_atoms = nullptr;        // [natom]
};

void sqd_state_free(nnpe_ctx* ctx) {
  SqdState* S = ctx->sqd;
  if (!S) return;
  dev_free(S->d_task); dev_free(S->d_perm); dev_free(S->d_bstart); dev_free(S->d_val);
  dev_free(S->d_fb); dev_free(S->d_sq); dev_free(S->d_atoms);
  delete S;
  ctx->sqd = nullptr;
}

// SQD::SQD (src/grsq.cpp:362-491): the half space of reciprocal vectors (i, j, k >= 1), then
// (i, j >= 1, 0), then (i >= 1, 0, 0), |i|, |j|, |k| <= kcutoff, i^2 + j^2 + k^2 <= cuti2, with
// g = 2 pi / L, kcutoff = int(q_max / g), cuti2 = int((q_max / g)^2) and |G| = sqrt(n^2) g in the
// reference's float arithmetic; bin b >= 1 holds the vectors with b dq <= |G| < (b + 1) dq
// (dq = q_max / nq, the last bin everything from (nq - 1) dq up), bin 0 the rest.
extern "C" int nnpe_sqd_init(nnpe_ctx* ctx, int nq, float q_max, long long* n_gvec) {
  if (!ctx || nq <= 0 || !(q_max > 0)) return ctx_fail(ctx, NNPE_ERR_ARG, "nnpe_sqd_init: bad argument");
  if (!ctx->have_system) return ctx_fail(ctx, NNPE_ERR_STATE, "nnpe_sqd_init: no system");
  if (ctx->nbox != 1) return ctx_fail(ctx, NNPE_ERR_UNSUPPORTED, "nnpe_sqd_init: one periodic box only");
  if (ctx->ntype > kMaxType) return ctx_fail(ctx, NNPE_ERR_UNSUPPORTED, "nnpe_sqd_init: at most 4 elements");
  sqd_state_free(ctx);
  const float pi_f = 3.14159265358979323846f, two_pi_f = pi_f * 2;
  const float g = two_pi_f / ctx->bl;
  const float qg = q_max / g;
  const int kc = (int)qg;
  const int cuti2 = (int)(qg * qg);
  const float dq = q_max / nq;
  if (kc > 500) return ctx_fail(ctx, NNPE_ERR_UNSUPPORTED, "nnpe_sqd_init: q_max L / (2 pi) above 500");
  SqdState* S = ctx->sqd = new SqdState;
  S->nq = nq;
  S->run = ctx->ntype <= 2 ? 8 : 4;
  const int run = S->run;
  auto bin_of = [&](int n2) {
    const float gs = std::sqrt((float)n2) * g;
    int b = std::min(nq - 1, std::max(0, (int)(gs / dq)));
    while (b + 1 <= nq - 1 && gs >= (float)(b + 1) * dq) ++b;
    while (b >= 1 && !(gs >= (float)b * dq)) --b;
    return std::make_pair(b, gs);
  };
  std::vector<int4> tasks;
  std::vector<int> gbin;
  S->qq.assign(nq, 0.0);
  S->count.assign(nq, 0);
  auto pack = [](int h, int k, int l) { return (h + 512) | ((k + 512) << 10) | ((l + 512) << 20); };
  // a column of vectors start + t e (t = 0 .. n-1) along axis, cut into runs
  auto column = [&](int h, int k, int l, int axis, int n) {
    for (int t0 = 0; t0 < n; t0 += run) {
      const int len = std::min(run, n - t0);
      const int hh = h + (axis == 0 ? t0 : 0), kk = k + (axis == 1 ? t0 : 0), ll = l + (axis == 2 ? t0 : 0);
      tasks.push_back(make_int4(pack(hh, kk, ll), axis | (len << 2), (int)gbin.size(), 0));
      for (int t = 0; t < len; ++t) {
        const int a = hh + (axis == 0 ? t : 0), b = kk + (axis == 1 ? t : 0), c = ll + (axis == 2 ? t : 0);
        const auto bg = bin_of(a * a + b * b + c * c);
        gbin.push_back(bg.first);
        S->qq[bg.first] += bg.second;
        ++S->count[bg.first];
      }
    }
  };
  for (int i = -kc; i <= kc; ++i)
    for (int j = -kc; j <= kc; ++j) {
      int n = 0;
      while (n + 1 <= kc && i * i + j * j + (n + 1) * (n + 1) <= cuti2) ++n;
      if (n > 0) column(i, j, 1, 2, n);
    }
  for (int i = -kc; i <= kc; ++i) {
    int n = 0;
    while (n + 1 <= kc && i * i + (n + 1) * (n + 1) <= cuti2) ++n;
    if (n > 0) column(i, 1, 0, 1, n);
  }
  {
    int n = 0;
    while (n + 1 <= kc && (n + 1) * (n + 1) <= cuti2) ++n;
    if (n > 0) column(1, 0, 0, 0, n);
  }
  S->ngv = (long long)gbin.size();
  S->ntask = (int)tasks.size();
  if (S->ngv >= (1LL << 31)) return ctx_fail(ctx, NNPE_ERR_UNSUPPORTED, "nnpe_sqd_init: too many vectors");
  for (int b = 0; b < nq; ++b)
    if (S->count[b]) S->qq[b] /= double(S->count[b]);
  // vectors grouped by bin, in enumeration order within a bin
  std::vector<int> bstart(nq + 1, 0), fill(nq, 0), perm(std::max<long long>(1, S->ngv));
  for (int b = 0; b < nq; ++b) bstart[b + 1] = bstart[b] + (int)S->count[b];
  for (long long v = 0; v < S->ngv; ++v) {
    const int b = gbin[v];
    perm[bstart[b] + fill[b]++] = (int)v;
  }
  // normalisation (src/grsq.cpp:478-489) in FP64
  const int npair = ctx->npair;
  std::vector<double> fb(2 * npair);
  int ip = 0;
  for (int a = 0; a < ctx->ntype; ++a)
    for (int b = a; b < ctx->ntype; ++b, ++ip) {
      fb[ip] = double(ctx->natom) / (double(ctx->h_typecount[a]) * ctx->h_typecount[b]);
      fb[npair + ip] = a == b ? double(ctx->natom) / ctx->h_typecount[b] : 0.0;
    }
  NNPE_CHECK_HIP(ctx, dev_alloc(S->d_task, std::max(1, S->ntask)));
  NNPE_CHECK_HIP(ctx, dev_alloc(S->d_perm, perm.size()));
  NNPE_CHECK_HIP(ctx, dev_alloc(S->d_bstart, nq + 1));
  NNPE_CHECK_HIP(ctx, dev_alloc(S->d_val, size_t(std::max<long long>(1, S->ngv)) * npair));
  NNPE_CHECK_HIP(ctx, dev_alloc(S->d_fb, fb.size()));
  NNPE_CHECK_HIP(ctx, dev_alloc(S->d_sq, size_t(nq) * npair));
  NNPE_CHECK_HIP(ctx, dev_alloc(S->d_atoms, sqd_padded(ctx->natom)));
  if (S->ntask) NNPE_CHECK_HIP(ctx, hipMemcpy(S->d_task, tasks.data(), tasks.size() * sizeof(int4), hipMemcpyHostToDevice));
  if (S->ngv) NNPE_CHECK_HIP(ctx, hipMemcpy(S->d_perm, perm.data(), S->ngv * sizeof(int), hipMemcpyHostToDevice));
  NNPE_CHECK_HIP(ctx, hipMemcpy(S->d_bstart, bstart.data(), bstart.size() * sizeof(int), hipMemcpyHostToDevice));
  NNPE_CHECK_HIP(ctx, hipMemcpy(S->d_fb, fb.data(), fb.size() * sizeof(double), hipMemcpyHostToDevice));
  if (n_gvec) *n_gvec = S->ngv;
  return NNPE_OK;
}

int sqd_enqueue(nnpe_ctx* ctx) {
  SqdState* S = ctx->sqd;
  hipStream_t st = ctx->stream;
  const int npad = sqd_padded(ctx->natom);
  hipLaunchKernelGGL(k_sqd_prep, dim3(blocks_for(npad, 256)), dim3(256), 0, st, ctx->natom, npad, ctx->xq, ctx->typ,
                     (double)ctx->bl, S->d_atoms);
  NNPE_CHECK_LAUNCH(ctx);
  if (S->ntask > 0) {
    const dim3 grid(blocks_for(S->ntask, 256)), blk(256);
    switch (ctx->ntype) {
      case 1: hipLaunchKernelGGL((k_sqd<1, 8>), grid, blk, 0, st, S->ntask, S->d_task, npad, S->d_atoms, S->d_val); break;
      case 2: hipLaunchKernelGGL((k_sqd<2, 8>), grid, blk, 0, st, S->ntask, S->d_task, npad, S->d_atoms, S->d_val); break;
      case 3: hipLaunchKernelGGL((k_sqd<3, 4>), grid, blk, 0, st, S->ntask, S->d_task, npad, S->d_atoms, S->d_val); break;
      default: hipLaunchKernelGGL((k_sqd<4, 4>), grid, blk, 0, st, S->ntask, S->d_task, npad, S->d_atoms, S->d_val); break;
    }
    NNPE_CHECK_LAUNCH(ctx);
  }
  hipLaunchKernelGGL(k_sqd_reduce, dim3(S->nq), dim3(64), 0, st, ctx->npair, S->d_bstart, S->d_perm, S->d_val, S->d_fb,
                     S->d_fb + ctx->npair, S->d_sq);
  NNPE_CHECK_LAUNCH(ctx);
  return NNPE_OK;
}

extern "C" int nnpe_sqd_compute(nnpe_ctx* ctx) {
  if (!ctx) return NNPE_ERR_ARG;
  if (!ctx->sqd) return ctx_fail(ctx, NNPE_ERR_STATE, "nnpe_sqd_compute before nnpe_sqd_init");
  if (ctx->nbox != 1) return ctx_fail(ctx, NNPE_ERR_UNSUPPORTED, "nnpe_sqd_compute: one periodic box only");
  return sqd_enqueue(ctx);
}

extern "C" int nnpe_sqd_get(nnpe_ctx* ctx, double* qq, double* sq, long long* count) {
  if (!ctx) return NNPE_ERR_ARG;
  SqdState* S = ctx->sqd;
  if (!S) return ctx_fail(ctx, NNPE_ERR_STATE, "nnpe_sqd_get before nnpe_sqd_init");
  if (qq) std::memcpy(qq, S->qq.data(), sizeof(double) * S->nq);
  if (count) std::memcpy(count, S->count.data(), sizeof(long long) * S->nq);
  if (sq) {
    NNPE_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    NNPE_CHECK_HIP(ctx, hipMemcpy(sq, S->d_sq, sizeof(double) * S->nq * ctx->npair, hipMemcpyDeviceToHost));
  }
  return NNPE_OK;
}
