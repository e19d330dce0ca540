// mlp.hip -- per-element MLP energy and its input gradient on MFMA (replaces the TensorFlow
// call TFCALL::evaluate, src/pair_nnp.cpp:92-97, behind the same contract: descriptors in,
// per-atom energy and dE/dG out).
//
// Network of element e:  E_i = s_out * (w_out . a_L + b_out),  a_l = tanh(a_{l-1} W_l + b_l),
// a_0 = G_i.  Backward (hand-written): d_L = s_out * w_out * (1 - a_L^2),
// d_{l-1} = (d_l W_l^T) * (1 - a_{l-1}^2),  dE/dG = d_1 W_1^T.
//
// Two kernels of the same shape, picked once in nnpe_mlp_set: k_mlp for any widths and every
// form, and k_mlp_w, the 16-bit forms' kernel for hidden layers all 128 wide and descriptors at
// most 128 wide (the production network), trimmed for VALU work and with its weight loads
// issued one product ahead.
// Kernel shape: one workgroup of kMlpWaves waves owns a tile of 32 atoms of one element and runs
// the whole forward and backward for them; within a layer wave w computes the 32-column output
// blocks w, w + kMlpWaves, ... and a barrier separates the layers.  Activations live in the
// workgroup's LDS slab [32 x width] per layer (padded rows, no bank-conflicting strides); every
// GEMM is
// 32 x N x K on v_mfma_f32_32x32x16_{bf16,f16} (K = 16 per instruction) or
// v_mfma_f32_32x32x2f32 for the FP32 form, with FP32 accumulation.  Weights are read as MFMA B
// fragments straight from global memory (L2-resident: 86 KB per element in BF16), in two
// layouts: [out][in] for the forward products, [in][out] for the backward ones, so each lane's
// B fragment is one contiguous 16-byte load.
// Precision forms: FP32 (fp32 storage, fp32 MFMA), BF16 and FP16 (16-bit storage of
// activations and weights, fp32 accumulation, bias/tanh/epilogues in fp32).  The descriptor
// standardisation x = (G - shift) * scale runs in fp32 while the rows are staged, so 16-bit
// rounding applies to O(1) standardised inputs and never to raw descriptors whose spread is
// small against their mean (folding the standardisation into layer-1 weights would cancel
// catastrophically there).  Per-layer power-of-two scales keep FP16 operands in range:
// weight scale s_w[l], gradient scale s_d[l]; each accumulator is unscaled in fp32 before use.
// The BF16 form keeps bf16 operands on every MFMA but carries the weights and the backward
// gradients as hi + lo bf16 pairs (x = hi + lo, lo = bf16(x - hi), a 16-bit significand): forward
// products take two MFMAs (A.W_hi + A.W_lo), backward products three (d_hi.W_hi + d_lo.W_hi +
// d_hi.W_lo).  Activations and the standardised input stay single bf16.  Weight and gradient
// rounding dominate the error of a plain bf16 network; the pairs remove both for 2-3x the MFMA
// work of a layer, which the MFMA rate (16x the FP32 MFMA rate) absorbs.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "internal.h"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

namespace {

constexpr int kMaxLayers = 6;   // weight layers including the output layer
constexpr int kMaxWidth = 128;  // widest hidden layer: one 32-column block per wave
constexpr int kMaxType = 4;

struct MlpDev {
  int nh;                       // hidden layers (weight layers - 1)
  int D, Dp;                    // descriptor width and its 32-multiple padding
  int H[kMaxLayers];            // hidden widths
  int ld[kMaxLayers + 1];       // LDS row strides (elements): [0] = G slab, [1 + l] = layer l
  int slab_off[kMaxLayers + 1]; // element offsets of the slabs in LDS
  int tile_off[kMaxType + 1];   // first tile of each element
  int count[kMaxType];
  int type_off[kMaxType];
  const void* Wt[kMaxType][kMaxLayers];  // [out][in] storage type
  const void* W[kMaxType][kMaxLayers];   // [in][out] storage type
  const void* Wt_lo[kMaxType][kMaxLayers];  // BF16 form: low parts of the weight pairs
  const void* W_lo[kMaxType][kMaxLayers];
  int lo_off[kMaxLayers + 1];    // BF16 form: LDS slabs of the gradients' low parts (layers 1..nh)
  const float* b[kMaxType][kMaxLayers];
  const float* wout[kMaxType];
  const float* shift[kMaxType];  // input standardisation [D] per element
  const float* scale[kMaxType];
  float bout[kMaxType];
  float out_scale;
  float s_w[kMaxLayers], s_d[kMaxLayers];
  // their reciprocals (powers of two, so exact): the kernels multiply instead of dividing
  float r_w[kMaxLayers], r_dw[kMaxLayers];  // 1 / s_w[l], 1 / (s_d[l] s_w[l])
};

template <int KIND> struct Prec;
template <> struct Prec<0> {  // FP32
  typedef float T;
  typedef float F;  // one lane's A or B fragment of one MFMA
  static constexpr int KS = 2;
  static constexpr int NB = 16;  // MFMAs per batch of prefetched B fragments
  __device__ static F ld(const T* p) { return *p; }
  __device__ static f32x16 mma(F a, F b, f32x16 c) { return __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, 0, 0, 0); }
  __device__ static T cvt(float v) { return v; }
  __device__ static float back(T v) { return v; }
};
template <> struct Prec<1> {  // BF16
  typedef __bf16 T;
  typedef bf16x8 F;
  static constexpr int KS = 16;
  static constexpr int NB = 4;  // B fragments per batch (hi and lo: eight fragment loads in flight)
  __device__ static F ld(const T* p) { return *reinterpret_cast<const bf16x8*>(p); }
  __device__ static f32x16 mma(F a, F b, f32x16 c) { return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0); }
  __device__ static T cvt(float v) { return (__bf16)v; }
  __device__ static float back(T v) { return (float)v; }
};
template <> struct Prec<3> {  // BF16, one term: the weights rounded once, no remainder parts
  typedef __bf16 T;
  typedef bf16x8 F;
  static constexpr int KS = 16;
  static constexpr int NB = 8;
  __device__ static F ld(const T* p) { return *reinterpret_cast<const bf16x8*>(p); }
  __device__ static f32x16 mma(F a, F b, f32x16 c) { return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0); }
  __device__ static T cvt(float v) { return (__bf16)v; }
  __device__ static float back(T v) { return (float)v; }
};
template <> struct Prec<2> {  // FP16
  typedef _Float16 T;
  typedef f16x8 F;
  static constexpr int KS = 16;
  static constexpr int NB = 8;
  __device__ static F ld(const T* p) { return *reinterpret_cast<const f16x8*>(p); }
  __device__ static f32x16 mma(F a, F b, f32x16 c) { return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0); }
  __device__ static T cvt(float v) { return (_Float16)v; }
  __device__ static float back(T v) { return (float)v; }
};

// C/D fragment of a 32x32 MFMA: lane l, element j -> row 8*(j/4) + 4*(l/32) + j%4, col l%32
__device__ __forceinline__ int crow(int lane, int j) { return 8 * (j >> 2) + 4 * (lane >> 5) + (j & 3); }

// four consecutive activations in LDS (8-byte aligned for 16-bit types, 16-byte for fp32)
template <class T>
__device__ __forceinline__ void store4(T* p, const T* v) {
  if constexpr (sizeof(T) == 2) {
    ushort4 u;
    __builtin_memcpy(&u, v, 8);
    *reinterpret_cast<ushort4*>(p) = u;
  } else {
    *reinterpret_cast<float4*>(p) = make_float4(v[0], v[1], v[2], v[3]);
  }
}
template <class T>
__device__ __forceinline__ void load4(const T* p, T* v) {
  if constexpr (sizeof(T) == 2) {
    const ushort4 u = *reinterpret_cast<const ushort4*>(p);
    __builtin_memcpy(v, &u, 8);
  } else {
    const float4 f = *reinterpret_cast<const float4*>(p);
    v[0] = f.x; v[1] = f.y; v[2] = f.z; v[3] = f.w;
  }
}

// Weight matrices W [n][K] (n a multiple of 32) are stored in MFMA fragment order: fragment
// (cb, t) -- rows cb*32 .. cb*32+31, k = t*KS .. t*KS+KS-1 -- is 64 lanes x KS/2 consecutive
// elements, lane l holding row cb*32 + l%32 and k = t*KS + (KS/2)*(l/32) + h.  One wave's
// fragment load is then one contiguous 1 KiB (16-bit) or 256 B (fp32) run: 8 (or 2) whole
// cache lines instead of 32 partial ones, which is what bounds a layer when every tile
// re-streams the weights through its CU's L1.
template <int KS, class T>
__device__ __forceinline__ const T* frag_ptr(const T* W, int K, int cb, int t, int lane) {
  return W + ((size_t)cb * (K / KS) + t) * (64 * (KS / 2)) + lane * (KS / 2);
}

// Transposed output block: out^T[32 features of block cb x 32 rows] = W[cb rows of bm] * A^T,
// the weights `bm` (fragment order, from global) as the MFMA A operand and the activations A
// ([32 rows][K] in LDS, row stride lda) as the B operand.  Both fragments: lane l holds
// row/column l%32 and the KS/2 consecutive k starting at ks*KS + (KS/2)*(l/32).  Lane l of the
// result holds features crow(l, j) of activation row l%32, four consecutive features per group
// of j, so the epilogues store 4 values at once.  The weight fragments of NB consecutive MFMAs
// are loaded into registers before the MFMA chain, so a chain waits for one round trip to L2
// instead of one per MFMA.
template <int KIND>
__device__ __forceinline__ f32x16 gemm_block(const typename Prec<KIND>::T* A, int lda, const typename Prec<KIND>::T* bm,
                                             int K, int cb, int lane) {
  typedef Prec<KIND> P;
  typedef typename P::F F;
  constexpr int KS = P::KS, HALF = KS / 2, NB = P::NB;
  f32x16 acc;
#pragma unroll
  for (int j = 0; j < 16; ++j) acc[j] = 0.0f;
  const typename P::T* arow = A + (lane & 31) * lda + HALF * (lane >> 5);
  for (int kb = 0; kb < K; kb += NB * KS) {
    F b[NB], a[NB];
#pragma unroll
    for (int s = 0; s < NB; ++s)
      if (kb + s * KS < K) b[s] = P::ld(frag_ptr<KS>(bm, K, cb, kb / KS + s, lane));
#pragma unroll
    for (int s = 0; s < NB; ++s)
      if (kb + s * KS < K) a[s] = P::ld(arow + kb + s * KS);  // LDS reads issued together
#pragma unroll
    for (int s = 0; s < NB; ++s)
      if (kb + s * KS < K) acc = P::mma(b[s], a[s], acc);
  }
  return acc;
}

// BF16 pair products: sum of A.B over the hi/lo parts given (nullptr = part absent), B
// fragments prefetched as in gemm_block
__device__ __forceinline__ f32x16 gemm_block_pairs(const __bf16* Ahi, const __bf16* Alo, int lda, const __bf16* Bhi,
                                                   const __bf16* Blo, int K, int cb, int lane) {
  typedef Prec<1> P;
  typedef P::F F;
  constexpr int KS = P::KS, HALF = KS / 2, NB = P::NB;
  f32x16 acc;
#pragma unroll
  for (int j = 0; j < 16; ++j) acc[j] = 0.0f;
  const size_t ao = (size_t)(lane & 31) * lda + HALF * (lane >> 5);
  for (int kb = 0; kb < K; kb += NB * KS) {
    F bh[NB], bl[NB], ah[NB], al[NB];
#pragma unroll
    for (int s = 0; s < NB; ++s)
      if (kb + s * KS < K) {
        bh[s] = P::ld(frag_ptr<KS>(Bhi, K, cb, kb / KS + s, lane));
        if (Blo) bl[s] = P::ld(frag_ptr<KS>(Blo, K, cb, kb / KS + s, lane));
      }
    // LDS reads of the A fragments issued together, half a batch at a time (register budget)
#pragma unroll
    for (int h = 0; h < NB; h += NB / 2) {
#pragma unroll
      for (int s = h; s < h + NB / 2; ++s)
        if (kb + s * KS < K) {
          ah[s] = P::ld(Ahi + ao + kb + s * KS);
          if (Alo) al[s] = P::ld(Alo + ao + kb + s * KS);
        }
#pragma unroll
      for (int s = h; s < h + NB / 2; ++s)
        if (kb + s * KS < K) {
          acc = P::mma(bh[s], ah[s], acc);
          if (Blo) acc = P::mma(bl[s], ah[s], acc);
          if (Alo) acc = P::mma(bh[s], al[s], acc);
        }
    }
  }
  return acc;
}

constexpr int kMlpWaves = 4;

// tanh and its derivative from one exponential: with e = exp(-2|z|) (hardware exp2) and
// r = 1/(1 + e) (hardware reciprocal), tanh z = sign(z) (1 - e) r and 1 - tanh^2 z = 4 e r^2.
// Both keep their relative precision (a few ulp) for any |z|: the derivative of a saturated unit
// is never formed as 1 - a^2 from an activation that has been rounded to 8 or 11 bits.
__device__ __forceinline__ void tanh_d(float z, float& a, float& t) {
  const float e = __builtin_amdgcn_exp2f(-2.8853900817779268f * fabsf(z));  // exp(-2|z|) = 2^(-2|z| / ln 2)
  const float r = __builtin_amdgcn_rcpf(1.0f + e);
  a = copysignf((1.0f - e) * r, z);
  t = 4.0f * e * r * r;
}

// One workgroup of kMlpWaves waves per 32-atom tile of one element, NH hidden layers of width at
// most 32 * kMlpWaves, so wave w owns the output column block w of every layer.  The forward
// epilogue of hidden layer l keeps 1 - a_l^2 in FP32 registers (the backward epilogue of the
// same column block needs exactly those values); the last hidden layer's epilogue forms the
// energy partial sums and d_L = s_out w_out (1 - a_L^2) directly.
template <int KIND, int NH>
// three waves per SIMD: the LDS of three tiles fits a CU, so registers are capped to match
__global__ void __launch_bounds__(64 * kMlpWaves) __attribute__((amdgpu_waves_per_eu(3))) k_mlp(MlpDev P, const float* __restrict__ G, float* __restrict__ dG,
                                           float* __restrict__ e_atom, const int* __restrict__ type_idx) {
  typedef typename Prec<KIND>::T T;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  T* lds = reinterpret_cast<T*>(lds_raw);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  constexpr int NT = 64 * kMlpWaves;
  const int tile = blockIdx.x;
  int e = 0;
  while (e + 1 < kMaxType && tile >= P.tile_off[e + 1]) ++e;
  const int r0 = (tile - P.tile_off[e]) * 32;
  const int nrows = min(32, P.count[e] - r0);
  const int* rows = type_idx + P.type_off[e] + r0;
  const int D = P.D, Dp = P.Dp;
  T* Gs = lds + P.slab_off[0];
  const int ldg = P.ld[0];

  // stage the descriptor rows standardised in fp32, zero padding rows/cols; row ids go through
  // LDS so every global load of the loop is independent of the others (they overlap)
  __shared__ int srow[32];
  __shared__ float s_wout[32 * kMlpWaves];  // output-layer weights of the element
  __shared__ float s_epart[kMlpWaves][32];  // per-wave energy partial sums of the rows
  if (tid < 32) srow[tid] = tid < nrows ? rows[tid] : -1;
  for (int c = tid; c < P.H[NH - 1]; c += NT) s_wout[c] = P.wout[e][c];
  __syncthreads();
  const float* shv = P.shift[e];
  const float* scv = P.scale[e];
  if ((D & 3) == 0) {
    const int D4 = D >> 2, Dp4 = Dp >> 2;
    const float4* G4 = reinterpret_cast<const float4*>(G);
    const float4* sh4 = reinterpret_cast<const float4*>(shv);
    const float4* sc4 = reinterpret_cast<const float4*>(scv);
#pragma unroll 4
    for (int idx = tid; idx < 32 * Dp4; idx += NT) {
      const int r = idx / Dp4, c4 = idx - r * Dp4;
      const int row = srow[r];
      // loads at clamped in-bounds addresses, padding selected afterwards: no branch around
      // the loads, so the unrolled iterations' loads are all in flight together
      const bool ok = row >= 0 && c4 < D4;
      const int rr = row >= 0 ? row : srow[0], cc = min(c4, D4 - 1);
      const float4 g = G4[(size_t)rr * D4 + cc], sh = sh4[cc], sc = sc4[cc];
      float4 v = make_float4((g.x - sh.x) * sc.x, (g.y - sh.y) * sc.y, (g.z - sh.z) * sc.z, (g.w - sh.w) * sc.w);
      if (!ok) v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      T* d = Gs + r * ldg + 4 * c4;
      d[0] = Prec<KIND>::cvt(v.x); d[1] = Prec<KIND>::cvt(v.y); d[2] = Prec<KIND>::cvt(v.z); d[3] = Prec<KIND>::cvt(v.w);
    }
  } else {
#pragma unroll 4
    for (int idx = tid; idx < 32 * Dp; idx += NT) {
      const int r = idx / Dp, c = idx - r * Dp;
      const int row = srow[r];
      float v = 0.0f;
      if (row >= 0 && c < D) v = (G[(size_t)row * D + c] - shv[c]) * scv[c];
      Gs[r * ldg + c] = Prec<KIND>::cvt(v);
    }
  }
  __syncthreads();

  const int row = lane & 31, cb = wave, f0 = cb * 32 + 4 * (lane >> 5);  // this lane: features f0 + 8q + t
  float tv[NH > 1 ? NH - 1 : 1][16];  // 1 - a_l^2 of this lane's features, hidden layers 0 .. NH-2
  // forward
#pragma unroll
  for (int l = 0; l < NH; ++l) {
    const T* src = l == 0 ? Gs : lds + P.slab_off[l];
    const int lds_src = P.ld[l];
    const int K = l == 0 ? Dp : P.H[l - 1];
    T* dst = lds + P.slab_off[l + 1];
    const int ldd = P.ld[l + 1];
    const T* wt = reinterpret_cast<const T*>(P.Wt[e][l]);
    const float* bias = P.b[e][l];
    const float un = P.r_w[l];
    const bool have = cb < P.H[l] / 32;
    float es = 0.0f;  // last layer: this lane's part of its row's energy
    if (have) {
      float4 bq[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) bq[q] = *reinterpret_cast<const float4*>(bias + f0 + 8 * q);  // before the chain
      f32x16 acc;
      if constexpr (KIND == 1)
        acc = gemm_block_pairs(src, nullptr, lds_src, wt, reinterpret_cast<const T*>(P.Wt_lo[e][l]), K, cb, lane);
      else
        acc = gemm_block<KIND>(src, lds_src, wt, K, cb, lane);
      const float sc = P.out_scale * P.s_d[NH - 1];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const float bv[4] = {bq[q].x, bq[q].y, bq[q].z, bq[q].w};
        T o[4], lo[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          float a, d;
          tanh_d(acc[4 * q + t] * un + bv[t], a, d);
          if (l + 1 < NH) {
            o[t] = Prec<KIND>::cvt(a);
            tv[l < NH - 1 ? l : 0][4 * q + t] = d;
          } else {
            // d_L = s_out w_out (1 - a_L^2), scaled by s_d[NH-1]; the energy uses the FP32 a_L
            const float w = s_wout[f0 + 8 * q + t];
            es += a * w;
            const float v = sc * w * d;
            o[t] = Prec<KIND>::cvt(v);
            if constexpr (KIND == 1) lo[t] = Prec<KIND>::cvt(v - Prec<KIND>::back(o[t]));
          }
        }
        store4(dst + row * ldd + f0 + 8 * q, o);
        if constexpr (KIND == 1)
          if (l + 1 == NH) store4(lds + P.lo_off[NH] + row * ldd + f0 + 8 * q, lo);
      }
    }
    if (l + 1 == NH) {
      es += __shfl_xor(es, 32, 64);  // lanes l and l + 32 hold the two halves of row l % 32
      if (lane < 32) s_epart[wave][lane] = es;
    }
    __syncthreads();
  }
  // energy of each row: the waves' partial sums in a fixed order
  if (tid < nrows) {
    float s = 0.0f;
#pragma unroll
    for (int w = 0; w < kMlpWaves; ++w) s += s_epart[w][tid];
    e_atom[srow[tid]] = P.out_scale * (s + P.bout[e]);
  }
#pragma unroll
  for (int l = NH - 1; l >= 1; --l) {
    // d_{l-1} = (d_l W_l^T) * (1 - a_{l-1}^2): A = d_l [32 x H_l], B[k=out][n=in] = W_l [in][out]
    const T* dl = lds + P.slab_off[l + 1];
    const int ldl2 = P.ld[l + 1];
    T* prev = lds + P.slab_off[l];
    const int ldp = P.ld[l];
    const T* w = reinterpret_cast<const T*>(P.W[e][l]);
    const float un = P.r_dw[l];
    const float sd = P.s_d[l - 1];
    if (cb < P.H[l - 1] / 32) {
      f32x16 acc;
      if constexpr (KIND == 1)
        acc = gemm_block_pairs(dl, lds + P.lo_off[l + 1], ldl2, w, reinterpret_cast<const T*>(P.W_lo[e][l]), P.H[l], cb,
                               lane);
      else
        acc = gemm_block<KIND>(dl, ldl2, w, P.H[l], cb, lane);
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        T* p = prev + row * ldp + f0 + 8 * q;
        T hi[4], lo[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          const float v = acc[4 * q + t] * un * tv[l - 1][4 * q + t] * sd;
          hi[t] = Prec<KIND>::cvt(v);
          if constexpr (KIND == 1) lo[t] = Prec<KIND>::cvt(v - Prec<KIND>::back(hi[t]));
        }
        store4(p, hi);
        if constexpr (KIND == 1) store4(lds + P.lo_off[l] + row * ldp + f0 + 8 * q, lo);
      }
    }
    __syncthreads();
  }
  // dE/dG = (d_1 W_1^T) * scale: A = d_1 [32 x H_0], B = W_0 [Dp][H_0]
  {
    const T* d1 = lds + P.slab_off[1];
    const T* w = reinterpret_cast<const T*>(P.W[e][0]);
    const float un = P.r_dw[0];
    for (int cb2 = wave; cb2 < Dp / 32; cb2 += kMlpWaves) {
      f32x16 acc;
      if constexpr (KIND == 1)
        acc = gemm_block_pairs(d1, lds + P.lo_off[1], P.ld[1], w, reinterpret_cast<const T*>(P.W_lo[e][0]), P.H[0], cb2,
                               lane);
      else
        acc = gemm_block<KIND>(d1, P.ld[1], w, P.H[0], cb2, lane);
      const int f02 = cb2 * 32 + 4 * (lane >> 5);
      if (row < nrows) {
        float* out = dG + (size_t)srow[row] * D;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int f = f02 + 8 * q;
          if ((D & 3) == 0) {
            if (f < D) {  // D a multiple of 4: the 4 features are all in or all out
              const float4 sc = *reinterpret_cast<const float4*>(scv + f);
              *reinterpret_cast<float4*>(out + f) =
                  make_float4(acc[4 * q] * sc.x * un, acc[4 * q + 1] * sc.y * un, acc[4 * q + 2] * sc.z * un,
                              acc[4 * q + 3] * sc.w * un);
            }
          } else {
#pragma unroll
            for (int t = 0; t < 4; ++t)
              if (f + t < D) out[f + t] = acc[4 * q + t] * (scv[f + t] * un);
          }
        }
      }
    }
  }
}

// Uniform-width fast path of the 16-bit forms (FP16, one-term BF16): every hidden layer 128
// wide (one 32-column block per wave in every product), descriptors at most 128 wide (D a
// multiple of 4).  Same data layout, LDS slabs and per-lane fragment mapping as k_mlp; what the
// fixed width buys is VALU economy, which bounds this kernel (the MFMA work of a 32-row tile is a
// fifth of its VALU time): compile-time strides and trip counts (no per-fragment bounds tests or
// address arithmetic), descriptor staging with one column group per thread (no index division),
// and a cheaper epilogue.  With e = exp(-2|z|) and r = 1/(1 + e): a = sign(z) (r - e r) and the
// stored derivative is t' = e r^2 = (1 - a^2) / 4; the factor 4 is restored in the power-of-two
// scale constants of the backward products (exact), so each value costs one multiply less than
// forming 1 - a^2 = 4 e r^2 and (1 - e) r separately.
constexpr int kW = kMaxWidth;   // the hidden width of the fast path
constexpr int kLdW = kW + 8;    // its LDS row stride (16-bit elements)

// The B fragments of one wave's 32-column block cb of fragment-ordered weights bm [n][K]
// (K <= 128, a multiple of 16), one 16-byte load per fragment from L2
template <int KIND>
__device__ __forceinline__ void load_b(typename Prec<KIND>::F (&b)[8], const typename Prec<KIND>::T* bm, int K, int cb,
                                       int lane) {
  const typename Prec<KIND>::T* bp = bm + (size_t)cb * (K / 16) * 512 + lane * 8;
#pragma unroll
  for (int s = 0; s < 8; ++s)
    if (s * 16 < K) b[s] = Prec<KIND>::ld(bp + s * 512);
}

// 32 x 32 output block of A [32 x K] (LDS, row stride lda) times the B fragments b: the A
// fragments in two halves (register budget), then the MFMA chain
template <int KIND>
__device__ __forceinline__ f32x16 gemm_b(const typename Prec<KIND>::T* A, int lda, const typename Prec<KIND>::F (&b)[8],
                                         int K, int lane) {
  typedef Prec<KIND> PR;
  typedef typename PR::F F;
  f32x16 acc;
#pragma unroll
  for (int j = 0; j < 16; ++j) acc[j] = 0.0f;
  const typename PR::T* arow = A + (lane & 31) * lda + 8 * (lane >> 5);
#pragma unroll
  for (int h = 0; h < 8; h += 4) {
    F a[4];
#pragma unroll
    for (int s = 0; s < 4; ++s)
      if ((h + s) * 16 < K) a[s] = PR::ld(arow + (h + s) * 16);
#pragma unroll
    for (int s = 0; s < 4; ++s)
      if ((h + s) * 16 < K) acc = PR::mma(b[h + s], a[s], acc);
  }
  return acc;
}

// (three waves per SIMD: registers for the fragments in flight beside the live ones)
template <int KIND, int NH>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3)))
k_mlp_w(MlpDev P, const float* __restrict__ G, float* __restrict__ dG, float* __restrict__ e_atom,
        const int* __restrict__ type_idx) {
  typedef Prec<KIND> PR;
  typedef typename PR::T T;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  T* lds = reinterpret_cast<T*>(lds_raw);
  const int tid = threadIdx.x, lane = tid & 63, cb = tid >> 6;
  const int tile = blockIdx.x;
  int e = 0;
  while (e + 1 < kMaxType && tile >= P.tile_off[e + 1]) ++e;
  const int r0 = (tile - P.tile_off[e]) * 32;
  const int nrows = min(32, P.count[e] - r0);
  const int* rows = type_idx + P.type_off[e] + r0;
  const int D = P.D, Dp = P.Dp, ldg = P.ld[0];
  T* Gs = lds;
  T* act = lds + P.slab_off[1];  // slab s = 1 .. NH at act + (s - 1) * 32 * kLdW

  // B fragments of the next product, issued one product ahead: layer 0's before the staging,
  // each later product's right after the previous MFMA chain, so the L2 round trip overlaps an
  // epilogue and a barrier instead of stalling the chain
  typename PR::F b[8];
  load_b<KIND>(b, reinterpret_cast<const T*>(P.Wt[e][0]), Dp, cb, lane);

  __shared__ int srow[32];
  // the epilogues work on y = kz, k = 2 / ln 2, so that exp(-2|z|) = 2^(-|y|) is one instruction:
  // biases staged as k b; output weights staged twice, for the energy and pre-multiplied by the
  // scale constant of d_L
  constexpr float kTanh = 2.8853900817779268f;
  __shared__ __align__(16) float s_wout[kW];
  __shared__ __align__(16) float s_wd[kW];
  __shared__ __align__(16) float s_bias[NH][kW];
  __shared__ float s_epart[4][32];
  if (tid < 32) srow[tid] = tid < nrows ? rows[tid] : -1;
  if (tid < kW) {
    const float w = P.wout[e][tid];
    s_wout[tid] = w;
    s_wd[tid] = 4.0f * P.out_scale * P.s_d[NH - 1] * w;
  }
#pragma unroll
  for (int l = 0; l < NH; ++l)
    if (tid < kW) s_bias[l][tid] = kTanh * P.b[e][l][tid];
  __syncthreads();
  // standardised descriptors: thread -> column group c4 = tid % 32 of rows tid / 32 + 8 i
  {
    const int c4 = tid & 31, D4 = D >> 2;
    if (c4 < (Dp >> 2)) {
      const int cc = min(c4, D4 - 1);
      const float4 sh = reinterpret_cast<const float4*>(P.shift[e])[cc];
      const float4 sc = reinterpret_cast<const float4*>(P.scale[e])[cc];
      const float4* G4 = reinterpret_cast<const float4*>(G);
      float4 g[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int row = srow[(tid >> 5) + 8 * i];
        g[i] = G4[(size_t)(row >= 0 ? row : srow[0]) * D4 + cc];  // in-bounds clamp; padding zeroed below
      }
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int r = (tid >> 5) + 8 * i;
        float4 v = make_float4((g[i].x - sh.x) * sc.x, (g[i].y - sh.y) * sc.y, (g[i].z - sh.z) * sc.z,
                               (g[i].w - sh.w) * sc.w);
        if (c4 >= D4 || srow[r] < 0) v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        const T o[4] = {PR::cvt(v.x), PR::cvt(v.y), PR::cvt(v.z), PR::cvt(v.w)};
        store4(Gs + r * ldg + 4 * c4, o);
      }
    }
  }
  __syncthreads();

  const int row = lane & 31, f0 = cb * 32 + 4 * (lane >> 5);  // this lane: features f0 + 8q + t of row `row`
  float tv[NH > 1 ? NH - 1 : 1][16];                           // t' of the hidden layers 0 .. NH-2
  // forward
#pragma unroll
  for (int l = 0; l < NH; ++l) {
    const T* src = l == 0 ? Gs : act + (l - 1) * 32 * kLdW;
    T* dst = act + l * 32 * kLdW;
    const f32x16 acc = gemm_b<KIND>(src, l == 0 ? ldg : kLdW, b, l == 0 ? Dp : kW, lane);
    if (l + 1 < NH)
      load_b<KIND>(b, reinterpret_cast<const T*>(P.Wt[e][l + 1]), kW, cb, lane);
    else if (NH > 1)
      load_b<KIND>(b, reinterpret_cast<const T*>(P.W[e][NH - 1]), kW, cb, lane);
    else if (cb < (Dp >> 5))
      load_b<KIND>(b, reinterpret_cast<const T*>(P.W[e][0]), kW, cb, lane);
    const float kun = kTanh * P.r_w[l];
    float es = 0.0f;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const float4 bq = *reinterpret_cast<const float4*>(&s_bias[l][f0 + 8 * q]);
      const float bv[4] = {bq.x, bq.y, bq.z, bq.w};
      float wq[4] = {0.0f, 0.0f, 0.0f, 0.0f}, wdq[4] = {0.0f, 0.0f, 0.0f, 0.0f};
      if (l + 1 == NH) {
        const float4 w4 = *reinterpret_cast<const float4*>(&s_wout[f0 + 8 * q]);
        const float4 d4 = *reinterpret_cast<const float4*>(&s_wd[f0 + 8 * q]);
        wq[0] = w4.x; wq[1] = w4.y; wq[2] = w4.z; wq[3] = w4.w;
        wdq[0] = d4.x; wdq[1] = d4.y; wdq[2] = d4.z; wdq[3] = d4.w;
      }
      T o[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const float y = fmaf(acc[4 * q + t], kun, bv[t]);  // k z
        const float ex = __builtin_amdgcn_exp2f(-fabsf(y));  // exp(-2|z|)
        const float r = __builtin_amdgcn_rcpf(1.0f + ex);
        const float er = ex * r;
        // sign of y by integer ops (r - er >= 0; a copysign here is moved past the 16-bit rounding
        // below by the compiler, which then repacks the halves)
        const float a = __uint_as_float(__float_as_uint(r - er) | (__float_as_uint(y) & 0x80000000u));
        const float td = er * r;  // (1 - a^2) / 4
        if (l + 1 < NH) {
          o[t] = PR::cvt(a);
          tv[l < NH - 1 ? l : 0][4 * q + t] = td;
        } else {
          // d_L = s_out w_out (1 - a_L^2) = 4 s_out w_out t', scaled by s_d[NH-1]; energy from the FP32 a_L
          es += a * wq[t];
          o[t] = PR::cvt(wdq[t] * td);
        }
      }
      store4(dst + row * kLdW + f0 + 8 * q, o);
    }
    if (l + 1 == NH) {
      es += __shfl_xor(es, 32, 64);  // lanes l and l + 32 hold the two halves of row l % 32
      if (lane < 32) s_epart[cb][lane] = es;
    }
    __syncthreads();
  }
  if (tid < nrows) {
    float s = 0.0f;
#pragma unroll
    for (int w = 0; w < 4; ++w) s += s_epart[w][tid];
    e_atom[srow[tid]] = P.out_scale * (s + P.bout[e]);
  }
  // backward: d_{l-1} = (d_l W_l^T) (1 - a_{l-1}^2) = 4 (d_l W_l^T) t'_{l-1}
#pragma unroll
  for (int l = NH - 1; l >= 1; --l) {
    const f32x16 acc = gemm_b<KIND>(act + l * 32 * kLdW, kLdW, b, kW, lane);
    if (l > 1)
      load_b<KIND>(b, reinterpret_cast<const T*>(P.W[e][l - 1]), kW, cb, lane);
    else if (cb < (Dp >> 5))
      load_b<KIND>(b, reinterpret_cast<const T*>(P.W[e][0]), kW, cb, lane);
    const float c = 4.0f * P.r_dw[l] * P.s_d[l - 1];
    T* prev = act + (l - 1) * 32 * kLdW + row * kLdW + f0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      T o[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) o[t] = PR::cvt(acc[4 * q + t] * c * tv[l - 1][4 * q + t]);
      store4(prev + 8 * q, o);
    }
    __syncthreads();
  }
  // dE/dG = (d_1 W_1^T) * scale: column block cb of the padded descriptor width
  if (cb < (Dp >> 5)) {
    const f32x16 acc = gemm_b<KIND>(act, kLdW, b, kW, lane);
    const float un = P.r_dw[0];
    if (row < nrows) {
      float* out = dG + (size_t)srow[row] * D;
      const float* scv = P.scale[e];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int f = f0 + 8 * q;
        if (f < D) {  // D a multiple of 4: the 4 features are all in or all out
          const float4 sc = *reinterpret_cast<const float4*>(scv + f);
          *reinterpret_cast<float4*>(out + f) = make_float4(acc[4 * q] * sc.x * un, acc[4 * q + 1] * sc.y * un,
                                                            acc[4 * q + 2] * sc.z * un, acc[4 * q + 3] * sc.w * un);
        }
      }
    }
  }
}

// rows of each element, wave-aggregated: one atomic per wave and element (a per-row atomic on
// a handful of counters serialises)
__global__ void __launch_bounds__(256) k_type_index(int n, int ntype, const int* typ, const int* off, int* fill,
                                                    int* idx) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const int t = i < n ? typ[i] : -1;
  for (int e = 0; e < ntype; ++e) {
    const unsigned long long bal = __ballot(t == e);
    if (!bal) continue;
    int base = 0;
    if (lane == __ffsll((long long)bal) - 1) base = atomicAdd(&fill[e], __popcll(bal));
    base = __shfl(base, __ffsll((long long)bal) - 1, 64);
    if (t == e) {
      const int rank = __builtin_amdgcn_mbcnt_hi((unsigned)(bal >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)bal, 0));
      idx[off[e] + base + rank] = i;
    }
  }
}

__global__ void k_zero_small(int* p, int n) {
  if (threadIdx.x < n) p[threadIdx.x] = 0;
}

__global__ void k_esum(int n, const float* e, double* part, double* part2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  block_sum2_store<256>(i < n ? double(e[i]) : 0.0, 0.0, part, part2);
}

}  // namespace

struct MlpState {
  MlpDev dev{};
  int kind = 1;
  int nl = 0;
  std::vector<int> widths;
  size_t lds_bytes = 0;
  bool uniform = false;  // k_mlp_w: 16-bit form, every hidden layer kMaxWidth wide, D <= kMaxWidth, D % 4 == 0
  int ntiles = 0;
  std::vector<void*> bufs;
  int* d_type_off = nullptr;
};

void mlp_state_free(nnpe_ctx* ctx) {
  MlpState* S = ctx->mlp;
  if (!S) return;
  for (void* p : S->bufs) (void)hipFree(p);
  dev_free(S->d_type_off);
  delete S;
  ctx->mlp = nullptr;
}

// float -> bf16 bits, round to nearest even (finite inputs)
static uint16_t bf16_rne(float x) {
  uint32_t u;
  std::memcpy(&u, &x, 4);
  return uint16_t((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
static float bf16_value(uint16_t b) {
  const uint32_t u = uint32_t(b) << 16;
  float x;
  std::memcpy(&x, &u, 4);
  return x;
}

static float pow2_scale(double maxabs) {
  if (!(maxabs > 0) || !std::isfinite(maxabs)) return 1.0f;
  return float(std::ldexp(1.0, -int(std::floor(std::log2(maxabs)))));
}

const int* mlp_type_off(nnpe_ctx* ctx) { return ctx->type_idx && ctx->mlp ? ctx->mlp->d_type_off : nullptr; }

// rows of each element (rebuilt after every re-sort; the list rebuild does it in its copy-back
// kernel)
int type_index_enqueue(nnpe_ctx* ctx) {
  if (!ctx->type_idx) return NNPE_OK;
  MlpState* S = ctx->mlp;
  hipLaunchKernelGGL(k_zero_small, dim3(1), dim3(64), 0, ctx->stream, ctx->type_fill, ctx->ntype);
  hipLaunchKernelGGL(k_type_index, dim3(blocks_for(ctx->natom, 256)), dim3(256), 0, ctx->stream, ctx->natom, ctx->ntype,
                     ctx->typ, S->d_type_off, ctx->type_fill, ctx->type_idx);
  NNPE_CHECK_LAUNCH(ctx);
  return NNPE_OK;
}

extern "C" int nnpe_mlp_set(nnpe_ctx* ctx, int nlayer, const int* widths, const float* params, float out_scale,
                            int precision, const float* in_shift, const float* in_scale) {
  if (ctx) md_state_free(ctx);  // captured step graphs hold the old parameters
  if (!ctx || nlayer < 2 || nlayer > kMaxLayers || !widths || !params || precision < 0 || precision > 3)
    return ctx_fail(ctx, NNPE_ERR_ARG, "nnpe_mlp_set: bad argument (2..6 weight layers, precision 0..3)");
  if (!ctx->have_system) return ctx_fail(ctx, NNPE_ERR_STATE, "no system");
  if (ctx->ntype > kMaxType) return ctx_fail(ctx, NNPE_ERR_UNSUPPORTED, "at most 4 elements");
  if (widths[nlayer] != 1) return ctx_fail(ctx, NNPE_ERR_ARG, "the output layer must have width 1");
  for (int l = 1; l < nlayer; ++l)
    if (widths[l] % 32 || widths[l] <= 0 || widths[l] > kMaxWidth)
      return ctx_fail(ctx, NNPE_ERR_ARG, "hidden widths must be multiples of 32, at most 128");
  mlp_state_free(ctx);
  MlpState* S = ctx->mlp = new MlpState;
  S->kind = precision; S->nl = nlayer;
  S->widths.assign(widths, widths + nlayer + 1);
  MlpDev& d = S->dev;
  d.nh = nlayer - 1;
  d.D = widths[0];
  d.Dp = (d.D + 31) / 32 * 32;
  for (int l = 0; l < d.nh; ++l) d.H[l] = widths[l + 1];
  d.out_scale = out_scale;
  const int esz = precision == 0 ? 4 : 2;
  const int pad = precision == 0 ? 4 : 8;  // LDS row padding (elements)
  d.ld[0] = d.Dp + pad;
  for (int l = 0; l < d.nh; ++l) d.ld[l + 1] = d.H[l] + pad;
  // slabs: descriptors (layer-0 input), then one per hidden layer (activations, overwritten in
  // place by the gradients).  BF16 form: the gradients' low parts of layer s are live only
  // while layer s-1's gradient is formed, so they alternate between two buffers, the even
  // layers' in the descriptor slab's place (dead after the first forward layer)
  for (int s = 0; s <= kMaxLayers; ++s) d.lo_off[s] = 0;
  int ldmax = 0;
  for (int s = 1; s <= d.nh; ++s) ldmax = std::max(ldmax, d.ld[s]);
  int off = 32 * d.ld[0];
  if (precision == 1) off = std::max(off, 32 * ldmax);
  d.slab_off[0] = 0;
  for (int s = 1; s <= d.nh; ++s) { d.slab_off[s] = off; off += 32 * d.ld[s]; }
  if (precision == 1) {
    const int odd = off;
    off += 32 * ldmax;
    for (int s = 1; s <= d.nh; ++s) d.lo_off[s] = (s % 2) ? odd : 0;
  }
  S->lds_bytes = size_t(off) * esz;
  S->uniform = (precision == 2 || precision == 3) && d.Dp <= kMaxWidth && d.D % 4 == 0;
  for (int l = 0; l < d.nh; ++l) S->uniform = S->uniform && d.H[l] == kMaxWidth;
  if (S->lds_bytes > 64 * 1024) return ctx_fail(ctx, NNPE_ERR_UNSUPPORTED, "MLP tile does not fit the 64 KiB LDS budget");
  // weights: per element, per layer W [in][out] then b [out]
  size_t pofs = 0;
  std::vector<double> maxw(nlayer, 0.0), rowsum(nlayer, 0.0);
  for (int t = 0; t < ctx->ntype; ++t) {
    size_t o = pofs + 0;
    for (int l = 0; l < nlayer; ++l) {
      const int in = widths[l], out = widths[l + 1];
      for (int i = 0; i < in; ++i) {
        double rs = 0;
        for (int k = 0; k < out; ++k) {
          double w = std::fabs(params[o + size_t(i) * out + k]);
          maxw[l] = std::max(maxw[l], w);
          rs += w;
        }
        rowsum[l] = std::max(rowsum[l], rs);
      }
      o += size_t(in) * out + out;
    }
    pofs = o;
  }
  // per-layer scales (FP16 form only; BF16 and FP32 have the FP32 exponent range)
  for (int l = 0; l < kMaxLayers; ++l) { d.s_w[l] = 1.0f; d.s_d[l] = 1.0f; }
  if (precision == 2) {
    for (int l = 0; l < d.nh; ++l) d.s_w[l] = pow2_scale(maxw[l]);
    // gradient bounds: |d_L| <= s_out max|w_out|, |d_{l-1}| <= |d_l| * max_i sum_j |W_l[i][j]|
    double bnd = std::fabs(out_scale) * maxw[nlayer - 1];
    d.s_d[d.nh - 1] = pow2_scale(bnd);
    for (int l = d.nh - 1; l >= 1; --l) { bnd *= rowsum[l]; d.s_d[l - 1] = pow2_scale(bnd); }
  }
  for (int l = 0; l < kMaxLayers; ++l) { d.r_w[l] = 1.0f / d.s_w[l]; d.r_dw[l] = 1.0f / (d.s_d[l] * d.s_w[l]); }
  pofs = 0;
  for (int t = 0; t < ctx->ntype; ++t) {
    for (int l = 0; l < nlayer; ++l) {
      const int in = widths[l], out = widths[l + 1];
      const float* W = params + pofs;
      const float* b = W + size_t(in) * out;
      pofs += size_t(in) * out + out;
      if (l == nlayer - 1) {  // output layer: fp32 vector + scalar
        float* dw = nullptr;
        NNPE_CHECK_HIP(ctx, hipMalloc(reinterpret_cast<void**>(&dw), sizeof(float) * in));
        S->bufs.push_back(dw);
        NNPE_CHECK_HIP(ctx, hipMemcpy(dw, W, sizeof(float) * in, hipMemcpyHostToDevice));
        d.wout[t] = dw;
        d.bout[t] = b[0];
        continue;
      }
      const int inp = l == 0 ? d.Dp : in;
      const float sw = d.s_w[l];
      std::vector<float> wt(size_t(out) * inp, 0.0f), wn(size_t(inp) * out, 0.0f);
      for (int i = 0; i < in; ++i)
        for (int k = 0; k < out; ++k) {
          wt[size_t(k) * inp + i] = W[size_t(i) * out + k] * sw;
          wn[size_t(i) * out + k] = W[size_t(i) * out + k] * sw;
        }
      // lo = false: the value rounded to the storage type; lo = true (BF16 form): the rounding
      // remainder x - hi rounded to bf16.  h is [n][K] row-major; it is stored in MFMA fragment
      // order (frag_ptr): fragment (cb, t) = 64 lanes x KS/2 elements, lane l holding row
      // cb*32 + l%32, k = t*KS + (KS/2)*(l/32) + e
      const int KS = precision == 0 ? 2 : 16, HALF = KS / 2;
      auto upload = [&](const std::vector<float>& hrow, int n, int K, const void*& dst, bool lo) -> int {
        std::vector<float> h(hrow.size());
        for (int cb = 0; cb < n / 32; ++cb)
          for (int t = 0; t < K / KS; ++t)
            for (int ln = 0; ln < 64; ++ln)
              for (int e2 = 0; e2 < HALF; ++e2)
                h[((size_t(cb) * (K / KS) + t) * 64 + ln) * HALF + e2] =
                    hrow[size_t(cb * 32 + ln % 32) * K + t * KS + HALF * (ln / 32) + e2];
        void* p = nullptr;
        NNPE_CHECK_HIP(ctx, hipMalloc(&p, h.size() * esz));
        S->bufs.push_back(p);
        if (precision == 0) {
          NNPE_CHECK_HIP(ctx, hipMemcpy(p, h.data(), h.size() * 4, hipMemcpyHostToDevice));
        } else {
          std::vector<uint16_t> hb(h.size());
          for (size_t k = 0; k < h.size(); ++k) {
            if (precision == 1 || precision == 3) {
              const uint16_t hi = bf16_rne(h[k]);
              hb[k] = lo ? bf16_rne(h[k] - bf16_value(hi)) : hi;
            } else {
              _Float16 hv = (_Float16)h[k];
              std::memcpy(&hb[k], &hv, 2);
            }
          }
          NNPE_CHECK_HIP(ctx, hipMemcpy(p, hb.data(), hb.size() * 2, hipMemcpyHostToDevice));
        }
        dst = p;
        return NNPE_OK;
      };
      int rc = upload(wt, out, inp, d.Wt[t][l], false);
      if (rc) return rc;
      rc = upload(wn, inp, out, d.W[t][l], false);
      if (rc) return rc;
      if (precision == 1) {
        rc = upload(wt, out, inp, d.Wt_lo[t][l], true);
        if (rc) return rc;
        rc = upload(wn, inp, out, d.W_lo[t][l], true);
        if (rc) return rc;
      }
      float* db = nullptr;
      NNPE_CHECK_HIP(ctx, hipMalloc(reinterpret_cast<void**>(&db), sizeof(float) * out));
      S->bufs.push_back(db);
      NNPE_CHECK_HIP(ctx, hipMemcpy(db, b, sizeof(float) * out, hipMemcpyHostToDevice));
      d.b[t][l] = db;
    }
  }
  // input standardisation, identity where not given
  for (int t = 0; t < ctx->ntype; ++t) {
    std::vector<float> sh(d.D, 0.0f), sc(d.D, 1.0f);
    if (in_shift) std::copy(in_shift + size_t(t) * d.D, in_shift + size_t(t + 1) * d.D, sh.begin());
    if (in_scale) std::copy(in_scale + size_t(t) * d.D, in_scale + size_t(t + 1) * d.D, sc.begin());
    for (int k = 0; k < d.D; ++k)
      if (!std::isfinite(sh[k]) || !std::isfinite(sc[k])) return ctx_fail(ctx, NNPE_ERR_ARG, "non-finite input standardisation");
    float* p = nullptr;
    NNPE_CHECK_HIP(ctx, hipMalloc(reinterpret_cast<void**>(&p), sizeof(float) * 2 * d.D));
    S->bufs.push_back(p);
    NNPE_CHECK_HIP(ctx, hipMemcpy(p, sh.data(), sizeof(float) * d.D, hipMemcpyHostToDevice));
    NNPE_CHECK_HIP(ctx, hipMemcpy(p + d.D, sc.data(), sizeof(float) * d.D, hipMemcpyHostToDevice));
    d.shift[t] = p;
    d.scale[t] = p + d.D;
  }
  NNPE_CHECK_HIP(ctx, dev_alloc(ctx->type_idx, ctx->natom));
  NNPE_CHECK_HIP(ctx, dev_alloc(ctx->type_fill, ctx->ntype));
  int rc = mlp_retile(ctx);
  if (rc) return rc;
  rc = nnp_desc_reserve(ctx, d.D);
  if (rc) return rc;
  return type_index_enqueue(ctx);
}

// tiles of 32 atoms per element, from the element counts
int mlp_retile(nnpe_ctx* ctx) {
  MlpState* S = ctx->mlp;
  if (!S) return NNPE_OK;
  MlpDev& d = S->dev;
  std::vector<int> toff(ctx->ntype + 1, 0);
  ctx->h_type_off.assign(ctx->ntype, 0);
  int acc = 0;
  for (int t = 0; t < ctx->ntype; ++t) {
    ctx->h_type_off[t] = acc;
    acc += ctx->h_typecount[t];
    d.count[t] = ctx->h_typecount[t];
    d.type_off[t] = ctx->h_type_off[t];
    toff[t + 1] = toff[t] + (ctx->h_typecount[t] + 31) / 32;
  }
  for (int t = 0; t <= kMaxType; ++t) d.tile_off[t] = toff[std::min(t, ctx->ntype)];
  S->ntiles = toff[ctx->ntype];
  if (!S->d_type_off) NNPE_CHECK_HIP(ctx, dev_alloc(S->d_type_off, ctx->ntype));
  NNPE_CHECK_HIP(ctx, hipMemcpy(S->d_type_off, ctx->h_type_off.data(), sizeof(int) * ctx->ntype, hipMemcpyHostToDevice));
  return NNPE_OK;
}

int mlp_enqueue(nnpe_ctx* ctx, bool ev) {
  MlpState* S = ctx->mlp;
  const MlpDev& d = S->dev;
  ctx->dG_kind = S->kind;
  if (S->ntiles > 0) {
#define NNPE_MLP(K, NH)                                                                                      \
  hipLaunchKernelGGL((k_mlp<K, NH>), dim3(S->ntiles), dim3(64 * kMlpWaves), S->lds_bytes, ctx->stream, d, ctx->G, \
                     ctx->dG, ctx->e_atom, ctx->type_idx)
#define NNPE_MLP_K(K)                         \
  switch (d.nh) {                             \
    case 1: NNPE_MLP(K, 1); break;            \
    case 2: NNPE_MLP(K, 2); break;            \
    case 3: NNPE_MLP(K, 3); break;            \
    case 4: NNPE_MLP(K, 4); break;            \
    default: NNPE_MLP(K, 5); break;           \
  }
    if (S->uniform) {
#define NNPE_MLP_W(K)                                                                                         \
  switch (d.nh) {                                                                                             \
    case 1: hipLaunchKernelGGL((k_mlp_w<K, 1>), dim3(S->ntiles), dim3(256), S->lds_bytes, ctx->stream, d, ctx->G, ctx->dG, ctx->e_atom, ctx->type_idx); break; \
    case 2: hipLaunchKernelGGL((k_mlp_w<K, 2>), dim3(S->ntiles), dim3(256), S->lds_bytes, ctx->stream, d, ctx->G, ctx->dG, ctx->e_atom, ctx->type_idx); break; \
    case 3: hipLaunchKernelGGL((k_mlp_w<K, 3>), dim3(S->ntiles), dim3(256), S->lds_bytes, ctx->stream, d, ctx->G, ctx->dG, ctx->e_atom, ctx->type_idx); break; \
    case 4: hipLaunchKernelGGL((k_mlp_w<K, 4>), dim3(S->ntiles), dim3(256), S->lds_bytes, ctx->stream, d, ctx->G, ctx->dG, ctx->e_atom, ctx->type_idx); break; \
    default: hipLaunchKernelGGL((k_mlp_w<K, 5>), dim3(S->ntiles), dim3(256), S->lds_bytes, ctx->stream, d, ctx->G, ctx->dG, ctx->e_atom, ctx->type_idx); break; \
  }
      if (S->kind == 3) { NNPE_MLP_W(3) }
      else { NNPE_MLP_W(2) }
#undef NNPE_MLP_W
    } else if (S->kind == 0) { NNPE_MLP_K(0) }
    else if (S->kind == 1) { NNPE_MLP_K(1) }
    else if (S->kind == 3) { NNPE_MLP_K(3) }
    else { NNPE_MLP_K(2) }
#undef NNPE_MLP_K
#undef NNPE_MLP
    NNPE_CHECK_LAUNCH(ctx);
  }
  if (ev) {
    const int nb = blocks_for(ctx->natom, 256);
    hipLaunchKernelGGL(k_esum, dim3(nb), dim3(256), 0, ctx->stream, ctx->natom, ctx->e_atom, ev_part(ctx, NNPE_TERM_NN, 0),
                       ev_part(ctx, NNPE_TERM_NN, 1));
    NNPE_CHECK_LAUNCH(ctx);
    return ev_finish(ctx, NNPE_TERM_NN, nb);
  }
  return NNPE_OK;
}

extern "C" int nnpe_mlp_eval(nnpe_ctx* ctx, int ev) {
  if (!ctx) return NNPE_ERR_ARG;
  if (!ctx->mlp) return ctx_fail(ctx, NNPE_ERR_STATE, "nnpe_mlp_eval before nnpe_mlp_set");
  if (ctx->D_cur != ctx->mlp->dev.D) return ctx_fail(ctx, NNPE_ERR_STATE, "descriptor width differs from the network input");
  return mlp_enqueue(ctx, ev != 0);
}

// descriptor / gradient / per-atom energy exchange in atom-id order (the TFCALL input and
// output buffers of the reference, src/pair_nnp.cpp:490-496, :697-706)
__global__ void k_rows_to_ids(int n, int D, const int4* info, const float* src, float* dst) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int a = info[i].x;
  for (int k = 0; k < D; ++k) dst[(size_t)a * D + k] = src[(size_t)i * D + k];
}
__global__ void k_ids_to_rows(int n, int D, const int4* info, const float* src, float* dst) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int a = info[i].x;
  for (int k = 0; k < D; ++k) dst[(size_t)i * D + k] = src[(size_t)a * D + k];
}

extern "C" int nnpe_nnp_get(nnpe_ctx* ctx, int what, float* out, int* D) {
  if (!ctx || what < 0 || what > 2) return NNPE_ERR_ARG;
  if (!ctx->G || !ctx->D_cur) return ctx_fail(ctx, NNPE_ERR_STATE, "no descriptors computed");
  const int Dc = what == 2 ? 1 : ctx->D_cur;
  if (D) *D = Dc;
  if (!out) return NNPE_OK;
  float* tmp = nullptr;
  NNPE_CHECK_HIP(ctx, hipMallocAsync(reinterpret_cast<void**>(&tmp), sizeof(float) * size_t(ctx->natom) * Dc, ctx->stream));
  const float* src = what == 0 ? ctx->G : what == 1 ? ctx->dG : ctx->e_atom;
  hipLaunchKernelGGL(k_rows_to_ids, dim3(blocks_for(ctx->natom, 256)), dim3(256), 0, ctx->stream, ctx->natom, Dc, ctx->info, src, tmp);
  NNPE_CHECK_LAUNCH(ctx);
  NNPE_CHECK_HIP(ctx, hipMemcpyAsync(out, tmp, sizeof(float) * size_t(ctx->natom) * Dc, hipMemcpyDeviceToHost, ctx->stream));
  NNPE_CHECK_HIP(ctx, hipFreeAsync(tmp, ctx->stream));
  NNPE_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return NNPE_OK;
}

extern "C" int nnpe_nnp_set_dedg(nnpe_ctx* ctx, const float* dGin, int D) {
  if (!ctx || !dGin || D <= 0) return NNPE_ERR_ARG;
  int rc = nnp_desc_reserve(ctx, D);
  if (rc) return rc;
  float* tmp = nullptr;
  NNPE_CHECK_HIP(ctx, hipMallocAsync(reinterpret_cast<void**>(&tmp), sizeof(float) * size_t(ctx->natom) * D, ctx->stream));
  NNPE_CHECK_HIP(ctx, hipMemcpyAsync(tmp, dGin, sizeof(float) * size_t(ctx->natom) * D, hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(k_ids_to_rows, dim3(blocks_for(ctx->natom, 256)), dim3(256), 0, ctx->stream, ctx->natom, D, ctx->info, tmp, ctx->dG);
  NNPE_CHECK_LAUNCH(ctx);
  NNPE_CHECK_HIP(ctx, hipFreeAsync(tmp, ctx->stream));
  NNPE_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
  ctx->D_cur = D;
  ctx->dG_kind = 0;
  return NNPE_OK;
}
