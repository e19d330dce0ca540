// forces.hip -- tabulated LJ + shifted-force Coulomb pair forces and the bonded terms.
//
// Pair (LJTableCoulDSF::compute, src/pair.cpp:289-419): a group of kPairG lanes per atom walks
// its full list row (G consecutive entries per step), and the group's butterfly sum is written
// once per atom: no atomics, deterministic.  The F/r table sits in LDS.  The force arithmetic per
// pair is the reference's, in the same order:
//   fpair = lerp(F/r)*s_lj + qi*qj*K*(1/(r*rsq) + fc/r)*s_coul
// On energy steps every directed entry adds half of the pair energy and virial, evaluated in
// FP64 from the FP32 positions (the forces stay FP32 and identical to the other steps).
//
// Bonded (BOND / ANGLE / DIHEDRAL::compute, src/bond.cpp:87-116, src/angle.cpp:87-136,
// src/dihedral.cpp:90-147): one lane per atom.  Each term's energy is a polynomial of one
// scalar s of its geometry (bond length, cosine of the angle, cosine of the dihedral); the lane
// evaluates every term its atom belongs to and adds -U'(s) ds/dx for its own atom only, so each
// atom's bonded force has one writer and a fixed summation order, and all rows work at once.
#include <algorithm>
#include <cmath>
#include <mutex>

#include "internal.h"

namespace {

// special-pair factors (src/molecule.h:11): class 0 full, classes 1 and 2 excluded, class 3 scaled
constexpr float kSpecialLj3 = 0.5f, kSpecialCoul3 = float(5 / 6.);
constexpr float kElectricK = 332.06371f;                                      // src/environ.h:77
constexpr double kElectricK64 = 332.06371;
constexpr int kPairG = 32;  // lanes per atom
constexpr int kPairU = 4;   // list entries per lane in flight at once
// threads per workgroup: 16 atoms share one LDS copy of the table (256 threads: 0.1345 ms per
// MD step, 512: 0.1335, 1024: 0.1431, the pair kernel running beside the descriptor pass)
constexpr int kPairBlock = 512;

// cutoff constants of the energy pass in FP64 (LJTableCoulDSF ctor, src/pair.cpp:171-195)
struct PairEv64 {
  double cut, rdr, fc, vc;
};

__device__ __forceinline__ int pair_index(int a, int b, int ntype) {
  const int i = min(a, b), j = max(a, b);  // small non-negative values: 24-bit multiplies
  return __mul24(i, ntype) - (__mul24(i, i - 1) >> 1) + (j - i);
}

// TYE: the neighbour's element comes from its list entry (at most NNPE_ENTRY_MAX_TYPES
// elements) instead of a second gather, typ[j], per entry
template <bool EV, int G, bool TYE>
__global__ void __launch_bounds__(kPairBlock) k_pair(int n, int cap, const uint32_t* __restrict__ nbr,
                                              const int* __restrict__ cnt, const float4* __restrict__ xq,
                                              const int* __restrict__ typ, const BoxPar* __restrict__ boxes,
                                              const int* __restrict__ rbox, int ntype, int npair,
                                              float cutsq, float rdr, float fc, PairEv64 P64, int nb1,
                                              const float* __restrict__ tab, float4* __restrict__ f, double* pe,
                                              double* pw) {
#pragma clang fp contract(fast)
  extern __shared__ float s_tab[];  // [npair][nb1] F/r
  __shared__ float4 s_img[27];
  // TYE: the table row of each element pair, pair_index(a, b) * nb1, at s_prow[4a + b]; the
  // special-pair factors (LJ, Coulomb) of each class at s_fac[class]: two LDS reads per entry
  // in place of the index arithmetic and the selects
  __shared__ int s_prow[NNPE_ENTRY_MAX_TYPES * NNPE_ENTRY_MAX_TYPES];
  __shared__ float2 s_fac[4];
  for (int k = threadIdx.x; k < npair * nb1; k += blockDim.x) s_tab[k] = tab[k];
  image_table_fill(s_img);
  if (TYE && threadIdx.x < NNPE_ENTRY_MAX_TYPES * NNPE_ENTRY_MAX_TYPES) {
    const int a = threadIdx.x / NNPE_ENTRY_MAX_TYPES, b = threadIdx.x % NNPE_ENTRY_MAX_TYPES;
    s_prow[threadIdx.x] = a < ntype && b < ntype ? pair_index(a, b, ntype) * nb1 : 0;
  }
  if (threadIdx.x < 4) {
    const int c = threadIdx.x;
    s_fac[c] = c == 0 ? make_float2(1.0f, 1.0f) : (c == 3 ? make_float2(kSpecialLj3, kSpecialCoul3) : make_float2(0.0f, 0.0f));
  }
  __syncthreads();
  const int i = (xcd_block(gridDim.x) * kPairBlock + (int)threadIdx.x) / G;
  const int lane = threadIdx.x & (G - 1);
  double e_acc = 0, w_acc = 0;
  if (i < n) {
    const float4 xi = xq[i];
    const int ti = typ[i];
    const int* prow = s_prow + NNPE_ENTRY_MAX_TYPES * (TYE ? ti : 0);
    const BoxPar& B = boxes[rbox[i]];
    const float bl = B.bl;
    float fx = 0, fy = 0, fz = 0;
    double e = 0, w = 0;
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
        tr[u] = TYE ? NNPE_ENTRY_T(ec[u]) : typ[j];
      }
#pragma unroll
      for (int u = 0; u < kPairU; ++u) {
        if (k0 + u * G >= m) break;
        const uint32_t ent = ec[u];
        const int code = NNPE_ENTRY_CODE(ent);
        const float4 xj = image_pos_t(xr[u], code, bl, s_img);
        float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
        const float rsq = dist2_rn(dx, dy, dz);  // the cut as the reference rounds it
        if (!(rsq < cutsq)) continue;
        const float ir = __builtin_amdgcn_rsqf(rsq), r = rsq * ir;
        const int p = TYE ? 0 : pair_index(ti, tr[u], ntype);
        float ipw1 = r * rdr - 0.5f;
        const int t1 = int(ipw1), t2 = t1 + 1;
        const float ipw2 = ipw1 - float(t1);
        ipw1 = 1.0f - ipw2;
        const float* ft = s_tab + (TYE ? prow[tr[u]] : __mul24(p, nb1));
        const float2 fac = s_fac[NNPE_ENTRY_SP(ent)];
        const float slj = fac.x, scl = fac.y;
        float fpair = (ft[t1] * ipw1 + ft[t2] * ipw2) * slj;
        const float qq = xi.w * xj.w * kElectricK;
        fpair += qq * (ir * ir * ir + fc * ir) * scl;
        fx += (dx *= fpair); fy += (dy *= fpair); fz += (dz *= fpair);
        if (EV) {
          // Energies and virials sum many terms of both signs that cancel; they are evaluated
          // in FP64 (image shift, distance, table lerp, Coulomb) from the same FP32 inputs.
          const float4 mi = s_img[code];
          const double bd = double(bl);
          const double ex = double(xi.x) - fma(double(mi.x), bd, double(xr[u].x));
          const double ey = double(xi.y) - fma(double(mi.y), bd, double(xr[u].y));
          const double ez = double(xi.z) - fma(double(mi.z), bd, double(xr[u].z));
          const double rsq2 = ex * ex + ey * ey + ez * ez, r2 = sqrt(rsq2);
          double h1 = r2 * P64.rdr - 0.5;
          const int s1 = int(h1);
          const double h2 = h1 - double(s1);
          h1 = 1.0 - h2;
          const float* ut = tab + size_t(npair) * nb1 + (ft - s_tab);
          const double q2 = double(xi.w) * double(xj.w) * kElectricK64;
          const double u2 = (double(ut[s1]) * h1 + double(ut[s1 + 1]) * h2) * slj +
                            q2 * (1.0 / r2 + P64.vc + P64.fc * (P64.cut - r2)) * scl;
          const double fp2 = (double(ft[s1]) * h1 + double(ft[s1 + 1]) * h2) * slj +
                             q2 * (1.0 / (r2 * rsq2) + P64.fc / r2) * scl;
          e += 0.5 * u2;
          w += 0.5 * fp2 * rsq2;
        }
      }
    }
    fx = group_sum<G>(fx); fy = group_sum<G>(fy); fz = group_sum<G>(fz);
    if (lane == 0) f[i] = make_float4(fx, fy, fz, 0.0f);
    e_acc = e; w_acc = w;
  }
  if (EV) block_sum2_store<kPairBlock>(e_acc, w_acc, pe, pw);
}

// ---- bonded terms ------------------------------------------------------------------------------
// The energy functions and the stored coefficients are the reference's (src/bond.cpp:27-34,
// src/angle.cpp:26-33, src/dihedral.cpp:27-35), the evaluation is this file's own:
//   bond      s = r,          stored (r0, 2a, b, 2c):   U = (x^2 / 2) (2a (b - x)^2 + 2c),  x = s - r0
//   angle     s = cos(t),     stored (cos t0, 2a, b, 2c): the same quartic in x = s - cos t0
//   dihedral  s = cos(phi),   U = e0 + e1 s + ... + e5 s^5 (the reference's cubic, or the
//             Ryckaert-Bellemans quintic when the coefficients are given that way), phi the angle
//             between the planes (i, j, k) and (j, k, l), 0 for cis.

// dU/dx of the quartic well: x (2c + 2a (b - x)(b - 2x))
__device__ __forceinline__ float well_slope(float x, float4 c) {
  const float bx = c.z - x;
  return x * (c.w + c.y * bx * (bx - x));
}
__device__ __forceinline__ double well_energy64(double x, float4 c) {
  const double bx = double(c.z) - x;
  return 0.5 * x * x * (double(c.y) * bx * bx + double(c.w));
}
__device__ __forceinline__ double well_slope64(double x, float4 c) {
  const double bx = double(c.z) - x;
  return x * (double(c.w) + double(c.y) * bx * (bx - x));
}

struct V3 {
  float x, y, z;
};
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 operator*(float s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ float dot3(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 cross3(V3 a, V3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

struct D3 {
  double x, y, z;
};
__device__ __forceinline__ D3 to_d3(V3 a) { return {double(a.x), double(a.y), double(a.z)}; }
__device__ __forceinline__ D3 sub64(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ double dot64(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ D3 cross64(D3 a, D3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// cos(phi) of the dihedral i-j-k-l from the plane normals m = (j - i) x (k - j) and
// n = (k - j) x (l - k), and the gradient of cos(phi) with respect to atom `slot` (0..3):
//   P = d cos / d m = n / (|m||n|) - cos m / |m|^2,   Q = d cos / d n = m / (|m||n|) - cos n / |n|^2,
//   d cos / d x_i = -(b x P),  d cos / d x_j = b x P - P x a - c x Q,
//   d cos / d x_k = P x a + c x Q - Q x b,  d cos / d x_l = Q x b
// with a = x_j - x_i, b = x_k - x_j, c = x_l - x_k.
__device__ __forceinline__ float dihedral_cos_grad(V3 xi, V3 xj, V3 xk, V3 xl, int slot, V3* grad) {
  const V3 a = xj - xi, b = xk - xj, c = xl - xk;
  const V3 m = cross3(a, b), n = cross3(b, c);
  const float mm = dot3(m, m), nn = dot3(n, n);
  const float rmn = 1.0f / sqrtf(mm * nn);
  const float cs = dot3(m, n) * rmn;
  const V3 P = rmn * n - (cs / mm) * m;
  const V3 Q = rmn * m - (cs / nn) * n;
  const V3 bP = cross3(b, P), Pa = cross3(P, a), cQ = cross3(c, Q), Qb = cross3(Q, b);
  // slot weights of the four products: i (-1,0,0,0), j (1,-1,-1,0), k (0,1,1,-1), l (0,0,0,1)
  const float wb = slot == 0 ? -1.0f : (slot == 1 ? 1.0f : 0.0f);
  const float wa = slot == 1 ? -1.0f : (slot == 2 ? 1.0f : 0.0f);
  const float wq = slot == 3 ? 1.0f : (slot == 2 ? -1.0f : 0.0f);
  *grad = wb * bP + wa * Pa + wa * cQ + wq * Qb;
  return cs;
}

// Argument block of the bonded pass: per kind, the membership table and the terms.
struct BondedArgs {
  const int* start[3];
  const int* ref[3];
  const int* type[3];
  const int* atoms[3];
  const float* coef[3];
};

template <bool EV>
__global__ void __launch_bounds__(64) k_bonded(int n, BondedArgs A, const int4* __restrict__ info,
                                               const int* __restrict__ inv, const float4* __restrict__ xq,
                                               const int* __restrict__ xcode, const BoxPar* __restrict__ boxes,
                                               const int* __restrict__ rbox, float4* __restrict__ f, double* pe_b,
                                               double* pw_b, double* pe_a, double* pw_a, double* pe_d, double* pw_d) {
  const int i = xcd_block(gridDim.x) * 64 + (int)threadIdx.x;
  double eb = 0, wb = 0, ea = 0, ed = 0;
  if (i < n) {
    const int aid = info[i].x;
    const BoxPar& B = boxes[rbox[i]];
    // unwrapped molecule coordinates (x_mol): every atom of a term is in this atom's molecule
    auto mol_pos = [&](int a) -> V3 {
      const int r = inv[a];
      const float4 p = image_pos(xq[r], xcode[r], B.bl, B.m_bl);
      return {p.x, p.y, p.z};
    };
    V3 g = {0.0f, 0.0f, 0.0f};
    if (A.start[0]) {
      for (int k = A.start[0][aid], ke = A.start[0][aid + 1]; k < ke; ++k) {
        const int ref = A.ref[0][k], t = ref >> 2, slot = ref & 3;
        const float4 c = reinterpret_cast<const float4*>(A.coef[0])[A.type[0][t]];
        const V3 p0 = mol_pos(A.atoms[0][2 * t]), p1 = mol_pos(A.atoms[0][2 * t + 1]);
        const V3 d = p0 - p1;  // ds/dx_0 = d / r, ds/dx_1 = -d / r
        const float r = sqrtf(dot3(d, d));
        const float k1 = well_slope(r - c.x, c) / r;
        g = g + (slot == 0 ? -k1 : k1) * d;
        if (EV && slot == 0) {
          const D3 dd = sub64(to_d3(p0), to_d3(p1));
          const double r2 = sqrt(dot64(dd, dd));
          eb += well_energy64(r2 - double(c.x), c);
          wb -= well_slope64(r2 - double(c.x), c) * r2;
        }
      }
    }
    if (A.start[1]) {
      for (int k = A.start[1][aid], ke = A.start[1][aid + 1]; k < ke; ++k) {
        const int ref = A.ref[1][k], t = ref >> 2, slot = ref & 3;
        const float4 c = reinterpret_cast<const float4*>(A.coef[1])[A.type[1][t]];
        const int* at = A.atoms[1] + 3 * t;  // end, vertex, end
        const V3 pv = mol_pos(at[1]);
        const V3 u = mol_pos(at[0]) - pv, v = mol_pos(at[2]) - pv;
        const float iu = 1.0f / sqrtf(dot3(u, u)), iv = 1.0f / sqrtf(dot3(v, v));
        const V3 uh = iu * u, vh = iv * v;
        const float cs = dot3(uh, vh);
        const V3 g0 = iu * (vh - cs * uh), g2 = iv * (uh - cs * vh);  // d cos / d x_end
        const V3 gs = slot == 0 ? g0 : (slot == 2 ? g2 : -1.0f * (g0 + g2));
        g = g + (-well_slope(cs - c.x, c)) * gs;
        if (EV && slot == 0) {
          const D3 pv2 = to_d3(pv);
          const D3 u2 = sub64(to_d3(mol_pos(at[0])), pv2), v2 = sub64(to_d3(mol_pos(at[2])), pv2);
          const double c2 = dot64(u2, v2) / sqrt(dot64(u2, u2) * dot64(v2, v2));
          ea += well_energy64(c2 - double(c.x), c);
        }
      }
    }
    if (A.start[2]) {
      for (int k = A.start[2][aid], ke = A.start[2][aid + 1]; k < ke; ++k) {
        const int ref = A.ref[2][k], t = ref >> 2, slot = ref & 3;
        const float4* c4 = reinterpret_cast<const float4*>(A.coef[2]) + 3 * A.type[2][t];
        const float4 c0 = c4[0], c1 = c4[1], c2 = c4[2];  // (d0..d3), (d4, e0, e1, e2), (e3, e4, e5, -)
        const int* at = A.atoms[2] + 4 * t;
        const V3 xi = mol_pos(at[0]), xj = mol_pos(at[1]), xk = mol_pos(at[2]), xl = mol_pos(at[3]);
        V3 gs;
        const float cs = dihedral_cos_grad(xi, xj, xk, xl, slot, &gs);
        const float dU = (((c1.x * cs + c0.w) * cs + c0.z) * cs + c0.y) * cs + c0.x;
        g = g + (-dU) * gs;
        if (EV && slot == 0) {
          const D3 a = sub64(to_d3(xj), to_d3(xi)), b = sub64(to_d3(xk), to_d3(xj)), cc = sub64(to_d3(xl), to_d3(xk));
          const D3 m = cross64(a, b), nv = cross64(b, cc);
          const double s = dot64(m, nv) / sqrt(dot64(m, m) * dot64(nv, nv));
          ed += ((((double(c2.z) * s + double(c2.y)) * s + double(c2.x)) * s + double(c1.w)) * s + double(c1.z)) * s +
                double(c1.y);
        }
      }
    }
    float4 fo = f[i];
    fo.x += g.x; fo.y += g.y; fo.z += g.z;
    f[i] = fo;
  }
  if (EV) {
    block_sum2_store<64>(eb, wb, pe_b, pw_b);
    __syncthreads();
    block_sum2_store<64>(ea, 0.0, pe_a, pw_a);
    __syncthreads();
    block_sum2_store<64>(ed, 0.0, pe_d, pw_d);
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
  const int N = ctx->natom, nb1 = ctx->pt_nbin + 1, nblk = blocks_for((long long)N * kPairG, kPairBlock);
  PairEv64 P64;
  P64.cut = double(ctx->pt_cut);
  P64.rdr = double(ctx->pt_nbin) / P64.cut;
  P64.vc = -1.0 / P64.cut;
  P64.fc = P64.vc / P64.cut;
  const size_t lds = size_t(ctx->npair) * nb1 * sizeof(float);
  double* pe = ev_part(ctx, NNPE_TERM_PAIR, 0);
  double* pw = ev_part(ctx, NNPE_TERM_PAIR, 1);
#define NNPE_PAIR(EV, TYE)                                                                                           \
  hipLaunchKernelGGL((k_pair<EV, kPairG, TYE>), dim3(nblk), dim3(kPairBlock), lds, ctx->stream, N, ctx->cap_long, ctx->nbr_long, \
                     ctx->cnt_long, ctx->xq, ctx->typ, ctx->d_box, ctx->rbox, ctx->ntype, ctx->npair, ctx->pt_cutsq,         \
                     ctx->pt_rdr, ctx->pt_fc, P64, nb1, ctx->pt_tab, ctx->f, pe, pw)
  const bool tye = ctx->ntype <= NNPE_ENTRY_MAX_TYPES;
  if (ev) {
    if (tye) NNPE_PAIR(true, true); else NNPE_PAIR(true, false);
  } else {
    if (tye) NNPE_PAIR(false, true); else NNPE_PAIR(false, false);
  }
#undef NNPE_PAIR
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

// kind 0 bonds, 1 angles, 2 dihedrals in the reference's cubic form, 3 dihedrals as an energy
// polynomial of degree 5 in cos(phi) (Ryckaert-Bellemans); 3 replaces the dihedral terms.
extern "C" int nnpe_bonded_set(nnpe_ctx* ctx, int kind, int ntyp, const float* coef, int nterm, const int* terms) {
  if (ctx) md_state_free(ctx);  // captured step graphs hold the old parameters
  if (!ctx || kind < 0 || kind > 3 || ntyp < 0 || nterm < 0 || (nterm && (!coef || !terms)) || (ntyp && !coef))
    return ctx_fail(ctx, NNPE_ERR_ARG, "nnpe_bonded_set: bad argument");
  if (!ctx->have_system) return ctx_fail(ctx, NNPE_ERR_STATE, "no system");
  const int slot = kind == 3 ? 2 : kind;  // storage slot
  const int na = 2 + slot;                // atoms per term
  const int N = ctx->natom;
  // every atom of a term must belong to one molecule (the unwrapped molecule coordinates)
  std::vector<int> hmol(N);
  if (nterm) {
    std::vector<int4> hinfo(N);
    NNPE_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    NNPE_CHECK_HIP(ctx, hipMemcpy(hinfo.data(), ctx->info, sizeof(int4) * N, hipMemcpyDeviceToHost));
    for (auto& in : hinfo) hmol[in.x] = in.y;
  }
  std::vector<int> type(std::max(nterm, 1)), atoms(size_t(na) * std::max(nterm, 1)), start(N + 1, 0);
  for (int t = 0; t < nterm; ++t) {
    const int* p = terms + (na + 1) * t;
    if (p[0] < 0 || p[0] >= ntyp) return ctx_fail(ctx, NNPE_ERR_ARG, "auto generate error: term type out of range");
    for (int k = 1; k <= na; ++k)
      if (p[k] < 0 || p[k] >= N || hmol[p[k]] != hmol[p[1]]) return ctx_fail(ctx, NNPE_ERR_ARG, "term atoms outside one molecule");
    type[t] = p[0];
    for (int k = 0; k < na; ++k) {
      atoms[size_t(na) * t + k] = p[1 + k];
      start[p[1 + k] + 1]++;
    }
  }
  for (int a = 0; a < N; ++a) start[a + 1] += start[a];
  std::vector<int> ref(std::max(start[N], 1)), fill(start.begin(), start.end() - 1);
  for (int t = 0; t < nterm; ++t)
    for (int k = 0; k < na; ++k) ref[fill[atoms[size_t(na) * t + k]]++] = (t << 2) | k;
  // coefficients in the kernel's form
  const int cw = slot == 2 ? 12 : 4;
  std::vector<float> c(size_t(cw) * std::max(ntyp, 1), 0.0f);
  for (int y = 0; y < ntyp; ++y) {
    float* o = c.data() + size_t(cw) * y;
    if (slot < 2) {
      for (int k = 0; k < 4; ++k) o[k] = coef[4 * y + k];
    } else if (kind == 2) {
      // stored (k0, k1, k2, k3) (src/dihedral.cpp:27-35): U = k0 + k1 s + k2/2 s^2 + k3/3 s^3
      const float* k = coef + 4 * y;
      const float d[5] = {k[1], k[2], k[3], 0.0f, 0.0f}, e[6] = {k[0], k[1], k[2] / 2.0f, k[3] / 3.0f, 0.0f, 0.0f};
      for (int q = 0; q < 5; ++q) o[q] = d[q];
      for (int q = 0; q < 6; ++q) o[5 + q] = e[q];
    } else {
      // U = C0 + C1 s + ... + C5 s^5
      const float* e = coef + 6 * y;
      for (int q = 0; q < 5; ++q) o[q] = float(q + 1) * e[q + 1];
      for (int q = 0; q < 6; ++q) o[5 + q] = e[q];
    }
  }
  NNPE_CHECK_HIP(ctx, dev_alloc(ctx->bd_coef[slot], c.size()));
  NNPE_CHECK_HIP(ctx, hipMemcpy(ctx->bd_coef[slot], c.data(), c.size() * sizeof(float), hipMemcpyHostToDevice));
  if (nterm) {
    NNPE_CHECK_HIP(ctx, dev_alloc(ctx->bd_type[slot], type.size()));
    NNPE_CHECK_HIP(ctx, hipMemcpy(ctx->bd_type[slot], type.data(), type.size() * sizeof(int), hipMemcpyHostToDevice));
    NNPE_CHECK_HIP(ctx, dev_alloc(ctx->bd_atoms[slot], atoms.size()));
    NNPE_CHECK_HIP(ctx, hipMemcpy(ctx->bd_atoms[slot], atoms.data(), atoms.size() * sizeof(int), hipMemcpyHostToDevice));
    NNPE_CHECK_HIP(ctx, dev_alloc(ctx->bd_start[slot], start.size()));
    NNPE_CHECK_HIP(ctx, hipMemcpy(ctx->bd_start[slot], start.data(), start.size() * sizeof(int), hipMemcpyHostToDevice));
    NNPE_CHECK_HIP(ctx, dev_alloc(ctx->bd_ref[slot], ref.size()));
    NNPE_CHECK_HIP(ctx, hipMemcpy(ctx->bd_ref[slot], ref.data(), ref.size() * sizeof(int), hipMemcpyHostToDevice));
    ctx->bd_start_len[slot] = N + 1;
  } else {
    dev_free(ctx->bd_type[slot]); dev_free(ctx->bd_atoms[slot]); dev_free(ctx->bd_start[slot]); dev_free(ctx->bd_ref[slot]);
    ctx->bd_start_len[slot] = 0;
  }
  ctx->bd_ntyp[slot] = ntyp;
  ctx->bd_nterm[slot] = nterm;
  return NNPE_OK;
}

int bonded_enqueue(nnpe_ctx* ctx, bool ev) {
  BondedArgs A;
  int total = 0;
  for (int k = 0; k < 3; ++k) {
    const bool on = ctx->bd_nterm[k] > 0;
    if (on && ctx->bd_start_len[k] != ctx->natom + 1)
      return ctx_fail(ctx, NNPE_ERR_STATE, "bonded terms were set for another atom count");
    A.start[k] = on ? ctx->bd_start[k] : nullptr;
    A.ref[k] = ctx->bd_ref[k];
    A.type[k] = ctx->bd_type[k];
    A.atoms[k] = ctx->bd_atoms[k];
    A.coef[k] = ctx->bd_coef[k];
    total += ctx->bd_nterm[k];
  }
  if (total == 0) return NNPE_OK;
  const int nblk = blocks_for(ctx->natom, 64);
  NNPE_CHECK_HIP(ctx, ev_reserve(ctx, nblk));
  auto P = [&](int term, int w) { return ev ? ev_part(ctx, term, w) : nullptr; };
  if (ev)
    hipLaunchKernelGGL(k_bonded<true>, dim3(nblk), dim3(64), 0, ctx->stream, ctx->natom, A, ctx->info, ctx->inv, ctx->xq,
                       ctx->xmol_code, ctx->d_box, ctx->rbox, ctx->f, P(NNPE_TERM_BOND, 0), P(NNPE_TERM_BOND, 1),
                       P(NNPE_TERM_ANGLE, 0), P(NNPE_TERM_ANGLE, 1), P(NNPE_TERM_DIHEDRAL, 0), P(NNPE_TERM_DIHEDRAL, 1));
  else
    hipLaunchKernelGGL(k_bonded<false>, dim3(nblk), dim3(64), 0, ctx->stream, ctx->natom, A, ctx->info, ctx->inv, ctx->xq,
                       ctx->xmol_code, ctx->d_box, ctx->rbox, ctx->f, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
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

// ---- the bonded pass without allocation ---------------------------------------------------------
// The per-workgroup energy/virial partials of the three bonded terms live in one file-scope
// device array instead of the context's partial buffer (which nnpe_bonded_compute may grow).
// Each device holds its own copy at its own address, found once per device when a context is
// created on it.  Every call on a device shares that device's array: calls that overlap on one
// device race, and each call overwrites every partial it reads.
namespace {
constexpr int kScratchBlocks = 1 << 14;  // 64-atom workgroups: at most 1,048,576 atoms
constexpr int kScratchDevices = 64;
__device__ __align__(256) double scratch[6 * kScratchBlocks];  // [bond e, bond w, angle e, angle w, dihedral e, dihedral w]
std::mutex g_scratch_mu;
double* g_scratch_ptr[kScratchDevices] = {};  // device address of scratch, per device
}  // namespace

int bonded_scratch_setup(nnpe_ctx* ctx) {
  if (ctx->device < 0 || ctx->device >= kScratchDevices) return NNPE_OK;  // the entry point refuses such devices
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[ctx->device]) return NNPE_OK;
  void* p = nullptr;
  NNPE_CHECK_HIP(ctx, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));  // with the context's device current
  g_scratch_ptr[ctx->device] = static_cast<double*>(p);
  return NNPE_OK;
}

extern "C" int nnpe_bonded_compute_noalloc(nnpe_ctx* ctx, int ev) {
  if (!ctx) return NNPE_ERR_ARG;
  if (!ctx->have_list) return ctx_fail(ctx, NNPE_ERR_STATE, "bonded compute needs a list build (unwrapped images)");
  double* sc = nullptr;
  if (ctx->device >= 0 && ctx->device < kScratchDevices) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    sc = g_scratch_ptr[ctx->device];
  }
  if (!sc) return ctx_fail(ctx, NNPE_ERR_UNSUPPORTED, "no scratch array on this device");
  const int nblk = blocks_for(ctx->natom, 64);
  if (nblk > kScratchBlocks)
    return ctx_fail(ctx, NNPE_ERR_UNSUPPORTED, "nnpe_bonded_compute_noalloc: more than 1048576 atoms");
  BondedArgs A;
  int total = 0;
  for (int k = 0; k < 3; ++k) {
    const bool on = ctx->bd_nterm[k] > 0;
    if (on && ctx->bd_start_len[k] != ctx->natom + 1)
      return ctx_fail(ctx, NNPE_ERR_STATE, "bonded terms were set for another atom count");
    A.start[k] = on ? ctx->bd_start[k] : nullptr;
    A.ref[k] = ctx->bd_ref[k];
    A.type[k] = ctx->bd_type[k];
    A.atoms[k] = ctx->bd_atoms[k];
    A.coef[k] = ctx->bd_coef[k];
    total += ctx->bd_nterm[k];
  }
  if (total == 0) return NNPE_OK;
  auto P = [&](int term, int w) { return sc + (size_t)(2 * (term - NNPE_TERM_BOND) + w) * kScratchBlocks; };
  if (ev)
    hipLaunchKernelGGL(k_bonded<true>, dim3(nblk), dim3(64), 0, ctx->stream, ctx->natom, A, ctx->info, ctx->inv, ctx->xq,
                       ctx->xmol_code, ctx->d_box, ctx->rbox, ctx->f, P(NNPE_TERM_BOND, 0), P(NNPE_TERM_BOND, 1),
                       P(NNPE_TERM_ANGLE, 0), P(NNPE_TERM_ANGLE, 1), P(NNPE_TERM_DIHEDRAL, 0), P(NNPE_TERM_DIHEDRAL, 1));
  else
    hipLaunchKernelGGL(k_bonded<false>, dim3(nblk), dim3(64), 0, ctx->stream, ctx->natom, A, ctx->info, ctx->inv, ctx->xq,
                       ctx->xmol_code, ctx->d_box, ctx->rbox, ctx->f, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
  NNPE_CHECK_LAUNCH(ctx);
  if (!ev) return NNPE_OK;
  int rc = NNPE_OK;
  for (int t = NNPE_TERM_BOND; t <= NNPE_TERM_DIHEDRAL && !rc; ++t) rc = ev_finish_from(ctx, t, P(t, 0), P(t, 1), nblk);
  return rc;
}
