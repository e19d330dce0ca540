// neigh.hip -- cell grid, cell sort, ghost images and the full Verlet neighbour lists.
//
// Follows the reference's list construction (src/environ.cpp:143-155, :234-522,
// src/neighlist.cpp:35-331) with the same FP32 arithmetic so that the set of interacting
// (atom, atom, image) triples is bit-identical:
//   refold x -= rint(x/L)*L, cell = int((offset + x) * 1/dx), ghosts for atoms within the comm
//   cutoff of a face with x_ghost = x + (+-L), the per-cell stencil with "no distance check"
//   cells, and the test (xi-xj)^2+(yi-yj)^2+(zi-zj)^2 < (rc+skin)^2 without contraction.
// The GPU build differs in mechanics only: counting sorts with atomics followed by a per-cell
// sort on the atom id (deterministic order), ghosts as (source row, image code) records, and a
// full list in padded rows (entry k of atom i at nbr[i*cap + k]) written by wavefront-ballot
// compaction, instead of per-thread half lists.
#include <algorithm>
#include <cmath>

#include "internal.h"

namespace {

constexpr int kStencilNoCheck = 1;

// the rebuild's counters and per-cell counts cleared in one launch, the g(r) counts of a
// tally folded into the rebuild (gr_n == 0 otherwise), the fill counters of the per-element row
// index (nt_fill == 0 when there is none), and, after an MD step's skin check, its flag added to
// the skin-violation count and cleared
__global__ void k_build_reset(int G, int* cell_cnt, int* cell_fill, int* gcell_cnt, int* gcell_fill, DevCounters* c,
                              int gr_n, unsigned long long* gr_count, int gr_shell_n, unsigned long long* gr_shell,
                              int nt_fill, int* type_fill, int skin_accum) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < G) { cell_cnt[i] = 0; cell_fill[i] = 0; gcell_cnt[i] = 0; gcell_fill[i] = 0; }
  if (i < gr_n) gr_count[i] = 0ull;
  if (i < gr_shell_n) gr_shell[i] = 0ull;
  if (i < nt_fill) type_fill[i] = 0;
  if (i == 0) {
    c->nghost = 0; c->overflow_long = 0; c->overflow_short = 0; c->ghost_overflow = 0;
    c->cell_range = 0; c->max_long = 0; c->max_short = 0;
    if (skin_accum) {
      if (c->invalid) c->skin_violations += 1;
      c->invalid = 0;
    }
  }
}

__device__ __forceinline__ int cell_coord(float g_offset, float x, float g_rdx) {
  return int(__fmul_rn(__fadd_rn(g_offset, x), g_rdx));
}

// LOCAL::refold_and_tag (src/environ.cpp:234-262): wrap into the row's box and bin into its grid.
__global__ void k_refold(int n, bool fp64, const BoxPar* __restrict__ boxes, const int* __restrict__ rbox, float4* xq,
                         double4* xd, int* cell_out, int* cell_cnt, DevCounters* cnt) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const BoxPar B = boxes[rbox[i]];
  const float bl = B.bl, r_bl = B.r_bl;
  const int num_gx = B.num_gx;
  float4 p = xq[i];
  if (fp64) {
    double4 d = xd[i];
    double dbl = bl, drbl = 1.0 / double(bl);
    d.x -= rint(d.x * drbl) * dbl; d.y -= rint(d.y * drbl) * dbl; d.z -= rint(d.z * drbl) * dbl;
    xd[i] = d;
    p.x = float(d.x); p.y = float(d.y); p.z = float(d.z);
  } else {
    p.x = __fsub_rn(p.x, __fmul_rn(rintf(__fmul_rn(p.x, r_bl)), bl));
    p.y = __fsub_rn(p.y, __fmul_rn(rintf(__fmul_rn(p.y, r_bl)), bl));
    p.z = __fsub_rn(p.z, __fmul_rn(rintf(__fmul_rn(p.z, r_bl)), bl));
  }
  xq[i] = p;
  int cx = cell_coord(B.g_offset, p.x, B.g_rdx), cy = cell_coord(B.g_offset, p.y, B.g_rdx),
      cz = cell_coord(B.g_offset, p.z, B.g_rdx);
  // outside the box's home cells (which the list kernel covers) is an error like outside the grid
  if (cx < B.home_lo || cy < B.home_lo || cz < B.home_lo || cx > B.home_hi || cy > B.home_hi || cz > B.home_hi) {
    atomicAdd(&cnt->cell_range, 1);
    cx = min(max(cx, 0), num_gx - 1); cy = min(max(cy, 0), num_gx - 1); cz = min(max(cz, 0), num_gx - 1);
  }
  int c = B.cell_base + (cx * num_gx + cy) * num_gx + cz;
  cell_out[i] = c;
  atomicAdd(&cell_cnt[c], 1);
}

// exclusive scan of n counts into start[0..n] with one 1024-thread block
__global__ void __launch_bounds__(1024) k_scan_excl(const int* cnt, int* start, int n) {
  __shared__ int s[1024];
  int t = threadIdx.x;
  int chunk = (n + 1023) / 1024;
  int b = min(t * chunk, n), e = min(b + chunk, n);
  int sum = 0;
  for (int k = b; k < e; ++k) sum += cnt[k];
  s[t] = sum;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    int v = t >= off ? s[t - off] : 0;
    __syncthreads();
    s[t] += v;
    __syncthreads();
  }
  int run = s[t] - sum;
  for (int k = b; k < e; ++k) { start[k] = run; run += cnt[k]; }
  if (t == 1023) start[n] = s[1023];
}

__global__ void k_scatter(int n, const int* cell_of, const int* start, int* fill, int* order) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int c = cell_of[i];
  order[start[c] + atomicAdd(&fill[c], 1)] = i;
}

// deterministic order inside a cell: ascending atom id.  One wave per cell: a cell of up to 64
// entries is ranked in registers (each lane counts the smaller keys of the others through
// shuffles, keys are unique) and scattered to its rank; longer cells (rare) fall back to an
// insertion sort by lane 0.
__global__ void __launch_bounds__(256) k_cell_sort(int ncell, const int* start, int* order, const int4* info) {
  const int c = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, lane = threadIdx.x & 63;
  if (c >= ncell) return;  // whole wave
  const int b = start[c], n = start[c + 1] - b;
  if (n <= 1) return;
  if (n <= 64) {
    const int v = lane < n ? order[b + lane] : 0;
    const int key = lane < n ? info[v].x : 0x7FFFFFFF;
    int rank = 0;
    for (int k = 0; k < n; ++k) rank += __shfl(key, k, 64) < key;
    if (lane < n) order[b + rank] = v;
    return;
  }
  if (lane != 0) return;
  for (int k = b + 1; k < b + n; ++k) {
    int v = order[k], key = info[v].x, m = k - 1;
    while (m >= b && info[order[m]].x > key) { order[m + 1] = order[m]; --m; }
    order[m + 1] = v;
  }
}

// ENVIRON::sort data movement (src/environ.cpp:319-361): rows permuted into cell order
__global__ void k_gather(int n, bool fp64, const int* order, const int* cell_old, const float4* xq, const float4* v,
                         const float4* fprev, const float* rm, const int* typ, const int4* info, const double4* xd,
                         const double4* vd, const double4* ad, const int* rbox, float4* xq2, float4* v2, float4* fprev2,
                         float* rm2, int* typ2, int4* info2, double4* xd2, double4* vd2, double4* ad2, int* rbox2,
                         int* cell_new, int* inv) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int o = order[i];
  xq2[i] = xq[o]; v2[i] = v[o]; fprev2[i] = fprev[o]; rm2[i] = rm[o]; typ2[i] = typ[o]; rbox2[i] = rbox[o];
  int4 in = info[o];
  info2[i] = in;
  cell_new[i] = cell_old[o];
  inv[in.x] = i;
  if (fp64) { xd2[i] = xd[o]; vd2[i] = vd[o]; ad2[i] = ad[o]; }
}

// ... and, when the network keeps a row index per element (type_idx), the sorted rows of each
// element appended to it, wave-aggregated: one atomic per wave and element (fill counters
// cleared by the rebuild's reset)
__device__ __forceinline__ void copy_back_row(int i, int n, bool fp64, const float4* xq2, const float4* v2,
                                              const float4* fprev2, const float* rm2, const int* typ2, const int4* info2,
                                              const double4* xd2, const double4* vd2, const double4* ad2,
                                              const int* rbox2, float4* xq, float4* v, float4* fprev, float* rm, int* typ,
                                              int4* info, double4* xd, double4* vd, double4* ad, int* rbox, int ntype,
                                              const int* type_off, int* type_fill, int* type_idx) {
  const int t = i < n ? typ2[i] : -1;
  if (i < n) {
    xq[i] = xq2[i]; v[i] = v2[i]; fprev[i] = fprev2[i]; rm[i] = rm2[i]; typ[i] = t; info[i] = info2[i];
    rbox[i] = rbox2[i];
    if (fp64) { xd[i] = xd2[i]; vd[i] = vd2[i]; ad[i] = ad2[i]; }
  }
  if (!type_idx) return;  // uniform
  const int lane = threadIdx.x & 63;
  for (int e = 0; e < ntype; ++e) {
    const unsigned long long bal = __ballot(t == e);
    if (!bal) continue;
    const int first = __ffsll((long long)bal) - 1;
    int base = 0;
    if (lane == first) base = atomicAdd(&type_fill[e], __popcll(bal));
    base = __shfl(base, first, 64);
    if (t == e) {
      const int rank = __builtin_amdgcn_mbcnt_hi((unsigned)(bal >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)bal, 0));
      type_idx[type_off[e] + base + rank] = i;
    }
  }
}

// LOCAL::update_ghost_and_info x_mol part (src/environ.cpp:461-469): the image of an atom whose
// every coordinate lies within L/2 of its molecule's intra-0 atom, among the images the atom
// has ghosts for; the atom itself otherwise.
__device__ int xmol_code_of(int i, const BoxPar& B, const float4* xq, const int4* info, const int* mol_first,
                            const int* inv) {
  const float ghost_lim = B.ghost_lim, bl = B.bl, m_bl = B.m_bl, h_bl = B.h_bl;
  int4 in = info[i];
  int code_sel = 13;
  if (in.z != 0) {
    float4 p = xq[i];
    float4 r = xq[inv[mol_first[in.y]]];
    const float mg = -ghost_lim;
    int xs = p.x > ghost_lim ? -1 : 0, xe = p.x < mg ? 1 : 0;
    int ys = p.y > ghost_lim ? -1 : 0, ye = p.y < mg ? 1 : 0;
    int zs = p.z > ghost_lim ? -1 : 0, ze = p.z < mg ? 1 : 0;
    for (int a = xs; a <= xe; ++a)
      for (int b = ys; b <= ye; ++b)
        for (int c = zs; c <= ze; ++c) {
          if (a == 0 && b == 0 && c == 0) continue;
          int code = (a + 1) * 9 + (b + 1) * 3 + (c + 1);
          float4 g = image_pos(p, code, bl, m_bl);
          if (fabsf(__fsub_rn(g.x, r.x)) < h_bl && fabsf(__fsub_rn(g.y, r.y)) < h_bl && fabsf(__fsub_rn(g.z, r.z)) < h_bl)
            code_sel = code;
        }
  }
  return code_sel;
}

// LOCAL::generate_ghost + update_ghost_pos<true> (src/environ.cpp:401-442, :488-514)
__device__ __forceinline__ void ghost_gen_row(int i, int n, const BoxPar* __restrict__ boxes, const int* __restrict__ rbox,
                                              int cap, const float4* xq, int* g_src, int* g_code, int* g_cell,
                                              int* gcell_cnt, DevCounters* cnt, const int4* __restrict__ info,
                                              const int* __restrict__ mol_first, const int* __restrict__ inv,
                                              int* __restrict__ xmol_code) {
  const int lane = threadIdx.x & 63;
  BoxPar B{};
  float4 p = make_float4(0, 0, 0, 0);
  int xs = 0, xe = 0, ys = 0, ye = 0, zs = 0, ze = 0;
  if (i < n) {
    B = boxes[rbox[i]];
    p = xq[i];
    const float ghost_lim = B.ghost_lim, mg = -ghost_lim;
    xs = p.x > ghost_lim ? -1 : 0; xe = p.x < mg ? 1 : 0;
    ys = p.y > ghost_lim ? -1 : 0; ye = p.y < mg ? 1 : 0;
    zs = p.z > ghost_lim ? -1 : 0; ze = p.z < mg ? 1 : 0;
  }
  const int nimg = (xe - xs + 1) * (ye - ys + 1) * (ze - zs + 1) - 1;  // 0 for no image
  // slots: one atomic per wave, each lane's images at its exclusive prefix (the cell sort
  // orders the ghosts afterwards, so the slot order does not matter)
  int inc = nimg;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int t = __shfl_up(inc, o, 64);
    if (lane >= o) inc += t;
  }
  const int wtot = __shfl(inc, 63, 64);
  int base = 0;
  if (lane == 0 && wtot > 0) base = atomicAdd(&cnt->nghost, wtot);
  base = __shfl(base, 0, 64);
  if (i < n) xmol_code[i] = xmol_code_of(i, B, xq, info, mol_first, inv);
  if (nimg == 0) return;
  int slot = base + inc - nimg;
  const float bl = B.bl, m_bl = B.m_bl, g_offset = B.g_offset, g_rdx = B.g_rdx;
  const int num_gx = B.num_gx;
  for (int a = xs; a <= xe; ++a)
    for (int b = ys; b <= ye; ++b)
      for (int c = zs; c <= ze; ++c) {
        if (a == 0 && b == 0 && c == 0) continue;
        const int code = (a + 1) * 9 + (b + 1) * 3 + (c + 1);
        const int sl = slot++;
        if (sl >= cap) { atomicAdd(&cnt->ghost_overflow, 1); continue; }
        float4 g = image_pos(p, code, bl, m_bl);
        int cx = cell_coord(g_offset, g.x, g_rdx), cy = cell_coord(g_offset, g.y, g_rdx), cz = cell_coord(g_offset, g.z, g_rdx);
        if (cx < 0 || cy < 0 || cz < 0 || cx >= num_gx || cy >= num_gx || cz >= num_gx) {
          atomicAdd(&cnt->cell_range, 1);
          cx = min(max(cx, 0), num_gx - 1); cy = min(max(cy, 0), num_gx - 1); cz = min(max(cz, 0), num_gx - 1);
        }
        const int gc = B.cell_base + (cx * num_gx + cy) * num_gx + cz;
        g_src[sl] = i; g_code[sl] = code; g_cell[sl] = gc;
        atomicAdd(&gcell_cnt[gc], 1);
      }
}

// The gathered rows copied back and the ghost images of each row generated in one pass: the
// ghost pass reads the gathered copies (the same values), so it need not wait for the copy.
__global__ void k_copy_back_ghosts(int n, bool fp64, const float4* xq2, const float4* v2, const float4* fprev2,
                                   const float* rm2, const int* typ2, const int4* info2, const double4* xd2,
                                   const double4* vd2, const double4* ad2, const int* rbox2, float4* xq, float4* v,
                                   float4* fprev, float* rm, int* typ, int4* info, double4* xd, double4* vd, double4* ad,
                                   int* rbox, int ntype, const int* type_off, int* type_fill, int* type_idx,
                                   const BoxPar* __restrict__ boxes, int cap, int* g_src, int* g_code, int* g_cell,
                                   int* gcell_cnt, DevCounters* cnt, const int* __restrict__ mol_first,
                                   const int* __restrict__ inv, int* __restrict__ xmol_code) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  copy_back_row(i, n, fp64, xq2, v2, fprev2, rm2, typ2, info2, xd2, vd2, ad2, rbox2, xq, v, fprev, rm, typ, info, xd, vd, ad,
                rbox, ntype, type_off, type_fill, type_idx);
  ghost_gen_row(i, n, boxes, rbox2, cap, xq2, g_src, g_code, g_cell, gcell_cnt, cnt, info2, mol_first, inv, xmol_code);
}

__global__ void k_ghost_scatter(int cap, const DevCounters* cnt, const int* g_src, const int* g_code, const int* g_cell,
                                const int* gstart, int* gfill, int* gs_src, int* gs_code) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  int ng = min(cnt->nghost, cap);
  if (s >= ng) return;
  int c = g_cell[s];
  int pos = gstart[c] + atomicAdd(&gfill[c], 1);
  gs_src[pos] = g_src[s];
  gs_code[pos] = g_code[s];
}

// ENVIRON::sort_ghost order made deterministic: ascending (source atom id, image code)
// ghosts of a cell in (atom id, image code) order, ranked as k_cell_sort
__global__ void __launch_bounds__(256) k_ghost_cell_sort(int ncell, const int* gstart, int* gs_src, int* gs_code,
                                                         const int4* info) {
  const int c = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, lane = threadIdx.x & 63;
  if (c >= ncell) return;  // whole wave
  const int b = gstart[c], n = gstart[c + 1] - b;
  if (n <= 1) return;
  if (n <= 64) {
    const int sv = lane < n ? gs_src[b + lane] : 0, cd = lane < n ? gs_code[b + lane] : 0;
    const long long key = lane < n ? (long long)info[sv].x * 32 + cd : 0x7FFFFFFFFFFFFFFFLL;
    int rank = 0;
    for (int k = 0; k < n; ++k) rank += __shfl(key, k, 64) < key;
    if (lane < n) { gs_src[b + rank] = sv; gs_code[b + rank] = cd; }
    return;
  }
  if (lane != 0) return;
  for (int k = b + 1; k < b + n; ++k) {
    int sv = gs_src[k], cd = gs_code[k];
    long long key = (long long)info[sv].x * 32 + cd;
    int m = k - 1;
    while (m >= b && (long long)info[gs_src[m]].x * 32 + gs_code[m] > key) {
      gs_src[m + 1] = gs_src[m]; gs_code[m + 1] = gs_code[m]; --m;
    }
    gs_src[m + 1] = sv; gs_code[m + 1] = cd;
  }
}


// NEIGHLIST::build (src/neighlist.cpp:190-331) as a full list: every real neighbour j != i and
// every ghost image within (rc+skin), or in a stencil cell flagged "no check", with the
// special class of the pair; optionally a second (short) list of the entries within a smaller
// cutoff.  One workgroup per home cell with wavefront-ballot compaction: all atoms of a cell
// share one stencil, so the workgroup stages the candidates (real rows and ghost images of the
// stencil cells, in stencil order) in LDS once per batch; each wave then takes one home atom at
// a time, its 64 lanes test 64 candidates per step, and a ballot + mbcnt gives every accepted
// lane its slot, so a wave appends up to 64 entries to the atom's row with one coalesced store.
// Entry order in a row: stencil cell, real before ghost, then position in the cell.
constexpr int kBuildThreads = 512;  // workgroup size of the list build
constexpr int kBuildBatch = 1152;  // candidates staged per pass (768 .. 2048 measured)
constexpr int kBuildHome = 64;     // home atoms per pass
constexpr int kBuildMaxRanges = 2 * 343;
constexpr int kShortMaxType = 4;   // short lists: one per neighbour element

template <bool kGr>
struct BuildRowArgs {
  float4 xi;
  int mol_i, h, tn;
  uint32_t home_ent;  // the home atom's own (real) candidate entry, the one candidate never taken
  float cutsq, cutsq_s;
  int ntype, stride, cap, cap_s;
  uint32_t* rowl;
  uint32_t* nbr_s;
  int ti_gr;
  float gr_rmaxsq, gr_rdr;
  int gr_nbin;
  unsigned int* s_gh;
};

// dword 3 of a raw (unswizzled, stride 0) buffer resource on CDNA: 32-bit data format
constexpr int kBufDword3 = 0x00020000;

// One home atom against the staged candidates: 64 candidates per step, ballot compaction into
// the full row and the per-element short rows.  kBigMol: molecules of more than 64 atoms read
// the special class from the molecule type's table; smaller ones take it from the lane masks
// (bit 0 of class in sp0, bit 1 in sp1, indexed by intra id).
// The rows are written through buffer resources sized to their capacity: the hardware range
// check drops an entry past the end (the count still grows and reports the overflow), so a
// store is a shift and the store itself.  Each element's short row has its own resource, and a
// lane stores under its element's exec mask, which keeps the per-lane row selection out of
// the step.  The special class is looked up only in steps where a lane sees its own molecule.
template <bool kShort, bool kGr, bool kBigMol, int NT>
__device__ __forceinline__ void build_row(const BuildRowArgs<kGr>& A, const float4* c_pos, const uint32_t* c_ent,
                                          const uint32_t* c_aux, const int* sp_row, unsigned long long sp0,
                                          unsigned long long sp1, int lane, int& nl, int (&ns)[kShortMaxType]) {
  const __amdgpu_buffer_rsrc_t rl = __builtin_amdgcn_make_buffer_rsrc(A.rowl, 0, A.cap * 4, kBufDword3);
  __amdgpu_buffer_rsrc_t rsh[NT];
#pragma unroll
  for (int el = 0; el < NT; ++el)
    rsh[el] = kShort ? __builtin_amdgcn_make_buffer_rsrc(A.nbr_s + ((size_t)el * A.stride + A.h) * A.cap_s, 0,
                                                         A.cap_s * 4, kBufDword3)
                     : rl;
  for (int u0 = 0; u0 < A.tn; u0 += 64) {
    // slots from tn up to the next multiple of 64 hold padding candidates placed far away (never
    // taken, never tallied), so the step has no bound test and no branch around its LDS reads
    const int u = u0 + lane;
    const uint32_t ce = c_ent[u];
    const float4 xj = c_pos[u];
    const uint32_t ax = c_aux[u];
    uint32_t e = ce & 0x7FFFFFFFu;
    const bool valid = e != A.home_ent;
    const bool nock = (int)ce < 0;
    const float rsq = dist2_rn(__fsub_rn(A.xi.x, xj.x), __fsub_rn(A.xi.y, xj.y), __fsub_rn(A.xi.z, xj.z));
    const bool take = valid && (nock || rsq < A.cutsq);
    const bool same_mol = __float_as_int(xj.w) == A.mol_i;
    if (__ballot(same_mol)) {
      const int jm = (int)(ax & 0xFFFFFFu);
      int sp = 0;
      if (kBigMol) {
        if (same_mol) sp = sp_row[jm];
      } else {
        // intra ids of the home molecule are < 64; other lanes' shifts are discarded
        sp = same_mol ? (int)(((sp0 >> (jm & 63)) & 1ull) | (((sp1 >> (jm & 63)) & 1ull) << 1)) : 0;
      }
      e |= (uint32_t)sp << 30;
    }
    // a real pair once, from its lower row, with weight 2; a ghost image from each side
    if (kGr) {
      const int j = NNPE_ENTRY_J(ce);
      const bool ghost = NNPE_ENTRY_CODE(ce) != 13;
      if (valid && (ghost || j > A.h) && rsq < A.gr_rmaxsq) {
        const int b = gr_bin_of(rsq, A.gr_rdr);
        if (b < A.gr_nbin) {
          const int tj = (int)(ax >> 24);
          const int a = min(A.ti_gr, tj), bb = max(A.ti_gr, tj);
          atomicAdd(&A.s_gh[(a * A.ntype - (a * (a - 1)) / 2 + (bb - a)) * A.gr_nbin + b], ghost ? 1u : 2u);
        }
      }
    }
    const unsigned long long bal = __ballot(take);
    if (take) {
      const unsigned rank = __builtin_amdgcn_mbcnt_hi((unsigned)(bal >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)bal, (unsigned)nl));
      __builtin_amdgcn_raw_buffer_store_b32(e, rl, (int)(rank * 4u), 0, 0);
    }
    nl += __popcll(bal);
    if (kShort) {
      const bool take_s = take && rsq < A.cutsq_s;
      const int tj = (int)(ax >> 24);
#pragma unroll
      for (int el = 0; el < NT; ++el) {
        const bool mine = take_s && tj == el;
        const unsigned long long bm = __ballot(mine);
        if (mine) {
          const unsigned rs = __builtin_amdgcn_mbcnt_hi((unsigned)(bm >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)bm, (unsigned)ns[el]));
          __builtin_amdgcn_raw_buffer_store_b32(e, rsh[el], (int)(rs * 4u), 0, 0);
        }
        ns[el] += __popcll(bm);
      }
    }
  }
}

// the short rows' element count as a template argument (1 when there are no short rows)
template <bool kShort, bool kGr, bool kBigMol>
__device__ __forceinline__ void build_row_nt(const BuildRowArgs<kGr>& A, const float4* c_pos, const uint32_t* c_ent,
                                             const uint32_t* c_aux, const int* sp_row, unsigned long long sp0,
                                             unsigned long long sp1, int lane, int& nl, int (&ns)[kShortMaxType]) {
  if (!kShort) { build_row<false, kGr, kBigMol, 1>(A, c_pos, c_ent, c_aux, sp_row, sp0, sp1, lane, nl, ns); return; }
  switch (A.ntype) {
    case 1: build_row<kShort, kGr, kBigMol, 1>(A, c_pos, c_ent, c_aux, sp_row, sp0, sp1, lane, nl, ns); break;
    case 2: build_row<kShort, kGr, kBigMol, 2>(A, c_pos, c_ent, c_aux, sp_row, sp0, sp1, lane, nl, ns); break;
    case 3: build_row<kShort, kGr, kBigMol, 3>(A, c_pos, c_ent, c_aux, sp_row, sp0, sp1, lane, nl, ns); break;
    default: build_row<kShort, kGr, kBigMol, 4>(A, c_pos, c_ent, c_aux, sp_row, sp0, sp1, lane, nl, ns); break;
  }
}

template <bool kShort, bool kGr>
__global__ void __launch_bounds__(kBuildThreads) k_build_cells(int nhome, const int* __restrict__ home_cells,
                                                     const BoxPar* __restrict__ boxes,
                                                     const int* __restrict__ cell_box, const int4* __restrict__ stencil,
                                                     int nst, float cutsq, float cutsq_s,
                                                     const float4* __restrict__ xq, const int4* __restrict__ info,
                                                     const int* __restrict__ cstart, const int* __restrict__ gstart,
                                                     const int* __restrict__ gs_src, const int* __restrict__ gs_code,
                                                     const int* __restrict__ mt_natom, const int* __restrict__ mt_off,
                                                     const int* __restrict__ mt_sp, const int* __restrict__ typ,
                                                     int ntype, int stride, int cap, int cap_s,
                                                     uint32_t* __restrict__ nbr, int* __restrict__ cnt_out,
                                                     uint32_t* __restrict__ nbr_s, int* __restrict__ cnt_s_out,
                                                     float4* __restrict__ xprev, DevCounters* counters,
                                                     float gr_rmaxsq, float gr_rdr, int gr_nbin, int npair,
                                                     unsigned long long* __restrict__ gr_copies) {
  // kGr: the full g(r) tally of the rebuild's positions rides along -- a real candidate pair
  // within r_max adds 2 from its lower row, a ghost image 1 (2 per unordered pair, the
  // reference's rule), into the workgroup's LDS histogram
  // dynamic LDS, sized at launch so that the g(r) variant still fits four workgroups per CU:
  // [npair][gr_nbin] counts when kGr, then the 2 nst candidate ranges (begin, prefix + total,
  // flags: bit 0 ghost, bit 1 no distance check)
  extern __shared__ unsigned int s_dyn[];
  unsigned int* const s_gh = s_dyn;
  int* const r_begin = reinterpret_cast<int*>(s_dyn + (kGr ? npair * gr_nbin : 0));
  int* const r_pref = r_begin + 2 * nst;
  unsigned char* const r_flag = reinterpret_cast<unsigned char*>(r_pref + 2 * nst + 1);
  __shared__ float4 c_pos[kBuildBatch];              // x, y, z, molecule id (int bits)
  __shared__ uint32_t c_ent[kBuildBatch];            // entry (special class 0) | nocheck << 31
  __shared__ uint32_t c_aux[kBuildBatch];            // intra id | element << 24
  __shared__ int h_nl[kBuildHome], h_ns[kBuildHome][kShortMaxType];
  __shared__ int s_wsum[kBuildThreads / 64 + 1];
  __shared__ int s_max[2];  // longest full and short row of this workgroup
  // one workgroup per cell that can hold atoms (the home cells of every box), so each XCD's
  // contiguous share of the grid is real work
  const int hk = xcd_block(gridDim.x);
  if (hk >= nhome) return;
  const int c = home_cells[hk];
  const int hb = cstart[c], he = cstart[c + 1];
  if (hb == he) return;  // uniform across the workgroup
  const int tid = threadIdx.x, lane = tid & 63;
  // wave index as a wave-uniform value: the home atom, its row pointers and counters then live
  // in scalar registers
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  if (tid < 2) s_max[tid] = 0;
  if (kGr)
    for (int k = tid; k < npair * gr_nbin; k += blockDim.x) s_gh[k] = 0u;
  const BoxPar B = boxes[cell_box[c]];
  const float bl = B.bl, m_bl = B.m_bl;
  const int num_gx = B.num_gx, lc = c - B.cell_base;
  const int cz = lc % num_gx, cy = (lc / num_gx) % num_gx, cx = lc / (num_gx * num_gx);
  // candidate ranges of the stencil, then their exclusive prefix sum
  for (int s = tid; s < nst; s += blockDim.x) {
    const int4 st = stencil[s];
    const int ax = cx + st.x, ay = cy + st.y, az = cz + st.z;
    int nr = 0, ng = 0, br = 0, bg = 0;
    if (ax >= 0 && ay >= 0 && az >= 0 && ax < num_gx && ay < num_gx && az < num_gx) {
      const int nc = B.cell_base + (ax * num_gx + ay) * num_gx + az;
      br = cstart[nc]; nr = cstart[nc + 1] - br;
      bg = gstart[nc]; ng = gstart[nc + 1] - bg;
    }
    const unsigned char nock = (st.w & kStencilNoCheck) ? 2 : 0;
    r_begin[2 * s] = br; r_pref[2 * s] = nr; r_flag[2 * s] = nock;
    r_begin[2 * s + 1] = bg; r_pref[2 * s + 1] = ng; r_flag[2 * s + 1] = nock | 1;
  }
  __syncthreads();
  const int nr2 = 2 * nst;
  const int total = block_scan_excl<kBuildThreads>(r_pref, nr2, s_wsum);
  for (int h0 = hb; h0 < he; h0 += kBuildHome) {
    const int h1 = min(he, h0 + kBuildHome);
    for (int k = tid; k < h1 - h0; k += blockDim.x) {
      h_nl[k] = 0;
      for (int e = 0; e < kShortMaxType; ++e) h_ns[k][e] = 0;
    }
    for (int t0 = 0; t0 < total; t0 += kBuildBatch) {
      const int tn = min(kBuildBatch, total - t0);
      const int tpad = (tn + 63) & ~63;  // <= kBuildBatch, a multiple of 64
      __syncthreads();
      for (int u = tid; u < tpad; u += blockDim.x) {
        if (u >= tn) {
          // padding: beyond every cutoff from any home position, no molecule, no flag
          c_pos[u] = make_float4(1.0e18f, 1.0e18f, 1.0e18f, __int_as_float(-1));
          c_ent[u] = 0u;
          c_aux[u] = 0u;
          continue;
        }
        const int t = t0 + u;
        int lo = 0, hi = nr2 - 1;  // last range with r_pref[k] <= t
        while (lo < hi) {
          const int mid = (lo + hi + 1) >> 1;
          if (r_pref[mid] <= t) lo = mid; else hi = mid - 1;
        }
        const int pos = r_begin[lo] + (t - r_pref[lo]);
        const unsigned char fl = r_flag[lo];
        int row, code;
        float4 p;
        if (fl & 1) { row = gs_src[pos]; code = gs_code[pos]; p = image_pos(xq[row], code, bl, m_bl); }
        else { row = pos; code = 13; p = xq[pos]; }
        const int4 in = info[row];
        const int tr = typ[row];
        c_pos[u] = make_float4(p.x, p.y, p.z, __int_as_float(in.y));
        c_ent[u] = NNPE_MAKE_ENTRY_T(row, ntype <= NNPE_ENTRY_MAX_TYPES ? tr : 0, code, 0) | ((fl & 2) ? 0x80000000u : 0u);
        c_aux[u] = (uint32_t)in.z | ((uint32_t)tr << 24);
      }
      __syncthreads();
      for (int h = h0 + wave; h < h1; h += kBuildThreads / 64) {
        const float4 xi = xq[h];
        const int4 ii = info[h];
        const int th = typ[h];
        const int ti_gr = kGr ? th : 0;
        const uint32_t hent = NNPE_MAKE_ENTRY_T(h, ntype <= NNPE_ENTRY_MAX_TYPES ? th : 0, 13, 0);
        const int mtn = mt_natom[ii.w];
        const int* sp_row = mt_sp + mt_off[ii.w] + ii.z * mtn;
        int nl = h_nl[h - h0];
        int ns[kShortMaxType];
#pragma unroll
        for (int e = 0; e < kShortMaxType; ++e) ns[e] = h_ns[h - h0][e];
        uint32_t* rowl = nbr + (size_t)h * cap;
        BuildRowArgs<kGr> A{xi, ii.y, h, tn, hent, cutsq, cutsq_s, ntype, stride, cap, cap_s, rowl, nbr_s, ti_gr,
                            gr_rmaxsq, gr_rdr, gr_nbin, s_gh};
        if (mtn <= 64) {
          // the home atom's special classes against its molecule, as two lane masks: no memory
          // load in the candidate loop, so no wait on the row stores already in flight
          const int v = lane < mtn ? sp_row[lane] : 0;
          const unsigned long long b0 = __ballot(v & 1), b1 = __ballot(v & 2);
          build_row_nt<kShort, kGr, false>(A, c_pos, c_ent, c_aux, nullptr, b0, b1, lane, nl, ns);
        } else {
          build_row_nt<kShort, kGr, true>(A, c_pos, c_ent, c_aux, sp_row, 0ull, 0ull, lane, nl, ns);
        }
        if (lane == 0) {
          h_nl[h - h0] = nl;
#pragma unroll
          for (int el = 0; el < kShortMaxType; ++el) h_ns[h - h0][el] = ns[el];
        }
      }
    }
    __syncthreads();
    for (int k = tid; k < h1 - h0; k += blockDim.x) {
      const int h = h0 + k, nl = h_nl[k];
      xprev[h] = xq[h];
      cnt_out[h] = min(nl, cap);
      if (nl > cap) atomicMax(&counters->overflow_long, nl);
      atomicMax(&s_max[0], nl);
      if (kShort) {
        for (int el = 0; el < ntype; ++el) {
          const int ns = h_ns[k][el];
          cnt_s_out[(size_t)el * stride + h] = min(ns, cap_s);
          if (ns > cap_s) atomicMax(&counters->overflow_short, ns);
          atomicMax(&s_max[1], ns);
        }
      }
    }
    __syncthreads();
  }
  // one device-scope atomic per workgroup (a per-atom atomic on one address serialises)
  if (tid == 0) {
    atomicMax(&counters->max_long, s_max[0]);
    if (kShort) atomicMax(&counters->max_short, s_max[1]);
  }
  if (kGr) {
    // the workgroup histogram goes into one of kGrShellCopies private copies of the counts (the
    // hardware block index picks it, so each copy takes the adds of workgroups of one XCD):
    // one copy would take every workgroup's device-scope add to the same few hundred bins
    __syncthreads();  // every wave's histogram adds are done
    unsigned long long* dst = gr_copies + (size_t)(blockIdx.x % kGrShellCopies) * (npair * gr_nbin);
    for (int k = tid; k < npair * gr_nbin; k += blockDim.x) {
      const unsigned int v = s_gh[k];
      if (v) atomicAdd(&dst[k], (unsigned long long)v);
    }
  }
}

// NEIGHLIST::isvalid (src/neighlist.cpp:178-187)
__global__ void k_isvalid(int n, float dx2, const float4* xq, const float4* xprev, DevCounters* cnt) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  bool bad = false;
  if (i < n) {
    float4 a = xq[i], b = xprev[i];
    float d = dist2_rn(__fsub_rn(a.x, b.x), __fsub_rn(a.y, b.y), __fsub_rn(a.z, b.z));
    bad = d > dx2;
  }
  if (__any(bad) && (threadIdx.x & 63) == 0) atomicOr(&cnt->invalid, 1);
}

__global__ void k_clear_invalid(DevCounters* cnt) { cnt->invalid = 0; }

// list export in atom ids
__global__ void k_export(int n, int cap, const uint32_t* nbr, const int* cnt, const long long* offs, const int4* info,
                         int* ai, int* aj, int* code, int* sp) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  long long o = offs[i];
  int a = info[i].x;
  for (int k = 0; k < cnt[i]; ++k) {
    uint32_t e = nbr[(size_t)i * cap + k];
    ai[o + k] = a;
    aj[o + k] = info[NNPE_ENTRY_J(e)].x;
    code[o + k] = NNPE_ENTRY_CODE(e);
    sp[o + k] = NNPE_ENTRY_SP(e);
  }
}

// relative stencil with the reference's masks (src/neighlist.cpp:35-142): all cells whose
// minimum squared cell distance is < ceil((cutoff/dx)^2); "no check" when the maximum is too.
std::vector<int4> make_stencil(float cutoff, float g_rdx, int num_gx) {
  float t = cutoff * g_rdx;
  const int max_d = int(std::ceil(t * t));
  std::vector<int4> out;
  auto sq = [](int v) { return v * v; };
  for (int dx = -(num_gx - 1); dx <= num_gx - 1; ++dx) {
    int mx, mxx;
    if (dx == 0) { mx = 0; mxx = 1; } else { int a = std::abs(dx); if ((mx = sq(a - 1)) >= max_d) continue; mxx = sq(a + 1); }
    for (int dy = -(num_gx - 1); dy <= num_gx - 1; ++dy) {
      int my, mxy;
      if (dy == 0) { my = mx; mxy = mxx + 1; } else { int a = std::abs(dy); if ((my = mx + sq(a - 1)) >= max_d) continue; mxy = mxx + sq(a + 1); }
      for (int dz = -(num_gx - 1); dz <= num_gx - 1; ++dz) {
        int mz, mxz;
        if (dz == 0) { mz = my; mxz = mxy + 1; } else { int a = std::abs(dz); if ((mz = my + sq(a - 1)) >= max_d) continue; mxz = mxy + sq(a + 1); }
        (void)mz;
        out.push_back(make_int4(dx, dy, dz, mxz < max_d ? kStencilNoCheck : 0));
      }
    }
  }
  return out;
}

}  // namespace

extern "C" int nnpe_neigh_init(nnpe_ctx* ctx, float pair_cutoff, float skin, float gr_cutoff, float short_cutoff) {
  if (ctx) md_state_free(ctx);  // captured step graphs hold the old parameters
  if (!ctx) return NNPE_ERR_ARG;
  if (!ctx->have_system) return ctx_fail(ctx, NNPE_ERR_STATE, "nnpe_neigh_init before nnpe_system_set");
  if (!(pair_cutoff > 0) || !(skin >= 0) || short_cutoff > pair_cutoff)
    return ctx_fail(ctx, NNPE_ERR_ARG, "nnpe_neigh_init: bad cutoffs");
  ctx->pair_cutoff = pair_cutoff; ctx->skin = skin; ctx->gr_cutoff = gr_cutoff; ctx->short_cutoff = short_cutoff;
  // src/run.cpp:76-79
  const float comm = skin + std::max(gr_cutoff, pair_cutoff);
  ctx->comm_cutoff = comm;
  // one shell of images: atoms within comm of both faces (L/2 < comm < L) get images on both
  // sides, which the reference's single-sided ghosts cannot (it needs L/2 > comm)
  if (!(ctx->bl > comm))
    return ctx_fail(ctx, NNPE_ERR_UNSUPPORTED, "box too small: L must exceed skin + max(rc, r_gr) (one image shell)");
  const BoxPar G0 = box_geometry(ctx->bl, pair_cutoff, skin, gr_cutoff);
  ctx->ghost_lim = G0.ghost_lim;
  ctx->g_dx = float(0.5 * (skin + pair_cutoff));
  ctx->g_rdx = G0.g_rdx;
  ctx->g_offset = G0.g_offset;
  ctx->num_gx = G0.num_gx;
  ctx->num_gxsq = ctx->num_gx * ctx->num_gx;
  ctx->num_grid = ctx->num_gxsq * ctx->num_gx;
  const float lc = pair_cutoff + skin;
  ctx->cutsq_list = lc * lc;  // NEIGHLIST ctor: cutsq = sqr(cutoff + skin)
  const float sc = short_cutoff + skin;
  ctx->cutsq_short = short_cutoff > 0 ? sc * sc : 0.0f;
  const float hs = skin / 2;
  ctx->dx2 = hs * hs;
  auto sp = make_stencil(lc, ctx->g_rdx, ctx->num_gx);
  auto sg = make_stencil(gr_cutoff + skin, ctx->g_rdx, ctx->num_gx);
  ctx->n_stencil_pair = int(sp.size()); ctx->n_stencil_gr = int(sg.size());
  NNPE_CHECK_HIP(ctx, dev_alloc(ctx->stencil_pair, sp.size()));
  NNPE_CHECK_HIP(ctx, hipMemcpy(ctx->stencil_pair, sp.data(), sp.size() * sizeof(int4), hipMemcpyHostToDevice));
  NNPE_CHECK_HIP(ctx, dev_alloc(ctx->stencil_gr, sg.size()));
  NNPE_CHECK_HIP(ctx, hipMemcpy(ctx->stencil_gr, sg.data(), sg.size() * sizeof(int4), hipMemcpyHostToDevice));
  // g(r) cells outside the pair stencil: the shell pass of a tally folded into a rebuild
  std::vector<int4> so;
  for (const int4& g : sg) {
    bool in = false;
    for (const int4& q : sp) in = in || (q.x == g.x && q.y == g.y && q.z == g.z);
    if (!in) so.push_back(g);
  }
  ctx->h_stencil_gr_outer = so;
  const int G = ctx->num_grid;
  NNPE_CHECK_HIP(ctx, dev_alloc(ctx->cell_start, G + 1)); NNPE_CHECK_HIP(ctx, dev_alloc(ctx->cell_cnt, G));
  NNPE_CHECK_HIP(ctx, dev_alloc(ctx->cell_fill, G));
  NNPE_CHECK_HIP(ctx, dev_alloc(ctx->gcell_start, G + 1)); NNPE_CHECK_HIP(ctx, dev_alloc(ctx->gcell_cnt, G));
  NNPE_CHECK_HIP(ctx, dev_alloc(ctx->gcell_fill, G));
  NNPE_CHECK_HIP(ctx, hipMemset(ctx->cell_start, 0, sizeof(int) * (G + 1)));
  NNPE_CHECK_HIP(ctx, hipMemset(ctx->gcell_start, 0, sizeof(int) * (G + 1)));
  // ghost capacity: the shell volume fraction with a 30% margin (+4096), at most 26 images/atom
  double vin = double(ctx->bl), vout = double(ctx->bl) + 2.0 * comm;
  double frac = (vout * vout * vout - vin * vin * vin) / (vin * vin * vin);
  long long gcap = (long long)(1.3 * frac * ctx->natom) + 4096;
  gcap = std::min<long long>(gcap, 26LL * ctx->natom + 64);
  ctx->ghost_cap = int(gcap);
  NNPE_CHECK_HIP(ctx, dev_alloc(ctx->g_src, gcap)); NNPE_CHECK_HIP(ctx, dev_alloc(ctx->g_code, gcap));
  NNPE_CHECK_HIP(ctx, dev_alloc(ctx->g_cell, gcap)); NNPE_CHECK_HIP(ctx, dev_alloc(ctx->gs_src, gcap));
  NNPE_CHECK_HIP(ctx, dev_alloc(ctx->gs_code, gcap));
  // list capacities: 1.5x the ideal-gas count of the shell at the system's mean density (+64),
  // rounded to 64; grown on overflow
  double rho = ctx->natom / (vin * vin * vin);
  auto capfor = [&](float rc) {
    double c = 4.0 / 3.0 * 3.141592653589793 * double(rc) * rc * rc * rho;
    int k = int(1.5 * c) + 64;
    return (k + 63) / 64 * 64;
  };
  ctx->cap_long = capfor(lc);
  ctx->cap_short = short_cutoff > 0 ? capfor(sc) : 0;
  NNPE_CHECK_HIP(ctx, dev_alloc(ctx->nbr_long, size_t(ctx->cap_long) * ctx->stride));
  NNPE_CHECK_HIP(ctx, dev_alloc(ctx->cnt_long, ctx->stride));
  NNPE_CHECK_HIP(ctx, hipMemset(ctx->cnt_long, 0, sizeof(int) * ctx->stride));
  if (ctx->cap_short) {
    // one short list per neighbour element: rows of element e at nbr_short + e * cap_short * stride
    NNPE_CHECK_HIP(ctx, dev_alloc(ctx->nbr_short, size_t(ctx->ntype) * ctx->cap_short * ctx->stride));
    NNPE_CHECK_HIP(ctx, dev_alloc(ctx->cnt_short, size_t(ctx->ntype) * ctx->stride));
    NNPE_CHECK_HIP(ctx, hipMemset(ctx->cnt_short, 0, sizeof(int) * ctx->ntype * ctx->stride));
  }
  ctx->nbox = 1;
  ctx->h_box.assign(1, G0);
  int rcb = box_upload_single(ctx);
  if (rcb) return rcb;
  ctx->have_grid = true;
  ctx->have_list = false;
  const int rco = gr_outer_refresh(ctx);
  if (rco) return rco;
  return nnp_act_reserve(ctx);
}

// ENVIRON::set_bl + init_grid + LOCAL::set_ghost_lim (src/environ.h:105, src/environ.cpp:143-155,
// src/run.cpp:76-82) for one box of edge bl, in the reference's FP32 operations
BoxPar box_geometry(float bl, float pair_cutoff, float skin, float gr_cutoff) {
  BoxPar B{};
  B.bl = bl; B.r_bl = 1.0f / bl; B.h_bl = bl * .5f; B.m_bl = -bl;
  const float comm = skin + std::max(gr_cutoff, pair_cutoff);
  B.ghost_lim = B.h_bl - comm;
  const float gl = float(0.5 * (skin + pair_cutoff));
  B.g_rdx = 1.0f / gl;
  B.g_offset = comm + B.h_bl;
  B.num_gx = int(std::ceil((B.g_offset * 2) * B.g_rdx));
  B.cell_base = 0;
  // a refolded coordinate lies in [-L/2, L/2] (to a few roundings): its cell, by the list
  // kernels' FP32 operations, lies between those of the slightly widened edges
  const float edge = B.h_bl * (1.0f + 4e-6f);
  B.home_lo = std::max(0, int((B.g_offset + (-edge)) * B.g_rdx));
  B.home_hi = std::min(B.num_gx - 1, int((B.g_offset + edge) * B.g_rdx));
  return B;
}

int home_cells_upload(nnpe_ctx* ctx) {
  std::vector<int> hc;
  for (const BoxPar& b : ctx->h_box)
    for (int x = b.home_lo; x <= b.home_hi; ++x)
      for (int y = b.home_lo; y <= b.home_hi; ++y)
        for (int z = b.home_lo; z <= b.home_hi; ++z) hc.push_back(b.cell_base + (x * b.num_gx + y) * b.num_gx + z);
  ctx->n_home_cells = int(hc.size());
  NNPE_CHECK_HIP(ctx, dev_alloc(ctx->home_cells, hc.size()));
  if (!hc.empty()) NNPE_CHECK_HIP(ctx, hipMemcpy(ctx->home_cells, hc.data(), sizeof(int) * hc.size(), hipMemcpyHostToDevice));
  return NNPE_OK;
}

// one box: every cell belongs to box 0
int box_upload_single(nnpe_ctx* ctx) {
  NNPE_CHECK_HIP(ctx, dev_alloc(ctx->d_box, ctx->h_box.size()));
  NNPE_CHECK_HIP(ctx, hipMemcpy(ctx->d_box, ctx->h_box.data(), sizeof(BoxPar) * ctx->h_box.size(), hipMemcpyHostToDevice));
  NNPE_CHECK_HIP(ctx, dev_alloc(ctx->cell_box, ctx->num_grid));
  NNPE_CHECK_HIP(ctx, hipMemset(ctx->cell_box, 0, sizeof(int) * ctx->num_grid));
  const int rh = home_cells_upload(ctx);
  if (rh) return rh;
  return NNPE_OK;
}

// Enqueue the rebuild kernels (no host synchronisation; graph-capturable).
int neigh_enqueue_build(nnpe_ctx* ctx, bool gr_tally, bool skin_accum, bool defer_shell) {
  const int N = ctx->natom, G = ctx->num_grid, bs = 256;
  hipStream_t s = ctx->stream;
  if (ctx->n_stencil_pair > kBuildMaxRanges / 2) return ctx_fail(ctx, NNPE_ERR_UNSUPPORTED, "pair stencil too large");
  if (ctx->cap_short && ctx->ntype > kShortMaxType) return ctx_fail(ctx, NNPE_ERR_UNSUPPORTED, "short lists: at most 4 elements");
  if (ctx->ntype > 255) return ctx_fail(ctx, NNPE_ERR_UNSUPPORTED, "list build: at most 255 elements");
  gr_tally = gr_tally && gr_fusable(ctx);
  const int gr_n = gr_tally ? ctx->npair * ctx->gr_nbin : 0;
  const int gr_shell_n = gr_tally ? gr_shell_size(ctx) : 0;
  const int* type_off = mlp_type_off(ctx);  // the per-element row index is refilled by k_copy_back_ghosts
  hipLaunchKernelGGL(k_build_reset, dim3(blocks_for(std::max(G, gr_shell_n), bs)), dim3(bs), 0, s, G, ctx->cell_cnt,
                     ctx->cell_fill, ctx->gcell_cnt, ctx->gcell_fill, ctx->counters, gr_n, ctx->gr_count, gr_shell_n,
                     ctx->gr_shell, type_off ? ctx->ntype : 0, ctx->type_fill, int(skin_accum));
  hipLaunchKernelGGL(k_refold, dim3(blocks_for(N, bs)), dim3(bs), 0, s, N, ctx->fp64, ctx->d_box, ctx->rbox, ctx->xq, ctx->xd,
                     ctx->t_cell, ctx->cell_cnt, ctx->counters);
  hipLaunchKernelGGL(k_scan_excl, dim3(1), dim3(1024), 0, s, ctx->cell_cnt, ctx->cell_start, G);
  hipLaunchKernelGGL(k_scatter, dim3(blocks_for(N, bs)), dim3(bs), 0, s, N, ctx->t_cell, ctx->cell_start, ctx->cell_fill, ctx->order);
  hipLaunchKernelGGL(k_cell_sort, dim3(blocks_for((long long)G * 64, 256)), dim3(256), 0, s, G, ctx->cell_start, ctx->order,
                     ctx->info);
  hipLaunchKernelGGL(k_gather, dim3(blocks_for(N, bs)), dim3(bs), 0, s, N, ctx->fp64, ctx->order, ctx->t_cell, ctx->xq, ctx->v,
                     ctx->fprev, ctx->rm, ctx->typ, ctx->info, ctx->xd, ctx->vd, ctx->ad, ctx->rbox, ctx->t_xq, ctx->t_v,
                     ctx->t_fprev, ctx->t_rm, ctx->t_typ, ctx->t_info, ctx->t_xd, ctx->t_vd, ctx->t_ad, ctx->t_rbox, ctx->cell,
                     ctx->inv);
  // copy the permuted rows back so every array keeps its buffer (captured graphs replay with
  // fixed pointers), and generate the ghost images from the gathered copies
  hipLaunchKernelGGL(k_copy_back_ghosts, dim3(blocks_for(N, bs)), dim3(bs), 0, s, N, ctx->fp64, ctx->t_xq, ctx->t_v,
                     ctx->t_fprev, ctx->t_rm, ctx->t_typ, ctx->t_info, ctx->t_xd, ctx->t_vd, ctx->t_ad, ctx->t_rbox, ctx->xq,
                     ctx->v, ctx->fprev, ctx->rm, ctx->typ, ctx->info, ctx->xd, ctx->vd, ctx->ad, ctx->rbox, ctx->ntype,
                     type_off, ctx->type_fill, type_off ? ctx->type_idx : nullptr, ctx->d_box, ctx->ghost_cap, ctx->g_src,
                     ctx->g_code, ctx->g_cell, ctx->gcell_cnt, ctx->counters, ctx->mol_first, ctx->inv, ctx->xmol_code);
  hipLaunchKernelGGL(k_scan_excl, dim3(1), dim3(1024), 0, s, ctx->gcell_cnt, ctx->gcell_start, G);
  hipLaunchKernelGGL(k_ghost_scatter, dim3(blocks_for(ctx->ghost_cap, bs)), dim3(bs), 0, s, ctx->ghost_cap, ctx->counters,
                     ctx->g_src, ctx->g_code, ctx->g_cell, ctx->gcell_start, ctx->gcell_fill, ctx->gs_src, ctx->gs_code);
  hipLaunchKernelGGL(k_ghost_cell_sort, dim3(blocks_for((long long)G * 64, 256)), dim3(256), 0, s, G, ctx->gcell_start,
                     ctx->gs_src, ctx->gs_code, ctx->info);
  // dynamic LDS of the list kernel: the g(r) counts, then the candidate ranges (two per stencil cell)
  const size_t nr2 = 2 * size_t(ctx->n_stencil_pair);
  const size_t gh = (gr_tally ? sizeof(unsigned int) * ctx->npair * ctx->gr_nbin : 0) + sizeof(int) * (2 * nr2 + 1) + nr2;
#define NNPE_BUILD(SHORT, GR)                                                                                             \
  hipLaunchKernelGGL((k_build_cells<SHORT, GR>), dim3(ctx->n_home_cells), dim3(kBuildThreads), gh, s, ctx->n_home_cells,           \
                     ctx->home_cells, ctx->d_box, ctx->cell_box, ctx->stencil_pair,                                               \
                     ctx->n_stencil_pair, ctx->cutsq_list, SHORT ? ctx->cutsq_short : 0.0f, ctx->xq, ctx->info,            \
                     ctx->cell_start, ctx->gcell_start, ctx->gs_src, ctx->gs_code, ctx->mt_natom, ctx->mt_special_off,      \
                     ctx->mt_special, ctx->typ, ctx->ntype, ctx->stride, ctx->cap_long, SHORT ? ctx->cap_short : 0,          \
                     ctx->nbr_long, ctx->cnt_long, SHORT ? ctx->nbr_short : nullptr, SHORT ? ctx->cnt_short : nullptr,      \
                     ctx->xprev, ctx->counters, ctx->gr_rmaxsq, ctx->gr_rdr, ctx->gr_nbin, ctx->npair, ctx->gr_shell)
  if (ctx->cap_short) {
    if (gr_tally) NNPE_BUILD(true, true); else NNPE_BUILD(true, false);
  } else {
    if (gr_tally) NNPE_BUILD(false, true); else NNPE_BUILD(false, false);
  }
#undef NNPE_BUILD
  NNPE_CHECK_LAUNCH(ctx);
  ctx->nbuild++;
  if (gr_tally && !defer_shell) {
    const int rc = gr_shell_finish_enqueue(ctx);
    if (rc) return rc;
  }
  return NNPE_OK;
}

// Check the build's counters; grow capacities and rebuild on list/ghost overflow.
int neigh_check_and_grow(nnpe_ctx* ctx, bool* rebuilt) {
  *rebuilt = false;
  int rc = counters_read(ctx);
  if (rc) return rc;
  DevCounters c = *ctx->h_counters;
  if (c.cell_range) return ctx_fail(ctx, NNPE_ERR_CELL_RANGE, "atoms outside the cell grid");
  bool grow = false;
  if (c.ghost_overflow || c.nghost > ctx->ghost_cap) {
    ctx->ghost_cap = c.nghost + c.nghost / 4 + 1024;
    NNPE_CHECK_HIP(ctx, dev_alloc(ctx->g_src, ctx->ghost_cap)); NNPE_CHECK_HIP(ctx, dev_alloc(ctx->g_code, ctx->ghost_cap));
    NNPE_CHECK_HIP(ctx, dev_alloc(ctx->g_cell, ctx->ghost_cap)); NNPE_CHECK_HIP(ctx, dev_alloc(ctx->gs_src, ctx->ghost_cap));
    NNPE_CHECK_HIP(ctx, dev_alloc(ctx->gs_code, ctx->ghost_cap));
    grow = true;
  }
  if (c.overflow_long) {
    ctx->cap_long = (c.overflow_long + c.overflow_long / 4 + 63) / 64 * 64;
    NNPE_CHECK_HIP(ctx, dev_alloc(ctx->nbr_long, size_t(ctx->cap_long) * ctx->stride));
    grow = true;
  }
  if (c.overflow_short) {
    ctx->cap_short = (c.overflow_short + c.overflow_short / 4 + 63) / 64 * 64;
    NNPE_CHECK_HIP(ctx, dev_alloc(ctx->nbr_short, size_t(ctx->ntype) * ctx->cap_short * ctx->stride));
    rc = nnp_act_reserve(ctx);
    if (rc) return rc;
    grow = true;
  }
  if (grow) {
    md_state_free(ctx);
    // the positions were refolded and sorted already; a second build sees the same set
    int r = neigh_enqueue_build(ctx);
    if (r) return r;
    r = counters_read(ctx);
    if (r) return r;
    c = *ctx->h_counters;
    if (c.ghost_overflow) return ctx_fail(ctx, NNPE_ERR_GHOST_OVERFLOW, "ghost overflow after growth");
    if (c.overflow_long || c.overflow_short) return ctx_fail(ctx, NNPE_ERR_NEIGH_OVERFLOW, "nei_list overflow after growth");
    *rebuilt = true;
  }
  return NNPE_OK;
}

extern "C" int nnpe_neigh_build(nnpe_ctx* ctx) {
  if (!ctx) return NNPE_ERR_ARG;
  if (!ctx->have_grid) return ctx_fail(ctx, NNPE_ERR_STATE, "nnpe_neigh_build before nnpe_neigh_init");
  int rc = neigh_enqueue_build(ctx);
  if (rc) return rc;
  bool rebuilt;
  rc = neigh_check_and_grow(ctx, &rebuilt);
  if (rc) return rc;
  ctx->have_list = true;
  return NNPE_OK;
}

extern "C" int nnpe_neigh_isvalid(nnpe_ctx* ctx, int* valid) {
  if (!ctx || !valid) return NNPE_ERR_ARG;
  if (!ctx->have_list) { *valid = 0; return NNPE_OK; }
  hipLaunchKernelGGL(k_clear_invalid, dim3(1), dim3(1), 0, ctx->stream, ctx->counters);
  hipLaunchKernelGGL(k_isvalid, dim3(blocks_for(ctx->natom, 256)), dim3(256), 0, ctx->stream, ctx->natom, ctx->dx2, ctx->xq,
                     ctx->xprev, ctx->counters);
  NNPE_CHECK_LAUNCH(ctx);
  int rc = counters_read(ctx);
  if (rc) return rc;
  *valid = ctx->h_counters->invalid ? 0 : 1;
  return NNPE_OK;
}

extern "C" int nnpe_neigh_size(nnpe_ctx* ctx, int which, long long* n_entries, int* n_ghost) {
  if (!ctx || (which != 0 && which != 1)) return NNPE_ERR_ARG;
  if (!ctx->have_list) return ctx_fail(ctx, NNPE_ERR_STATE, "no list");
  if (which == 1 && !ctx->cap_short) return ctx_fail(ctx, NNPE_ERR_STATE, "no short list");
  const int nl = which ? ctx->ntype : 1;  // the short list is one list per neighbour element
  long long s = 0;
  std::vector<int> h(ctx->natom);
  NNPE_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
  for (int el = 0; el < nl; ++el) {
    const int* dc = which ? ctx->cnt_short + size_t(el) * ctx->stride : ctx->cnt_long;
    NNPE_CHECK_HIP(ctx, hipMemcpy(h.data(), dc, sizeof(int) * ctx->natom, hipMemcpyDeviceToHost));
    for (int v : h) s += v;
  }
  if (n_entries) *n_entries = s;
  int rc = counters_read(ctx);
  if (rc) return rc;
  if (n_ghost) *n_ghost = ctx->h_counters->nghost;
  return NNPE_OK;
}

extern "C" int nnpe_neigh_export(nnpe_ctx* ctx, int which, int* ai, int* aj, int* code, int* sp, long long cap,
                                 long long* n_out) {
  if (!ctx || !ai || !aj || !code || !sp || (which != 0 && which != 1)) return NNPE_ERR_ARG;
  if (!ctx->have_list) return ctx_fail(ctx, NNPE_ERR_STATE, "no list");
  if (which == 1 && !ctx->cap_short) return ctx_fail(ctx, NNPE_ERR_STATE, "no short list");
  const int N = ctx->natom;
  const int nl = which ? ctx->ntype : 1;  // the short list is one list per neighbour element
  std::vector<int> h(size_t(N) * nl);
  NNPE_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
  for (int el = 0; el < nl; ++el) {
    const int* dcs = which ? ctx->cnt_short + size_t(el) * ctx->stride : ctx->cnt_long;
    NNPE_CHECK_HIP(ctx, hipMemcpy(h.data() + size_t(el) * N, dcs, sizeof(int) * N, hipMemcpyDeviceToHost));
  }
  std::vector<long long> offs(size_t(N) * nl);
  long long tot = 0;
  for (size_t k = 0; k < offs.size(); ++k) { offs[k] = tot; tot += h[k]; }
  if (n_out) *n_out = tot;
  if (tot > cap) return ctx_fail(ctx, NNPE_ERR_ARG, "export capacity too small");
  long long* doffs = nullptr;
  int *dai = nullptr, *daj = nullptr, *dc = nullptr, *ds = nullptr;
  NNPE_CHECK_HIP(ctx, dev_alloc(doffs, offs.size()));
  NNPE_CHECK_HIP(ctx, dev_alloc(dai, tot)); NNPE_CHECK_HIP(ctx, dev_alloc(daj, tot));
  NNPE_CHECK_HIP(ctx, dev_alloc(dc, tot)); NNPE_CHECK_HIP(ctx, dev_alloc(ds, tot));
  NNPE_CHECK_HIP(ctx, hipMemcpy(doffs, offs.data(), sizeof(long long) * offs.size(), hipMemcpyHostToDevice));
  for (int el = 0; el < nl; ++el) {
    const int capx = which ? ctx->cap_short : ctx->cap_long;
    const uint32_t* nb = which ? ctx->nbr_short + size_t(el) * ctx->cap_short * ctx->stride : ctx->nbr_long;
    const int* dcs = which ? ctx->cnt_short + size_t(el) * ctx->stride : ctx->cnt_long;
    hipLaunchKernelGGL(k_export, dim3(blocks_for(N, 256)), dim3(256), 0, ctx->stream, N, capx, nb, dcs,
                       doffs + size_t(el) * N, ctx->info, dai, daj, dc, ds);
    NNPE_CHECK_LAUNCH(ctx);
  }
  NNPE_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
  NNPE_CHECK_HIP(ctx, hipMemcpy(ai, dai, sizeof(int) * tot, hipMemcpyDeviceToHost));
  NNPE_CHECK_HIP(ctx, hipMemcpy(aj, daj, sizeof(int) * tot, hipMemcpyDeviceToHost));
  NNPE_CHECK_HIP(ctx, hipMemcpy(code, dc, sizeof(int) * tot, hipMemcpyDeviceToHost));
  NNPE_CHECK_HIP(ctx, hipMemcpy(sp, ds, sizeof(int) * tot, hipMemcpyDeviceToHost));
  dev_free(doffs); dev_free(dai); dev_free(daj); dev_free(dc); dev_free(ds);
  return NNPE_OK;
}
