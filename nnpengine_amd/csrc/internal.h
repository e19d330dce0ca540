// internal.h -- device context of the nnpe engine (not part of the ABI).
//
// HBM layout (one context = one system):
//  * real atoms sorted by cell; per-atom SoA arrays of 16-byte vectors so a wave's loads are
//    coalesced 1 KiB transactions: xq = (x, y, z, q) FP32, v, f, fprev (xyz + pad);
//    FP64 integration state xd/vd/ad as double4 when that mode is on;
//  * info = (atom id, molecule id, intra id, molecule type), typ, 1/m, cell, inverse map;
//  * ghosts are not materialised: a ghost is (source row, image code) sorted by ghost cell,
//    its position is xq[src] + image shift computed in FP32 exactly as the reference's
//    x_ghost = x + (+-L) (src/environ.cpp:431-433, :505-507);
//  * neighbour lists are padded rows: entry k of atom i at nbr[i * cap + k], read by a group of
//    lanes per atom (consecutive words); an entry packs (j row : 23 bits, element of j : 2 bits,
//    image code : 5 bits, special class : 2 bits).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>
#include <string>
#include <vector>

#include "../../include/nnpe.h"

// neighbour-list entry: row j (23 bits) | element of j (2 bits) | image code (5) | special class (2)
#define NNPE_ENTRY_J(e) ((int)((e) & 0x7FFFFFu))
#define NNPE_ENTRY_T(e) ((int)(((e) >> 23) & 3u))
#define NNPE_ENTRY_CODE(e) ((int)(((e) >> 25) & 31u))
#define NNPE_ENTRY_SP(e) ((int)((e) >> 30))
#define NNPE_MAKE_ENTRY(j, code, sp) ((uint32_t)(j) | ((uint32_t)(code) << 25) | ((uint32_t)(sp) << 30))
#define NNPE_MAKE_ENTRY_T(j, t, code, sp) (NNPE_MAKE_ENTRY(j, code, sp) | ((uint32_t)(t) << 23))
#define NNPE_MAX_ATOMS (1 << 23)
// entries carry the neighbour's element when there are at most this many elements (else 0, and
// the kernels that need it read typ[j])
#define NNPE_ENTRY_MAX_TYPES 4

// device counters, one small block read back by the host when needed
struct DevCounters {
  int nghost;          // ghosts generated by the last build
  int overflow_long;   // max entries of any atom when > capacity
  int overflow_short;
  int ghost_overflow;
  int cell_range;      // atoms outside the grid
  int invalid;         // set by the isvalid kernel
  int max_long, max_short;
  int skin_violations; // fixed-interval rebuilds that came after an atom moved > skin/2
  int log_count;       // MD log records written
};

// Per-box parameters.  A context holds one periodic box, or (serving) many independent boxes
// side by side: every row carries its box index (rbox), every cell belongs to one box (cell_box),
// and the kernels take L, the cell grid and the ghost limit of the row's or cell's box, so boxes
// of different sizes advance together in the same launches without interacting.
struct BoxPar {
  float bl, r_bl, h_bl, m_bl;           // L, 1/L, L/2, -L
  float g_offset, g_rdx, ghost_lim, pad0;
  int num_gx, cell_base;                // cells per axis, first global cell of the box
  int home_lo, home_hi;                 // cell coordinates a refolded atom can take, per axis
};

struct NnpState;  // descriptor tables and moment buffers (nnp.hip)
struct MlpState;  // per-element network weights (mlp.hip)
struct RmdfState; // S(q) refinement of the pair tables (rmdf.hip)
struct SqdState;  // direct k-space S(q) (sqd.hip)

struct nnpe_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  // MD steps: the classical terms run on stream2 beside the NNP descriptor and network on
  // stream (fork/join by the two events; default priority, like stream)
  hipStream_t stream2 = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  std::string err;

  // ---- system ---------------------------------------------------------------------------
  int natom = 0, ntype = 0, npair = 0, nmol = 0, nmoltype = 0, stride = 0;
  float bl = 0, r_bl = 0, h_bl = 0, m_bl = 0;
  bool have_system = false;
  float4 *xq = nullptr, *v = nullptr, *f = nullptr, *fprev = nullptr, *xprev = nullptr;
  // second position buffer: inside a captured rebuild interval the NNP force pass of one step
  // writes the next step's positions here while other atoms still read this step's, and the
  // two pointers trade places from step to step (md.hip)
  float4* xq_alt = nullptr;
  double4 *xd = nullptr, *vd = nullptr, *ad = nullptr;
  float* rm = nullptr;
  int* typ = nullptr;
  int4* info = nullptr;     // atom id, mol id, intra id, molecule type
  int* inv = nullptr;       // atom id -> row
  int* cell = nullptr;      // cell of each row
  int* xmol_code = nullptr; // image code of the unwrapped-molecule position of each row
  int nbox = 1;
  BoxPar* d_box = nullptr;  // [nbox]
  std::vector<BoxPar> h_box;
  int* rbox = nullptr;      // box of each row (permuted with the rows)
  int* t_rbox = nullptr;
  int* cell_box = nullptr;  // box of each cell [num_grid]
  int* home_cells = nullptr;  // cells that can hold atoms, in cell order [n_home_cells]
  int n_home_cells = 0;
  // permutation scratch
  float4 *t_xq = nullptr, *t_v = nullptr, *t_fprev = nullptr, *t_xprev = nullptr;
  double4 *t_xd = nullptr, *t_vd = nullptr, *t_ad = nullptr;
  float* t_rm = nullptr;
  int *t_typ = nullptr, *t_cell = nullptr, *order = nullptr;
  int4* t_info = nullptr;
  int* mol_first = nullptr;                      // atom id of each molecule's intra-0 atom
  int *mt_natom = nullptr, *mt_special_off = nullptr, *mt_special = nullptr;
  std::vector<int> h_mt_natom;

  // ---- grid / ghosts / lists -----------------------------------------------------------
  bool have_grid = false, have_list = false;
  float pair_cutoff = 0, skin = 0, gr_cutoff = 0, short_cutoff = 0, comm_cutoff = 0;
  float cutsq_list = 0, cutsq_short = 0, dx2 = 0;
  int num_gx = 0, num_gxsq = 0, num_grid = 0;
  float g_dx = 0, g_rdx = 0, g_offset = 0, ghost_lim = 0;
  int *cell_start = nullptr, *cell_cnt = nullptr, *cell_fill = nullptr;   // [num_grid (+1)]
  int *gcell_start = nullptr, *gcell_cnt = nullptr, *gcell_fill = nullptr;
  int ghost_cap = 0;
  int *g_src = nullptr, *g_code = nullptr, *g_cell = nullptr;  // unsorted ghosts
  int *gs_src = nullptr, *gs_code = nullptr;                  // sorted by cell
  int4* stencil_pair = nullptr; int n_stencil_pair = 0;
  int4* stencil_gr = nullptr; int n_stencil_gr = 0;
  // g(r) stencil cells outside the pair stencil (all of them on the host; on the device those an
  // atom of the home cell can reach within the tally's range)
  std::vector<int4> h_stencil_gr_outer;
  int4* stencil_gr_outer = nullptr; int n_stencil_gr_outer = 0;
  int cap_long = 0, cap_short = 0;
  uint32_t *nbr_long = nullptr, *nbr_short = nullptr;
  int *cnt_long = nullptr, *cnt_short = nullptr;
  DevCounters* counters = nullptr;     // device
  DevCounters* h_counters = nullptr;   // pinned host mirror
  int* scan_tmp = nullptr;             // block sums for the cell scans
  int nbuild = 0;

  // ---- energy / virial partial sums (deterministic two-stage reduction) ----------------
  double* ev_partial = nullptr;   // [NNPE_NTERM][2][max_blocks]
  double* ev_sum = nullptr;       // [NNPE_NTERM][2] accumulated
  double* h_ev = nullptr;         // pinned
  int ev_max_blocks = 0;

  // ---- pair table --------------------------------------------------------------------------
  bool have_pair = false;
  float pt_cut = 0, pt_cutsq = 0, pt_rdr = 0, pt_fc = 0, pt_vc = 0;
  int pt_nbin = 0;
  float* pt_tab = nullptr;  // [npair][nbin+1] force then [npair][nbin+1] potential

  // ---- bonded (kind 0 bonds, 1 angles, 2 dihedrals) -----------------------------------------
  // Terms as given: [nterm][2 + kind] atom ids, kind k has 2 + k atoms, and a type per term.
  // Coefficients: 4 floats per type for bonds and angles (the classes' stored form); 12 per
  // dihedral type (derivative polynomial d0..d4 and energy polynomial e0..e5 in cos phi, one pad).
  // Atom membership: bd_start[k][atom id .. +1] delimits bd_ref[k], the (term << 2 | slot) of
  // every term the atom belongs to, in term order.
  int bd_ntyp[3] = {0, 0, 0}, bd_nterm[3] = {0, 0, 0};
  float* bd_coef[3] = {nullptr, nullptr, nullptr};
  int* bd_type[3] = {nullptr, nullptr, nullptr};
  int* bd_atoms[3] = {nullptr, nullptr, nullptr};
  int* bd_start[3] = {nullptr, nullptr, nullptr};
  int* bd_ref[3] = {nullptr, nullptr, nullptr};
  int bd_start_len[3] = {0, 0, 0};  // natom + 1 of the atom table each membership table was built for

  // ---- verlet --------------------------------------------------------------------------------
  double dt = 1.0;
  bool fp64 = false;

  // ---- g(r) / S(q) --------------------------------------------------------------------------
  int gr_nbin = 0, gr_nbin_inner = 0;
  float gr_dr = 0, gr_rdr = 0, gr_rmaxsq = 0, gr_rmaxsq_inner = 0;
  unsigned long long* gr_count = nullptr;  // [npair][gr_nbin]
  unsigned long long* gr_shell = nullptr;  // [copies][npair][gr_nbin]: outer-shell adds of a fused tally
  double* gr_h = nullptr;
  double* gr_norm = nullptr;
  int sq_nq = 0;
  double sq_qmin = 0, sq_qmax = 0;
  double* sq_out = nullptr;
  double* sq_ftmx = nullptr;  // [nq][nbin]
  std::vector<int> h_typecount;

  // ---- NNP: descriptors G and dE/dG by row, per-atom network energies ---------------------
  NnpState* nnp = nullptr;
  MlpState* mlp = nullptr;
  RmdfState* rmdf = nullptr;
  SqdState* sqd = nullptr;
  float* G = nullptr;
  float* dG = nullptr;
  int dG_kind = 0;  // precision form of the network that wrote dG (0: FP32 or given by the caller)
  int D_cur = 0, D_cap = 0;
  float* e_atom = nullptr;              // [natom] by row
  int* type_idx = nullptr;              // rows grouped by element: [ntype][typecount]
  int* type_fill = nullptr;             // [ntype]
  std::vector<int> h_type_off;          // element offsets into type_idx

  // ---- MD driver -----------------------------------------------------------------------------
  long long md_step = 0;                // steps completed by nnpe_md_run / the step calls
  void* md = nullptr;                   // captured step graphs (md.hip)
  bool md_fork = true;                  // NNP steps: classical terms on stream2 beside the NNP passes
  bool md_fork_early = true;            // ... branching before the descriptors (else after them)
  bool md_fuse_x = true;                // captured intervals: position update in the NNP force pass
  double* md_log = nullptr;             // device log records
  int md_log_cap = 0;
};

// enqueue helpers (no host synchronisation; safe inside stream capture)
// skin_accum: the MD step's skin check ran before this rebuild; its flag is added to the
// skin-violation count and cleared by the rebuild's reset kernel.  defer_shell: a g(r) tally
// folded into the rebuild leaves its shell pass and fold (gr_shell_finish_enqueue) to the caller
int neigh_enqueue_build(nnpe_ctx* ctx, bool gr_tally = false, bool skin_accum = false, bool defer_shell = false);
int pair_enqueue(nnpe_ctx* ctx, bool ev);
int bonded_enqueue(nnpe_ctx* ctx, bool ev);
int verlet_enqueue_x(nnpe_ctx* ctx);
// ... with the MD step's skin check of the updated positions against those of the last build
int verlet_enqueue_x_skin(nnpe_ctx* ctx);
int verlet_enqueue_v(nnpe_ctx* ctx);
int ke_enqueue(nnpe_ctx* ctx);
int gr_enqueue(nnpe_ctx* ctx, bool full);
// full g(r) tally folded into a list rebuild (neigh.hip, grsq.hip): the rebuild counts the pairs
// of its candidates, the outer cells follow in a shell pass
bool gr_fusable(nnpe_ctx* ctx);
// private copies of the full tally's counts that a tally folded into a rebuild adds into (the
// rebuild kernel's workgroup histograms and the shell pass), summed by the fold kernel
constexpr int kGrShellCopies = 32;
int gr_shell_size(nnpe_ctx* ctx);  // entries of the shell copies (cleared by the rebuild's reset)
int gr_outer_refresh(nnpe_ctx* ctx);  // device outer stencil for the current grid and g(r) range
int gr_shell_finish_enqueue(nnpe_ctx* ctx);
int sq_enqueue(nnpe_ctx* ctx);
int nnp_tbl_desc_enqueue(nnpe_ctx* ctx);
int nnp_tbl_force_enqueue(nnpe_ctx* ctx, bool ev);
int sf_desc_enqueue(nnpe_ctx* ctx);
int sf_force_enqueue(nnpe_ctx* ctx, bool ev);
int sf_adjoint_enqueue(nnpe_ctx* ctx);
// kick: fused FP64 Verlet half-kick; adv (with kick): 1 the next step's position update, 2 the
// positions copied, into ctx->xq_alt
int sf_force_only_enqueue(nnpe_ctx* ctx, bool ev, bool kick, int adv = 0);
int mlp_enqueue(nnpe_ctx* ctx, bool ev);
void md_state_free(nnpe_ctx* ctx);
int neigh_check_and_grow(nnpe_ctx* ctx, bool* rebuilt);
int verlet_switch_precision(nnpe_ctx* ctx, bool fp64);

// NNP helpers shared by nnp.hip / mlp.hip / md.hip
NnpState* nnp_state_of(nnpe_ctx* ctx);
void nnp_state_free(nnpe_ctx* ctx);
void mlp_state_free(nnpe_ctx* ctx);
int nnp_desc_reserve(nnpe_ctx* ctx, int D);
int sf_dim(int ntype);
int type_index_enqueue(nnpe_ctx* ctx);
// element offsets into type_idx on the device when the network keeps a row index per element
// (the rebuild then fills type_idx in its copy-back kernel), else nullptr
const int* mlp_type_off(nnpe_ctx* ctx);
// NNP pair tables being refined: force / potential [npair][4][nbin+1]; nonzero if not set
int nnp_tbl_tables(nnpe_ctx* ctx, float** ftab, float** utab, int* nbin, float* dr);
// symmetry-function active lists sized to the short lists' capacity
int nnp_act_reserve(nnpe_ctx* ctx);

// direct k-space S(q) (sqd.hip)
void sqd_state_free(nnpe_ctx* ctx);

// RMDF (rmdf.hip)
void rmdf_state_free(nnpe_ctx* ctx);
int rmdf_enqueue(nnpe_ctx* ctx, bool potential);
bool rmdf_active(nnpe_ctx* ctx);
float rmdf_gamma(nnpe_ctx* ctx);

// ---- helpers -------------------------------------------------------------------------------
int ctx_fail(nnpe_ctx* ctx, int code, const std::string& msg);
#define NNPE_CHECK_HIP(ctx, call)                                                               \
  do {                                                                                          \
    hipError_t _e = (call);                                                                     \
    if (_e != hipSuccess) return ctx_fail((ctx), NNPE_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(_e)); \
  } while (0)
#define NNPE_CHECK_LAUNCH(ctx) NNPE_CHECK_HIP(ctx, hipGetLastError())

template <class T>
inline hipError_t dev_alloc(T*& p, size_t n) {
  if (p) { (void)hipFree(p); p = nullptr; }
  if (n == 0) n = 1;
  return hipMalloc(reinterpret_cast<void**>(&p), n * sizeof(T));
}
template <class T>
inline void dev_free(T*& p) {
  if (p) { (void)hipFree(p); p = nullptr; }
}

inline int blocks_for(long long n, int bs) { return int((n + bs - 1) / bs); }

// image code -> shift component, as the reference writes the ghost shift vector
__device__ __forceinline__ float image_shift(int c, float bl, float m_bl) {
  return c == 1 ? 0.0f : (c > 1 ? bl : m_bl);
}
// code = 9a + 3b + c with a, b, c in 0..2, decoded by exact multiply-shift divisions on the
// code range 0..26 (a = code/9 = (57 code) >> 9, b = rem/3 = (11 rem) >> 5 for rem <= 8)
__device__ __forceinline__ float4 image_pos(float4 xs, int code, float bl, float m_bl) {
  const unsigned u = (unsigned)code;
  const unsigned a = __umul24(u, 57u) >> 9, rem = u - __umul24(9u, a);
  const unsigned b = __umul24(rem, 11u) >> 5, c = rem - __umul24(3u, b);
  float4 r = xs;
  r.x = __fadd_rn(xs.x, image_shift((int)a, bl, m_bl));
  r.y = __fadd_rn(xs.y, image_shift((int)b, bl, m_bl));
  r.z = __fadd_rn(xs.z, image_shift((int)c, bl, m_bl));
  return code == 13 ? xs : r;
}
// (a-b)^2 summed left to right with IEEE single rounding and no contraction, as the
// reference's sqr(dx)+sqr(dy)+sqr(dz) (src/neighlist.cpp:260)
// The same images through a 27-entry LDS table of shift multipliers (code 9a + 3b + c ->
// (a-1, b-1, c-1)): x + m L as one fused multiply-add, exact as the sum above because m L is
// -L, 0 or L.  (Code 13 gives x + 0, which differs from x only in the sign of a zero
// coordinate; no distance or force depends on that.)  The pair-loop kernels fill the table once
// per workgroup and save the decode of the code on every list entry.
__device__ __forceinline__ void image_table_fill(float4* t) {
  for (int k = threadIdx.x; k < 27; k += blockDim.x)
    t[k] = make_float4(float(k / 9 - 1), float((k / 3) % 3 - 1), float(k % 3 - 1), 0.0f);
}
__device__ __forceinline__ float4 image_pos_t(float4 xs, int code, float bl, const float4* t) {
  const float4 m = t[code];
  return make_float4(__fmaf_rn(m.x, bl, xs.x), __fmaf_rn(m.y, bl, xs.y), __fmaf_rn(m.z, bl, xs.z), xs.w);
}

__device__ __forceinline__ float dist2_rn(float dx, float dy, float dz) {
  return __fadd_rn(__fadd_rn(__fmul_rn(dx, dx), __fmul_rn(dy, dy)), __fmul_rn(dz, dz));
}

// ---- lanes-per-atom groups ------------------------------------------------------------------
// The pair-loop kernels give each atom a group of G lanes (G divides 64): lane l of the group
// takes entries l, l+G, l+2G, ... of the atom's list row, so a group reads G consecutive 4-byte
// entries per step, and the group's partial sums are combined by a fixed xor butterfly (every
// lane ends with the same, order-independent-of-timing total).
template <int CTRL>
__device__ __forceinline__ float dpp_f32(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
template <int G, class T>
__device__ __forceinline__ T group_sum(T v) {
  if constexpr (G == 16 && std::is_same<T, float>::value) {
    // DPP inside each 16-lane row (no LDS traffic): rotate by 8 and by 4 (row_ror), then swap
    // pairs at distance 2 and 1 (quad_perm).  After the rotate by 8 each lane holds a_l + a_(l^8),
    // which has period 8, so the rotate by 4 reads the same value as lane l^4 and so on: every
    // step adds the same operands in the same order as the xor butterfly below (bit-identical).
    v += dpp_f32<0x128>(v);
    v += dpp_f32<0x124>(v);
    v += dpp_f32<0x4E>(v);
    v += dpp_f32<0xB1>(v);
    return v;
  } else {
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, G);
    return v;
  }
}

// group_sum of N values at once, step by step across the values (the same sums, bit for bit,
// as N separate group_sum calls; independent DPP steps fill each other's hazard slots)
template <int G, int N>
__device__ __forceinline__ void group_sum_n(float* v) {
  if constexpr (G == 16) {
#pragma unroll
    for (int k = 0; k < N; ++k) v[k] += dpp_f32<0x128>(v[k]);
#pragma unroll
    for (int k = 0; k < N; ++k) v[k] += dpp_f32<0x124>(v[k]);
#pragma unroll
    for (int k = 0; k < N; ++k) v[k] += dpp_f32<0x4E>(v[k]);
#pragma unroll
    for (int k = 0; k < N; ++k) v[k] += dpp_f32<0xB1>(v[k]);
  } else {
#pragma unroll
    for (int k = 0; k < N; ++k) v[k] = group_sum<G>(v[k]);
  }
}

// Exclusive prefix sum of v[0..n) in place by the whole workgroup of NT threads (each thread
// scans a contiguous chunk, wave totals through LDS); v[n] receives the total, which is also
// returned.  s_wsum: NT/64 + 1 ints of LDS.  Ends with a barrier.
template <int NT>
__device__ __forceinline__ int block_scan_excl(int* v, int n, int* s_wsum) {
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int k = (n + NT - 1) / NT;
  const int b = min(tid * k, n), e = min(b + k, n);
  int sum = 0;
  for (int q = b; q < e; ++q) sum += v[q];
  int inc = sum;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int t = __shfl_up(inc, o, 64);
    if (lane >= o) inc += t;
  }
  if (lane == 63) s_wsum[w] = inc;
  __syncthreads();
  if (tid == 0) {
    int run = 0;
    for (int q = 0; q < NT / 64; ++q) { const int t = s_wsum[q]; s_wsum[q] = run; run += t; }
    s_wsum[NT / 64] = run;
  }
  __syncthreads();
  int run = s_wsum[w] + inc - sum;
  for (int q = b; q < e; ++q) { const int t = v[q]; v[q] = run; run += t; }
  const int total = s_wsum[NT / 64];
  if (tid == 0) v[n] = total;
  __syncthreads();
  return total;
}

// Workgroups are dealt to the 8 XCDs round-robin by workgroup id.  Renumber them so each XCD
// walks one contiguous range of blocks (rows are cell-sorted, so its atoms are spatially close
// and their neighbours' rows stay in that XCD's L2).
__device__ __forceinline__ int xcd_block(int nblocks) {
  const int b = blockIdx.x, x = b & 7, k = b >> 3, q = nblocks >> 3, r = nblocks & 7;
  return x * q + min(x, r) + k;
}

// neighbour list rows: entry k of row i at nbr[i * cap + k]
__device__ __forceinline__ const uint32_t* list_row(const uint32_t* nbr, int cap, int i) {
  return nbr + (size_t)i * cap;
}

// g(r) bin of a squared distance, int(sqrt(rsq) * 1/dr) in FP32 as the reference
// (src/grsq.cpp:118-121): correctly rounded sqrt, one rounded product
__device__ __forceinline__ int gr_bin_of(float rsq, float rdr) { return int(__fmul_rn(sqrtf(rsq), rdr)); }

__device__ __forceinline__ int box_cell_coord(float g_offset, float x, float g_rdx) {
  return int(__fmul_rn(__fadd_rn(g_offset, x), g_rdx));
}

// energy/virial: block reduction of one double per thread into ev_partial[slot][blockIdx]
template <int BS>
__device__ __forceinline__ void block_sum2_store(double a, double b, double* part_a, double* part_b) {
  __shared__ double sa[BS / 64], sb[BS / 64];
  for (int o = 32; o > 0; o >>= 1) {
    a += __shfl_down(a, o, 64);
    b += __shfl_down(b, o, 64);
  }
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) { sa[w] = a; sb[w] = b; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double ta = 0, tb = 0;
    for (int k = 0; k < BS / 64; ++k) { ta += sa[k]; tb += sb[k]; }
    part_a[blockIdx.x] = ta;
    part_b[blockIdx.x] = tb;
  }
}

// host-side entry points shared between translation units
hipError_t ev_reserve(nnpe_ctx* ctx, int nblocks);
// Velocity Verlet's second half-kick in FP64 for atom i (VERLET::updateV, src/verlet.cpp:29-97,
// FP64 form): a = f/m * accelF, v += (a_prev + a) * dt/2, a_prev <- a; returns a as the FP32
// force row (the force array holds the acceleration afterwards, as in the reference).  Shared by
// the Verlet kernel and the force kernel that finishes an atom's force, so both round alike.
__device__ __forceinline__ float4 vv_kick64(int i, double dt, const float* rm, double4* vd, double4* ad, float4 ff) {
#pragma clang fp contract(off)
  const double r = double(rm[i]) * 4.184e-4;
  const double4 ap = ad[i];
  const double4 an = make_double4(double(ff.x) * r, double(ff.y) * r, double(ff.z) * r, 0);
  double4 vv = vd[i];
  const double hdt = 0.5 * dt;
  vv.x += (ap.x + an.x) * hdt;
  vv.y += (ap.y + an.y) * hdt;
  vv.z += (ap.z + an.z) * hdt;
  vd[i] = vv;
  ad[i] = an;
  return make_float4(float(an.x), float(an.y), float(an.z), 0);
}

// arguments of a fused FP64 half-kick
struct VvArgs {
  double dt;
  const float* rm;
  double4* vd;
  double4* ad;
  // after the kick, the next step's position update of the same atom (adv 1: x += (v + a dt/2) dt
  // into xd and, as FP32, into xq_out), or (adv 2) this step's positions copied into xq_out
  double4* xd = nullptr;
  float4* xq_out = nullptr;
  int adv = 0;
};

// vv_kick64, then the next step's position update (the arithmetic of the separate update kernel,
// x += (v + a dt/2) dt, on the values just written), or the current position copied, into xq_out
__device__ __forceinline__ float4 vv_kick64_adv(int i, const VvArgs& V, float4 ff, float4 xcur) {
#pragma clang fp contract(off)
  const double r = double(V.rm[i]) * 4.184e-4;
  const double4 ap = V.ad[i];
  const double4 an = make_double4(double(ff.x) * r, double(ff.y) * r, double(ff.z) * r, 0);
  double4 vv = V.vd[i];
  const double hdt = 0.5 * V.dt;
  vv.x += (ap.x + an.x) * hdt;
  vv.y += (ap.y + an.y) * hdt;
  vv.z += (ap.z + an.z) * hdt;
  V.vd[i] = vv;
  V.ad[i] = an;
  if (V.adv == 1) {
    double4 x = V.xd[i];
    x.x += (vv.x + an.x * hdt) * V.dt;
    x.y += (vv.y + an.y * hdt) * V.dt;
    x.z += (vv.z + an.z * hdt) * V.dt;
    V.xd[i] = x;
    xcur.x = float(x.x); xcur.y = float(x.y); xcur.z = float(x.z);
  }
  V.xq_out[i] = xcur;
  return make_float4(float(an.x), float(an.y), float(an.z), 0);
}

int ev_finish(nnpe_ctx* ctx, int term, int nblocks);  // add partials of a term into ev_sum
// add per-block partials held elsewhere (the same fixed-order reduction) into ev_sum
int ev_finish_from(nnpe_ctx* ctx, int term, const double* part_e, const double* part_w, int nblocks);
// the allocation-free bonded entry point: address of its device array on the context's device
// (once per device, at context creation; forces.hip)
int bonded_scratch_setup(nnpe_ctx* ctx);
double* ev_part(nnpe_ctx* ctx, int term, int which);  // partial array for term (0 energy, 1 virial)
int counters_read(nnpe_ctx* ctx);
// upload ctx->h_box and a cell -> box map for the single-box layout
int box_upload_single(nnpe_ctx* ctx);
// the cells that can hold atoms of the boxes in h_box (their home ranges), the list kernel's grid
int home_cells_upload(nnpe_ctx* ctx);
// grid and ghost parameters of one box of edge bl (cell_base 0)
BoxPar box_geometry(float bl, float pair_cutoff, float skin, float gr_cutoff);
// MLP tiles from ctx->h_typecount (after the element counts change)
int mlp_retile(nnpe_ctx* ctx);
// position update skipping the rows of boxes flagged in box_skip (device, per box)
int verlet_enqueue_x_skip(nnpe_ctx* ctx, const unsigned char* box_skip);
