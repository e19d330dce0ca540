// serve.hip -- the serving form: a request loop over independent periodic boxes.
//
// (No reference counterpart: the reference runs one system per process.)  Requests are queued
// with their own topology and step count.  Admitted boxes live side by side in one resident
// arena -- a context whose rows, cells and ghosts belong to many boxes (per-row box index,
// per-box L and cell grid) -- so every launch of the MD step (list build, pair, bonded, NNP
// descriptors, MLP tiles, NNP forces, Verlet) advances all resident boxes at once and fills
// the GPU even when each box is small.  The arena advances in chunks of `rebuild_every` steps
// (the first step of a chunk rebuilds the lists; the rest replay captured hipGraphs).  Between
// chunks, finished boxes leave (their final state is copied out) and queued boxes are admitted
// first-in first-out while the arena's atom capacity -- its preallocated HBM footprint --
// allows; the surviving boxes' rows are moved on the device (no host round trip) and the new
// boxes' rows uploaded.  A box's per-row arithmetic does not depend on its neighbours in the
// arena (rows, lists, MLP rows and reductions are per atom), so its trajectory is the same alone
// and in any mix.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <deque>
#include <map>
#include <vector>

#include "internal.h"

namespace {

struct SReq {
  long long id = 0;
  int natom = 0, nmol = 0, nsteps = 0, done = 0;
  float bl = 0;
  std::vector<float> x, v, q, mass;
  std::vector<int> typ, mol, intra, moltype;
  std::vector<int> terms[3];  // bonds (t,i,j), angles (t,i,j,k), dihedrals (t,i,j,k,l), local atom ids
  std::vector<int> mol_first_local;
  // placement in the arena
  int box = -1, row_base = 0, mol_base = 0;
  bool fresh = true;
  std::vector<double> xo, vo;
};

// a run of rows of one surviving box: new rows [nb, nb + n) come from old rows [ob, ob + n)
struct Seg {
  int nb, ob, n, box, d_atom, d_mol, pad0, pad1;
};

__global__ void k_relayout(int nrows, int nseg, const Seg* __restrict__ seg, bool fp64, const float4* xq, const float4* v,
                           const float4* f, const float4* fprev, const float* rm, const int* typ, const int4* info,
                           const double4* xd, const double4* vd, const double4* ad, float4* t_xq, float4* t_v,
                           float4* t_f, float4* t_fprev, float* t_rm, int* t_typ, int4* t_info, int* t_rbox, double4* t_xd,
                           double4* t_vd, double4* t_ad) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  int lo = 0, hi = nseg - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (seg[mid].nb <= r) lo = mid; else hi = mid - 1;
  }
  const Seg S = seg[lo];
  const int o = S.ob + (r - S.nb);
  t_xq[r] = xq[o]; t_v[r] = v[o]; t_f[r] = f[o]; t_fprev[r] = fprev[o]; t_rm[r] = rm[o]; t_typ[r] = typ[o];
  int4 in = info[o];
  in.x += S.d_atom;
  in.y += S.d_mol;
  t_info[r] = in;
  t_rbox[r] = S.box;
  if (fp64) { t_xd[r] = xd[o]; t_vd[r] = vd[o]; t_ad[r] = ad[o]; }
}

__global__ void k_rows_back(int nrows, bool fp64, const float4* t_xq, const float4* t_v, const float4* t_f,
                            const float4* t_fprev, const float* t_rm, const int* t_typ, const int4* t_info, const int* t_rbox,
                            const double4* t_xd, const double4* t_vd, const double4* t_ad, float4* xq, float4* v, float4* f,
                            float4* fprev, float* rm, int* typ, int4* info, int* rbox, double4* xd, double4* vd, double4* ad) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  xq[r] = t_xq[r]; v[r] = t_v[r]; f[r] = t_f[r]; fprev[r] = t_fprev[r]; rm[r] = t_rm[r]; typ[r] = t_typ[r];
  info[r] = t_info[r]; rbox[r] = t_rbox[r];
  if (fp64) { xd[r] = t_xd[r]; vd[r] = t_vd[r]; ad[r] = t_ad[r]; }
}

template <class T>
hipError_t grow_alloc(T*& p, size_t& cap, size_t n) {
  if (p && n <= cap) return hipSuccess;
  cap = std::max(n, cap + cap / 4);
  return dev_alloc(p, cap);
}

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

struct nnpe_serve {
  nnpe_ctx* ctx = nullptr;
  float pair_cutoff = 0, skin = 0, nnp_cutoff = 0;
  int rebuild_every = 10, max_atoms = 0, max_boxes = 0, ntype = 0;
  bool fp64 = true;
  double dt = 0.5;
  std::vector<int> mt_natom;
  std::vector<float> coef[3];
  int ntyp[3] = {0, 0, 0};
  std::deque<SReq*> queue;
  std::vector<SReq*> res;
  std::map<long long, SReq*> done;
  long long next_id = 1, chunks = 0, relayouts = 0;
  size_t arena_bytes = 0;
  Seg* d_seg = nullptr;
  size_t seg_cap = 0;
  unsigned char* d_skip = nullptr;
  size_t skip_cap = 0;
  size_t cell_cap = 0, ghost_cap = 0;
  nnpe_md_params P{};
};

static int serve_fail(nnpe_serve* s, int code, const char* msg) { return ctx_fail(s ? s->ctx : nullptr, code, msg); }

extern "C" int nnpe_serve_create(const nnpe_serve_config* c, nnpe_serve** out) {
  if (!c || !out || c->max_atoms <= 0 || c->max_boxes <= 0 || c->ntype <= 0 || c->nmoltype <= 0 || !c->mt_natom ||
      !c->mt_special_off || !c->mt_special || !c->pair_force || !c->pair_potential || c->pair_nbin <= 0 ||
      c->rebuild_every <= 0 || !c->widths || !c->params || !c->sf_mu || !c->sf_eta)
    return NNPE_ERR_ARG;
  *out = nullptr;
  nnpe_serve* s = new nnpe_serve;
  int rc = nnpe_create(&s->ctx, c->device);
  if (rc) { delete s; return rc; }
  nnpe_ctx* ctx = s->ctx;
  s->pair_cutoff = c->pair_cutoff; s->skin = c->skin; s->nnp_cutoff = c->nnp_cutoff;
  s->rebuild_every = c->rebuild_every; s->max_atoms = c->max_atoms; s->max_boxes = c->max_boxes; s->ntype = c->ntype;
  s->fp64 = c->fp64 != 0; s->dt = c->dt;
  s->mt_natom.assign(c->mt_natom, c->mt_natom + c->nmoltype);
  const float* cf[3] = {c->bond_coef, c->angle_coef, c->dihedral_coef};
  const int nt[3] = {c->nbond_types, c->nangle_types, c->ndihedral_types};
  for (int k = 0; k < 3; ++k) {
    s->ntyp[k] = nt[k];
    if (nt[k] > 0 && cf[k]) s->coef[k].assign(cf[k], cf[k] + size_t(4) * nt[k]);
  }
  size_t free0 = 0, total = 0;
  (void)hipMemGetInfo(&free0, &total);
  // arena at capacity: max_atoms rows (one placeholder molecule per row until boxes arrive), a
  // placeholder box at liquid density so the list capacities start at a realistic size
  const int N = c->max_atoms;
  const float bl0 = std::max(float(std::cbrt(N / 0.1)), 2.5f * (c->skin + c->pair_cutoff));
  std::vector<float> x(size_t(3) * N, 0.0f), q(N, 0.0f), m(N, 1.0f);
  std::vector<int> typ(N, 0), mid(N), intra(N, 0), mtype(N, 0);
  for (int i = 0; i < N; ++i) mid[i] = i;
  rc = nnpe_system_set(ctx, N, c->ntype, bl0, x.data(), nullptr, q.data(), m.data(), typ.data(), mid.data(), intra.data(), N,
                       mtype.data(), c->nmoltype, c->mt_natom, c->mt_special_off, c->mt_special);
  if (!rc) rc = nnpe_neigh_init(ctx, c->pair_cutoff, c->skin, 0.0f, c->nnp_cutoff);
  if (!rc) rc = nnpe_pair_table(ctx, c->pair_cutoff, c->pair_nbin, c->pair_force, c->pair_potential);
  for (int k = 0; k < 3 && !rc; ++k)
    if (s->ntyp[k] > 0) rc = nnpe_bonded_set(ctx, k, s->ntyp[k], s->coef[k].data(), 0, nullptr);
  if (!rc) rc = nnpe_sf_set(ctx, c->nnp_cutoff, 16, c->sf_mu, c->sf_eta_r, 4, c->sf_eta);
  if (!rc) rc = nnpe_mlp_set(ctx, c->nlayer, c->widths, c->params, c->out_scale, c->precision, c->in_shift, c->in_scale);
  if (!rc) rc = nnpe_verlet_set(ctx, c->dt, c->fp64);
  if (rc) {
    nnpe_destroy(ctx);
    delete s;
    return rc;
  }
  s->cell_cap = size_t(ctx->num_grid);
  s->ghost_cap = size_t(ctx->ghost_cap);
  size_t free1 = 0;
  (void)hipMemGetInfo(&free1, &total);
  s->arena_bytes = free0 > free1 ? free0 - free1 : 0;
  ctx->natom = 0;
  ctx->nmol = 0;
  ctx->nbox = 0;
  ctx->have_list = false;
  s->P.dt = c->dt; s->P.fp64 = c->fp64; s->P.rebuild_every = c->rebuild_every;
  s->P.pair = 1; s->P.bonded = 1; s->P.nnp = 2; s->P.use_graph = 1;
  *out = s;
  return NNPE_OK;
}

extern "C" int nnpe_serve_destroy(nnpe_serve* s) {
  if (!s) return NNPE_OK;
  for (auto* r : s->queue) delete r;
  for (auto* r : s->res) delete r;
  for (auto& kv : s->done) delete kv.second;
  if (s->ctx) {
    (void)hipSetDevice(s->ctx->device);
    dev_free(s->d_seg);
    dev_free(s->d_skip);
    nnpe_destroy(s->ctx);
  }
  delete s;
  return NNPE_OK;
}

extern "C" int nnpe_serve_submit(nnpe_serve* s, int natom, float bl, const float* x, const float* v, const float* q,
                                 const float* mass, const int* typ, const int* mol_id, const int* intra_id, int nmol,
                                 const int* mol_type, int nbond, const int* bonds, int nangle, const int* angles, int ndih,
                                 const int* dihedrals, int nsteps, long long* id) {
  if (!s || natom <= 0 || nmol <= 0 || !x || !q || !mass || !typ || !mol_id || !intra_id || !mol_type || !id || nsteps <= 0)
    return serve_fail(s, NNPE_ERR_ARG, "nnpe_serve_submit: bad argument");
  if (nsteps % s->rebuild_every) return serve_fail(s, NNPE_ERR_ARG, "nsteps must be a multiple of the chunk length");
  if (natom > s->max_atoms) return serve_fail(s, NNPE_ERR_ARG, "box larger than the arena");
  const float comm = s->skin + s->pair_cutoff;
  if (!(bl > comm)) return serve_fail(s, NNPE_ERR_UNSUPPORTED, "box too small: L must exceed skin + rc (one image shell)");
  SReq* r = new SReq;
  r->natom = natom; r->nmol = nmol; r->nsteps = nsteps; r->bl = bl;
  r->x.assign(x, x + size_t(3) * natom);
  r->v.assign(size_t(3) * natom, 0.0f);
  if (v) r->v.assign(v, v + size_t(3) * natom);
  r->q.assign(q, q + natom); r->mass.assign(mass, mass + natom); r->typ.assign(typ, typ + natom);
  r->mol.assign(mol_id, mol_id + natom); r->intra.assign(intra_id, intra_id + natom);
  r->moltype.assign(mol_type, mol_type + nmol);
  r->mol_first_local.assign(nmol, -1);
  for (int i = 0; i < natom; ++i) {
    if (typ[i] < 0 || typ[i] >= s->ntype || mol_id[i] < 0 || mol_id[i] >= nmol) {
      delete r;
      return serve_fail(s, NNPE_ERR_ARG, "atom type / molecule id out of range");
    }
    if (intra_id[i] == 0) r->mol_first_local[mol_id[i]] = i;
  }
  for (int m2 = 0; m2 < nmol; ++m2)
    if (r->mol_first_local[m2] < 0 || mol_type[m2] < 0 || mol_type[m2] >= int(s->mt_natom.size())) {
      delete r;
      return serve_fail(s, NNPE_ERR_ARG, "molecule without an intra-0 atom or with a bad type");
    }
  const int nterm[3] = {nbond, nangle, ndih};
  const int* tp[3] = {bonds, angles, dihedrals};
  for (int k = 0; k < 3; ++k) {
    if (nterm[k] < 0 || (nterm[k] > 0 && !tp[k])) { delete r; return serve_fail(s, NNPE_ERR_ARG, "bonded terms missing"); }
    const int w = 3 + k;
    int prev = -1;  // terms grouped by molecule, ascending
    for (int e = 0; e < nterm[k]; ++e) {
      const int* p = tp[k] + size_t(w) * e;
      bool ok = p[0] >= 0 && p[0] < s->ntyp[k] && p[1] >= 0 && p[1] < natom;
      for (int a = 2; a < w && ok; ++a) ok = p[a] >= 0 && p[a] < natom && mol_id[p[a]] == mol_id[p[1]];
      ok = ok && mol_id[p[1]] >= prev;
      if (!ok) { delete r; return serve_fail(s, NNPE_ERR_ARG, "bonded term with a bad type, atoms outside one molecule or out of molecule order"); }
      prev = mol_id[p[1]];
    }
    if (nterm[k] > 0) r->terms[k].assign(tp[k], tp[k] + size_t(w) * nterm[k]);
  }
  r->id = s->next_id++;
  *id = r->id;
  s->queue.push_back(r);
  return NNPE_OK;
}

// Lay the resident boxes out again: survivors keep their device rows (moved), new boxes are
// uploaded; per-box grids, element counts, MLP tiles, molecule table and bonded terms follow.
static int serve_relayout(nnpe_serve* s, const std::vector<SReq*>& keep, const std::vector<SReq*>& add) {
  nnpe_ctx* ctx = s->ctx;
  md_state_free(ctx);
  NNPE_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
  const bool fp64 = ctx->fp64;
  // placement
  std::vector<SReq*> all(keep);
  all.insert(all.end(), add.begin(), add.end());
  std::vector<Seg> segs;
  int row = 0, mol = 0, cell = 0;
  ctx->h_box.assign(all.size(), BoxPar{});
  ctx->h_typecount.assign(ctx->ntype, 0);
  double rho_max = 0;
  double ghost_est = 0;
  const float comm = s->skin + s->pair_cutoff;
  for (size_t b = 0; b < all.size(); ++b) {
    SReq* r = all[b];
    BoxPar g = box_geometry(r->bl, s->pair_cutoff, s->skin, 0.0f);
    g.cell_base = cell;
    ctx->h_box[b] = g;
    if (b < keep.size()) {
      Seg sg{};
      sg.nb = row; sg.ob = r->row_base; sg.n = r->natom; sg.box = int(b);
      sg.d_atom = row - r->row_base; sg.d_mol = mol - r->mol_base;
      segs.push_back(sg);
    }
    r->box = int(b); r->row_base = row; r->mol_base = mol;
    row += r->natom; mol += r->nmol;
    cell += g.num_gx * g.num_gx * g.num_gx;
    for (int t : r->typ) ctx->h_typecount[t]++;
    const double vol = double(r->bl) * r->bl * r->bl;
    rho_max = std::max(rho_max, r->natom / vol);
    const double vout = double(r->bl) + 2.0 * comm;
    ghost_est += std::min(26.0 * r->natom, 1.3 * r->natom * (vout * vout * vout - vol) / vol);
  }
  const int nkeep_rows = keep.empty() ? 0 : keep.back()->row_base + keep.back()->natom;
  // move the survivors' rows into the scratch rows, then upload the new boxes there
  if (!segs.empty()) {
    NNPE_CHECK_HIP(ctx, grow_alloc(s->d_seg, s->seg_cap, segs.size()));
    NNPE_CHECK_HIP(ctx, hipMemcpy(s->d_seg, segs.data(), sizeof(Seg) * segs.size(), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_relayout, dim3(blocks_for(nkeep_rows, 256)), dim3(256), 0, ctx->stream, nkeep_rows, int(segs.size()),
                       s->d_seg, fp64, ctx->xq, ctx->v, ctx->f, ctx->fprev, ctx->rm, ctx->typ, ctx->info, ctx->xd, ctx->vd,
                       ctx->ad, ctx->t_xq, ctx->t_v, ctx->t_xprev, ctx->t_fprev, ctx->t_rm, ctx->t_typ, ctx->t_info,
                       ctx->t_rbox, ctx->t_xd, ctx->t_vd, ctx->t_ad);
    NNPE_CHECK_LAUNCH(ctx);
  }
  if (!add.empty()) {
    const int n = row - nkeep_rows;
    std::vector<float4> hx(n), hv(n), hz(n, make_float4(0, 0, 0, 0));
    std::vector<double4> hxd(n), hvd(n), hzd(n, make_double4(0, 0, 0, 0));
    std::vector<float> hrm(n);
    std::vector<int> htyp(n), hbox(n);
    std::vector<int4> hinfo(n);
    int k = 0;
    for (SReq* r : add)
      for (int i = 0; i < r->natom; ++i, ++k) {
        hx[k] = make_float4(r->x[3 * i], r->x[3 * i + 1], r->x[3 * i + 2], r->q[i]);
        hv[k] = make_float4(r->v[3 * i], r->v[3 * i + 1], r->v[3 * i + 2], 0);
        hxd[k] = make_double4(r->x[3 * i], r->x[3 * i + 1], r->x[3 * i + 2], 0);
        hvd[k] = make_double4(r->v[3 * i], r->v[3 * i + 1], r->v[3 * i + 2], 0);
        hrm[k] = 1.0f / r->mass[i];
        htyp[k] = r->typ[i];
        hinfo[k] = make_int4(r->row_base + i, r->mol_base + r->mol[i], r->intra[i], r->moltype[r->mol[i]]);
        hbox[k] = r->box;
      }
    const size_t o = nkeep_rows;
    NNPE_CHECK_HIP(ctx, hipMemcpy(ctx->t_xq + o, hx.data(), sizeof(float4) * n, hipMemcpyHostToDevice));
    NNPE_CHECK_HIP(ctx, hipMemcpy(ctx->t_v + o, hv.data(), sizeof(float4) * n, hipMemcpyHostToDevice));
    NNPE_CHECK_HIP(ctx, hipMemcpy(ctx->t_xprev + o, hz.data(), sizeof(float4) * n, hipMemcpyHostToDevice));
    NNPE_CHECK_HIP(ctx, hipMemcpy(ctx->t_fprev + o, hz.data(), sizeof(float4) * n, hipMemcpyHostToDevice));
    NNPE_CHECK_HIP(ctx, hipMemcpy(ctx->t_rm + o, hrm.data(), sizeof(float) * n, hipMemcpyHostToDevice));
    NNPE_CHECK_HIP(ctx, hipMemcpy(ctx->t_typ + o, htyp.data(), sizeof(int) * n, hipMemcpyHostToDevice));
    NNPE_CHECK_HIP(ctx, hipMemcpy(ctx->t_info + o, hinfo.data(), sizeof(int4) * n, hipMemcpyHostToDevice));
    NNPE_CHECK_HIP(ctx, hipMemcpy(ctx->t_rbox + o, hbox.data(), sizeof(int) * n, hipMemcpyHostToDevice));
    if (fp64) {
      NNPE_CHECK_HIP(ctx, hipMemcpy(ctx->t_xd + o, hxd.data(), sizeof(double4) * n, hipMemcpyHostToDevice));
      NNPE_CHECK_HIP(ctx, hipMemcpy(ctx->t_vd + o, hvd.data(), sizeof(double4) * n, hipMemcpyHostToDevice));
      NNPE_CHECK_HIP(ctx, hipMemcpy(ctx->t_ad + o, hzd.data(), sizeof(double4) * n, hipMemcpyHostToDevice));
    }
  }
  if (row > 0) {
    hipLaunchKernelGGL(k_rows_back, dim3(blocks_for(row, 256)), dim3(256), 0, ctx->stream, row, fp64, ctx->t_xq, ctx->t_v,
                       ctx->t_xprev, ctx->t_fprev, ctx->t_rm, ctx->t_typ, ctx->t_info, ctx->t_rbox, ctx->t_xd, ctx->t_vd,
                       ctx->t_ad, ctx->xq, ctx->v, ctx->f, ctx->fprev, ctx->rm, ctx->typ, ctx->info, ctx->rbox, ctx->xd, ctx->vd,
                       ctx->ad);
    NNPE_CHECK_LAUNCH(ctx);
  }
  ctx->natom = row;
  ctx->nmol = mol;
  ctx->nbox = int(all.size());
  // boxes, cells, ghosts, list capacities
  if (!ctx->h_box.empty()) {
    NNPE_CHECK_HIP(ctx, dev_alloc(ctx->d_box, ctx->h_box.size()));
    NNPE_CHECK_HIP(ctx, hipMemcpy(ctx->d_box, ctx->h_box.data(), sizeof(BoxPar) * ctx->h_box.size(), hipMemcpyHostToDevice));
  }
  const int G = std::max(cell, 1);
  if (size_t(G) > s->cell_cap) {
    s->cell_cap = size_t(G) + G / 4;
    const size_t cc = s->cell_cap;
    NNPE_CHECK_HIP(ctx, dev_alloc(ctx->cell_start, cc + 1)); NNPE_CHECK_HIP(ctx, dev_alloc(ctx->cell_cnt, cc));
    NNPE_CHECK_HIP(ctx, dev_alloc(ctx->cell_fill, cc));
    NNPE_CHECK_HIP(ctx, dev_alloc(ctx->gcell_start, cc + 1)); NNPE_CHECK_HIP(ctx, dev_alloc(ctx->gcell_cnt, cc));
    NNPE_CHECK_HIP(ctx, dev_alloc(ctx->gcell_fill, cc));
    NNPE_CHECK_HIP(ctx, hipMemset(ctx->cell_start, 0, sizeof(int) * (cc + 1)));
    NNPE_CHECK_HIP(ctx, hipMemset(ctx->gcell_start, 0, sizeof(int) * (cc + 1)));
  }
  {
    std::vector<int> cb(G, 0);
    for (size_t b = 0; b < all.size(); ++b) {
      const BoxPar& g = ctx->h_box[b];
      std::fill(cb.begin() + g.cell_base, cb.begin() + g.cell_base + g.num_gx * g.num_gx * g.num_gx, int(b));
    }
    NNPE_CHECK_HIP(ctx, dev_alloc(ctx->cell_box, cb.size()));
    NNPE_CHECK_HIP(ctx, hipMemcpy(ctx->cell_box, cb.data(), sizeof(int) * cb.size(), hipMemcpyHostToDevice));
  }
  const int rh = home_cells_upload(ctx);
  if (rh) return rh;
  ctx->num_grid = cell;
  const long long gneed = (long long)ghost_est + 4096;
  if (gneed > ctx->ghost_cap) {
    ctx->ghost_cap = int(gneed + gneed / 4);
    NNPE_CHECK_HIP(ctx, dev_alloc(ctx->g_src, ctx->ghost_cap)); NNPE_CHECK_HIP(ctx, dev_alloc(ctx->g_code, ctx->ghost_cap));
    NNPE_CHECK_HIP(ctx, dev_alloc(ctx->g_cell, ctx->ghost_cap)); NNPE_CHECK_HIP(ctx, dev_alloc(ctx->gs_src, ctx->ghost_cap));
    NNPE_CHECK_HIP(ctx, dev_alloc(ctx->gs_code, ctx->ghost_cap));
  }
  auto capfor = [&](float rc) {
    double c2 = 4.0 / 3.0 * 3.141592653589793 * double(rc) * rc * rc * rho_max;
    int k2 = int(1.5 * c2) + 64;
    return (k2 + 63) / 64 * 64;
  };
  const int cl = capfor(s->pair_cutoff + s->skin), cs = capfor(s->nnp_cutoff + s->skin);
  if (cl > ctx->cap_long) {
    ctx->cap_long = cl;
    NNPE_CHECK_HIP(ctx, dev_alloc(ctx->nbr_long, size_t(ctx->cap_long) * ctx->stride));
  }
  if (cs > ctx->cap_short) {
    ctx->cap_short = cs;
    NNPE_CHECK_HIP(ctx, dev_alloc(ctx->nbr_short, size_t(ctx->ntype) * ctx->cap_short * ctx->stride));
    int rca = nnp_act_reserve(ctx);
    if (rca) return rca;
  }
  // molecules: atom id of each molecule's intra-0 atom
  {
    std::vector<int> mf(std::max(mol, 1), 0);
    for (SReq* r : all)
      for (int m2 = 0; m2 < r->nmol; ++m2) mf[r->mol_base + m2] = r->row_base + r->mol_first_local[m2];
    NNPE_CHECK_HIP(ctx, hipMemcpy(ctx->mol_first, mf.data(), sizeof(int) * mf.size(), hipMemcpyHostToDevice));
  }
  // element counts -> MLP tiles
  int rc = mlp_retile(ctx);
  if (rc) return rc;
  // bonded terms with arena atom ids (grouped by molecule: boxes in order, molecules ascending);
  // every kind is set, with no terms where the arena has none, so each per-molecule offset
  // table spans the arena's molecules
  for (int k = 0; k < 3; ++k) {
    const int w = 3 + k;
    std::vector<int> t;
    for (SReq* r : all) {
      const std::vector<int>& src = r->terms[k];
      for (size_t e = 0; e < src.size(); e += w) {
        t.push_back(src[e]);
        for (int a = 1; a < w; ++a) t.push_back(r->row_base + src[e + a]);
      }
    }
    rc = nnpe_bonded_set(ctx, k, s->ntyp[k], s->coef[k].data(), int(t.size() / w), t.empty() ? nullptr : t.data());
    if (rc) return rc;
  }
  ctx->have_list = false;
  s->relayouts++;
  return NNPE_OK;
}

extern "C" int nnpe_serve_advance(nnpe_serve* s, long long* done_ids, int cap, int* ndone) {
  if (!s || (cap > 0 && !done_ids)) return NNPE_ERR_ARG;
  nnpe_ctx* ctx = s->ctx;
  if (ndone) *ndone = 0;
  // leave: boxes that finished in the previous chunk were moved to `done` already
  std::vector<SReq*> keep(s->res);
  std::vector<SReq*> add;
  int atoms = 0;
  for (SReq* r : keep) atoms += r->natom;
  while (!s->queue.empty() && int(keep.size() + add.size()) < s->max_boxes && atoms + s->queue.front()->natom <= s->max_atoms) {
    add.push_back(s->queue.front());
    atoms += s->queue.front()->natom;
    s->queue.pop_front();
  }
  const bool changed = !add.empty() || ctx->nbox != int(keep.size());
  if (changed) {
    int rc = serve_relayout(s, keep, add);
    if (rc) return rc;
    s->res = keep;
    s->res.insert(s->res.end(), add.begin(), add.end());
    if (ctx->natom > 0) {
      // the chunk's first position update, except for the boxes that start now
      std::vector<unsigned char> skip(s->res.size(), 0);
      for (size_t b = 0; b < s->res.size(); ++b) skip[b] = s->res[b]->fresh ? 1 : 0;
      NNPE_CHECK_HIP(ctx, grow_alloc(s->d_skip, s->skip_cap, skip.size()));
      NNPE_CHECK_HIP(ctx, hipMemcpy(s->d_skip, skip.data(), skip.size(), hipMemcpyHostToDevice));
      rc = verlet_enqueue_x_skip(ctx, s->d_skip);
      if (rc) return rc;
      // the arena's step counter sits on a chunk boundary, so md_run rebuilds (eagerly, with
      // capacity growth) and runs the first step without a position update
      if (ctx->md_step % s->rebuild_every) ctx->md_step += s->rebuild_every - ctx->md_step % s->rebuild_every;
      if (ctx->md_step == 0) ctx->md_step = s->rebuild_every;
    }
  }
  if (ctx->natom == 0) return NNPE_OK;
  int nlog = 0;
  int rc = nnpe_md_run(ctx, &s->P, s->rebuild_every, nullptr, 0, &nlog);
  if (rc) return rc;
  s->chunks++;
  for (SReq* r : s->res) { r->fresh = false; r->done += s->rebuild_every; }
  // retire finished boxes: copy their rows out (rows of a box are contiguous)
  std::vector<SReq*> still;
  int nd = 0;
  for (SReq* r : s->res) {
    if (r->done < r->nsteps) { still.push_back(r); continue; }
    const int n = r->natom;
    std::vector<int4> hinfo(n);
    r->xo.assign(size_t(3) * n, 0.0);
    r->vo.assign(size_t(3) * n, 0.0);
    NNPE_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    NNPE_CHECK_HIP(ctx, hipMemcpy(hinfo.data(), ctx->info + r->row_base, sizeof(int4) * n, hipMemcpyDeviceToHost));
    if (ctx->fp64) {
      std::vector<double4> hx(n), hv(n);
      NNPE_CHECK_HIP(ctx, hipMemcpy(hx.data(), ctx->xd + r->row_base, sizeof(double4) * n, hipMemcpyDeviceToHost));
      NNPE_CHECK_HIP(ctx, hipMemcpy(hv.data(), ctx->vd + r->row_base, sizeof(double4) * n, hipMemcpyDeviceToHost));
      for (int k = 0; k < n; ++k) {
        const int a = hinfo[k].x - r->row_base;
        r->xo[3 * a] = hx[k].x; r->xo[3 * a + 1] = hx[k].y; r->xo[3 * a + 2] = hx[k].z;
        r->vo[3 * a] = hv[k].x; r->vo[3 * a + 1] = hv[k].y; r->vo[3 * a + 2] = hv[k].z;
      }
    } else {
      std::vector<float4> hx(n), hv(n);
      NNPE_CHECK_HIP(ctx, hipMemcpy(hx.data(), ctx->xq + r->row_base, sizeof(float4) * n, hipMemcpyDeviceToHost));
      NNPE_CHECK_HIP(ctx, hipMemcpy(hv.data(), ctx->v + r->row_base, sizeof(float4) * n, hipMemcpyDeviceToHost));
      for (int k = 0; k < n; ++k) {
        const int a = hinfo[k].x - r->row_base;
        r->xo[3 * a] = hx[k].x; r->xo[3 * a + 1] = hx[k].y; r->xo[3 * a + 2] = hx[k].z;
        r->vo[3 * a] = hv[k].x; r->vo[3 * a + 1] = hv[k].y; r->vo[3 * a + 2] = hv[k].z;
      }
    }
    s->done[r->id] = r;
    if (nd < cap) done_ids[nd] = r->id;
    ++nd;
  }
  s->res.swap(still);
  if (ndone) *ndone = nd;
  return NNPE_OK;
}

extern "C" int nnpe_serve_result(nnpe_serve* s, long long id, double* x, double* v) {
  if (!s) return NNPE_ERR_ARG;
  auto it = s->done.find(id);
  if (it == s->done.end()) return serve_fail(s, NNPE_ERR_STATE, "no finished request with this id");
  SReq* r = it->second;
  if (x) std::copy(r->xo.begin(), r->xo.end(), x);
  if (v) std::copy(r->vo.begin(), r->vo.end(), v);
  delete r;
  s->done.erase(it);
  return NNPE_OK;
}

extern "C" int nnpe_serve_stats(nnpe_serve* s, long long* out) {
  if (!s || !out) return NNPE_ERR_ARG;
  long long atoms = 0;
  for (SReq* r : s->res) atoms += r->natom;
  out[0] = (long long)s->res.size();
  out[1] = atoms;
  out[2] = (long long)s->queue.size();
  out[3] = (long long)s->done.size();
  out[4] = s->chunks;
  out[5] = s->relayouts;
  out[6] = (long long)s->arena_bytes;
  out[7] = s->max_atoms;
  return NNPE_OK;
}

extern "C" const char* nnpe_serve_last_error(nnpe_serve* s) { return s && s->ctx ? s->ctx->err.c_str() : ""; }
