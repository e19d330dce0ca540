// oracle.cpp -- CPU restatement of NNPEngine's per-step MD hot path.
//
// TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg load this library, and only as the test reference (never as the thing measured or shipped).
//
// Every algorithm is restated from the reference sources (paths relative to the reference repo),
// templated on the arithmetic type:
//   T = float  : the reference's own arithmetic (numtype=float, src/environ.h:4-12), used for
//                bit-exact neighbor lists and g(r) counts and as the FP32 force reference;
//   T = double : the FP64 oracle every GPU result states its error against.
// The reference runs 27 OpenMP threads; this restatement is its NUM_THREAD = 1 case (every real
// atom is owned by thread 0), whose half list + ghost entries give the same set of interacting
// (atom, atom, image) triples as the 27-thread run.
//
// Pinning (what checks this file): tests/test_oracle_*.py compare it against the reference built
// from source (oracle/_ref, fixtures under tests/golden/), against brute-force O(N^2) minimum-image
// searches, closed forms (analytic LJ + shifted-force Coulomb, harmonic bond/angle values,
// analytic NNPOTENTIAL descriptors), finite differences of energies, and FP64 autograd (torch CPU)
// for the MLP.
//
// Compile: g++ -O2 -std=c++17 -shared -fPIC -ffp-contract=off (see oracle/Makefile).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <type_traits>
#include <vector>

namespace {

template <class T> inline T sqr(T v) { return v * v; }
template <class T> inline T cub(T v) { return v * v * v; }

const double kElectricK = 332.06371;           // src/environ.h:77
const double kAccelF = 4.184e-4;               // src/environ.h:78
const double kEkinF = 1195.0286802753603;      // src/environ.h:79
const double kSpecialLj[4] = {1, 0, 0, 0.5};   // src/molecule.h:11
const double kSpecialCoul[4] = {1, 0, 0, 5 / 6.};

struct SysIn {  // flat system description, all per-atom arrays in atom-id (input) order
  int natom, ntype, nmol, nmoltype;
  float bl;
  const float *x, *v, *q, *mass;
  const int *typ, *mol_id, *intra_id, *mol_type;
  const int *mt_natom, *mt_special_off, *mt_special;
};

template <class T>
struct Eng {
  // ---- ENVIRON (src/environ.h:20-123) -------------------------------------------------------
  int natom = 0, ntype = 0, npair = 0, nmol = 0, nghost = 0, ntot = 0;
  T bl, r_bl, h_bl, m_bl;
  std::vector<T> x;           // [ntot*3]  real atoms (sorted by cell) then ghosts
  std::vector<T> v, f, f_prev;  // [natom*3]
  std::vector<T> q, r_m;      // q[ntot], r_m[natom]
  std::vector<int> typ, atom_id, mol_id, intra_id;  // [ntot]
  std::vector<int> x_grd, x_grd_ghost;
  std::vector<int> i_ghost;   // ghost -> source real row
  std::vector<int> g_code;    // ghost image code (ix+1)*9+(iy+1)*3+(iz+1)
  std::vector<int> xmol_row, fmol_row;  // by atom id: row of x used for x_mol, row of f
  std::vector<int> mol_type, mol_repre;  // per molecule
  std::vector<int> mt_natom, mt_special_off, mt_special;
  std::vector<int> pairtype;  // [ntype*ntype]
  std::vector<int> typecount;
  // grid (ENVIRON::init_grid, src/environ.cpp:143-155)
  int num_gx = 0, num_gxsq = 0, num_grid = 0;
  T g_dx, g_rdx, g_offset, ghost_lim, m_ghost_lim;
  std::vector<int> grid_start, grid_end, grid_start_ghost, grid_end_ghost;
  std::vector<std::vector<int>> bin_list_pair, bin_list_gr;
  // NEIGHLIST (src/neighlist.h:6-47), NUM_THREAD = 1
  T cutsq_list = 0, dx2 = 0;
  std::vector<int> nl_start, nl_num, nl_j, nl_sp;
  std::vector<T> x_prev;
  int nbuild = 0;
  bool list_full = false;

  static constexpr int mask_bin = 0xFFFFFF, mask_no_check = 1 << 24;

  void set_bl(T b) { bl = b; r_bl = T(1) / bl; h_bl = bl * T(.5); m_bl = -bl; }

  // ENVIRON::initialize (src/environ.cpp:22-141), from arrays instead of the text file.
  void init(const SysIn& s) {
    natom = s.natom; ntype = s.ntype; nmol = s.nmol;
    set_bl(T(s.bl));
    x.assign(size_t(natom) * 3, 0); v.assign(size_t(natom) * 3, 0);
    f.assign(size_t(natom) * 3, 0); f_prev.assign(size_t(natom) * 3, 0);
    q.resize(natom); r_m.resize(natom); typ.resize(natom); atom_id.resize(natom);
    mol_id.resize(natom); intra_id.resize(natom);
    for (int i = 0; i < natom; ++i) {
      for (int d = 0; d < 3; ++d) {
        x[3 * i + d] = T(s.x[3 * i + d]);
        v[3 * i + d] = s.v ? T(s.v[3 * i + d]) : T(0);
      }
      q[i] = T(s.q[i]);
      r_m[i] = T(1) / T(s.mass[i]);
      typ[i] = s.typ[i]; atom_id[i] = i; mol_id[i] = s.mol_id[i]; intra_id[i] = s.intra_id[i];
    }
    mol_type.assign(s.mol_type, s.mol_type + nmol);
    mol_repre.assign(nmol, -1);
    for (int i = 0; i < natom; ++i) if (intra_id[i] == 0) mol_repre[mol_id[i]] = i;
    mt_natom.assign(s.mt_natom, s.mt_natom + s.nmoltype);
    mt_special_off.assign(s.mt_special_off, s.mt_special_off + s.nmoltype);
    int tot = 0;
    for (int k = 0; k < s.nmoltype; ++k) tot = std::max(tot, s.mt_special_off[k] + s.mt_natom[k] * s.mt_natom[k]);
    mt_special.assign(s.mt_special, s.mt_special + tot);
    pairtype.assign(size_t(ntype) * ntype, 0);
    npair = 0;
    for (int i = 0; i < ntype; ++i) for (int j = i; j < ntype; ++j) pairtype[i * ntype + j] = pairtype[j * ntype + i] = npair++;
    typecount.assign(ntype, 0);
    for (int i = 0; i < natom; ++i) ++typecount[typ[i]];
    xmol_row.resize(natom); fmol_row.resize(natom);
    for (int i = 0; i < natom; ++i) xmol_row[i] = fmol_row[i] = i;
    ntot = natom; nghost = 0;
  }

  int special_of(int mi_type, int ia, int ib) const {
    return mt_special[mt_special_off[mi_type] + ia * mt_natom[mi_type] + ib];
  }

  // ENVIRON::init_grid + LOCAL::set_ghost_lim (src/environ.cpp:143-155, src/environ.h:142)
  void init_grid(T grid_length, T ghost_cutoff) {
    g_rdx = T(1) / (g_dx = grid_length);
    g_offset = ghost_cutoff + h_bl;
    num_gx = int(std::ceil((g_offset * T(2)) * g_rdx));
    num_gxsq = num_gx * num_gx;
    num_grid = num_gx * num_gxsq;
    grid_start.assign(num_grid, 0); grid_end.assign(num_grid, 0);
    grid_start_ghost.assign(num_grid, 0); grid_end_ghost.assign(num_grid, 0);
    ghost_lim = h_bl - ghost_cutoff; m_ghost_lim = -ghost_lim;
  }

  // NEIGHLIST::build_bin_list (src/neighlist.cpp:35-168), NEGH_USE_INNERCUT on.
  std::vector<std::vector<int>> build_bin_list(T cutoff) {
    std::vector<std::vector<int>> bl_out(num_grid);
    const int max_d = int(std::ceil(sqr(cutoff * g_rdx)));
    for (int _i = 0; _i < num_gx; ++_i) for (int _j = 0; _j < num_gx; ++_j) for (int _k = 0; _k < num_gx; ++_k) {
      std::vector<int>& il = bl_out[(_i * num_gx + _j) * num_gx + _k];
      for (int _l = 0; _l < num_gx; ++_l) {
        int mx, mxx, tpp;
        if (_l == _i) { mx = 0; mxx = 1; }
        else { tpp = std::abs(_l - _i); if ((mx = sqr(tpp - 1)) >= max_d) continue; mxx = sqr(tpp + 1); }
        for (int _m = 0; _m < num_gx; ++_m) {
          int my, mxy;
          if (_m == _j) { my = mx; mxy = mxx + 1; }
          else { tpp = std::abs(_m - _j); if ((my = mx + sqr(tpp - 1)) >= max_d) continue; mxy = mxx + sqr(tpp + 1); }
          for (int _n = 0; _n < num_gx; ++_n) {
            int mz, mxz;
            if (_n == _k) { mz = my; mxz = mxy + 1; }
            else { tpp = std::abs(_n - _k); if ((mz = my + sqr(tpp - 1)) >= max_d) continue; mxz = mxy + sqr(tpp + 1); }
            int index = (_l * num_gx + _m) * num_gx + _n;
            if (mxz < max_d) index |= mask_no_check;
            il.push_back(index);
          }
        }
      }
    }
    return bl_out;
  }

  // LOCAL::refold_and_tag (src/environ.cpp:234-262), cell part only (one owner thread).
  void refold_and_tag() {
    x_grd.assign(natom, 0);
    for (int i = 0; i < natom; ++i) {
      T* xl = &x[3 * i];
      int c[3];
      for (int d = 0; d < 3; ++d) xl[d] -= T(std::nearbyint(xl[d] * r_bl)) * bl;
      for (int d = 0; d < 3; ++d) c[d] = int((g_offset + xl[d]) * g_rdx);
      x_grd[i] = c[0] * num_gxsq + c[1] * num_gx + c[2];
    }
  }

  // ENVIRON::sort (src/environ.cpp:303-362): stable counting sort of real atoms by cell.
  void sort() {
    std::vector<int> cnt(num_grid, 0), order(natom);
    for (int i = 0; i < natom; ++i) ++cnt[x_grd[i]];
    int acc = 0;
    for (int c = 0; c < num_grid; ++c) { grid_start[c] = acc; acc += cnt[c]; grid_end[c] = acc; }
    std::vector<int> fill(grid_start);
    for (int i = 0; i < natom; ++i) order[fill[x_grd[i]]++] = i;
    auto perm3 = [&](std::vector<T>& a) { std::vector<T> b(size_t(natom) * 3); for (int k = 0; k < natom; ++k) for (int d = 0; d < 3; ++d) b[3 * k + d] = a[3 * order[k] + d]; std::copy(b.begin(), b.end(), a.begin()); };
    auto perm1 = [&](auto& a) { auto b = a; for (int k = 0; k < natom; ++k) b[k] = a[order[k]]; std::copy(b.begin(), b.begin() + natom, a.begin()); };
    x.resize(size_t(natom) * 3);
    perm3(x); perm3(v); perm3(f_prev);
    q.resize(natom); typ.resize(natom); atom_id.resize(natom); mol_id.resize(natom); intra_id.resize(natom);
    perm1(q); perm1(r_m); perm1(typ); perm1(atom_id); perm1(mol_id); perm1(intra_id); perm1(x_grd);
    for (int i = 0; i < natom; ++i) { fmol_row[atom_id[i]] = i; xmol_row[atom_id[i]] = i; }
    // mol_repre holds atom ids; nothing to permute.
  }

  // LOCAL::generate_ghost + ENVIRON::tidyup_ghost + update_ghost_pos<true> + ENVIRON::sort_ghost
  // + LOCAL::update_ghost_and_info (src/environ.cpp:157-176, :364-522).
  void make_ghosts() {
    std::vector<int> src, code;
    for (int i = 0; i < natom; ++i) {
      const T* xl = &x[3 * i];
      int xs = xl[0] > ghost_lim ? -1 : 0, xe = xl[0] < m_ghost_lim ? 1 : 0;
      int ys = xl[1] > ghost_lim ? -1 : 0, ye = xl[1] < m_ghost_lim ? 1 : 0;
      int zs = xl[2] > ghost_lim ? -1 : 0, ze = xl[2] < m_ghost_lim ? 1 : 0;
      if (xs != xe || ys != ye || zs != ze)
        for (int a = xs; a <= xe; ++a) for (int b = ys; b <= ye; ++b) for (int c = zs; c <= ze; ++c) {
          if (a == 0 && b == 0 && c == 0) continue;
          src.push_back(i); code.push_back((a + 1) * 9 + (b + 1) * 3 + (c + 1));
        }
    }
    nghost = int(src.size());
    ntot = natom + nghost;
    std::vector<T> gx(size_t(nghost) * 3);
    std::vector<int> gcell(nghost);
    for (int g = 0; g < nghost; ++g) {
      int c[3];
      for (int d = 0; d < 3; ++d) {
        gx[3 * g + d] = x[3 * src[g] + d] + image_shift(code[g], d);
        c[d] = int((g_offset + gx[3 * g + d]) * g_rdx);
      }
      gcell[g] = c[0] * num_gxsq + c[1] * num_gx + c[2];
    }
    // stable counting sort by cell
    std::vector<int> cnt(num_grid, 0), order(nghost);
    for (int g = 0; g < nghost; ++g) ++cnt[gcell[g]];
    int acc = natom;
    for (int c = 0; c < num_grid; ++c) { grid_start_ghost[c] = acc; acc += cnt[c]; grid_end_ghost[c] = acc; }
    std::vector<int> fill(num_grid);
    for (int c = 0; c < num_grid; ++c) fill[c] = grid_start_ghost[c] - natom;
    for (int g = 0; g < nghost; ++g) order[fill[gcell[g]]++] = g;
    i_ghost.resize(nghost); g_code.resize(nghost); x_grd_ghost.resize(nghost);
    x.resize(size_t(ntot) * 3); q.resize(ntot); typ.resize(ntot); atom_id.resize(ntot);
    mol_id.resize(ntot); intra_id.resize(ntot);
    for (int k = 0; k < nghost; ++k) {
      int g = order[k], row = natom + k, s = src[g];
      i_ghost[k] = s; g_code[k] = code[g]; x_grd_ghost[k] = gcell[g];
      for (int d = 0; d < 3; ++d) x[3 * row + d] = gx[3 * g + d];
      q[row] = q[s]; typ[row] = typ[s]; atom_id[row] = atom_id[s]; mol_id[row] = mol_id[s]; intra_id[row] = intra_id[s];
      if (intra_id[s]) {  // unwrapped molecule image (src/environ.cpp:461-469)
        const T* xr = &x[3 * xmol_row[mol_repre[mol_id[s]]]];
        const T* xg = &x[3 * row];
        if (std::fabs(xg[0] - xr[0]) < h_bl && std::fabs(xg[1] - xr[1]) < h_bl && std::fabs(xg[2] - xr[2]) < h_bl)
          xmol_row[atom_id[s]] = row;
      }
    }
  }

  T image_shift(int code, int d) const {
    int c = d == 0 ? code / 9 : (d == 1 ? (code / 3) % 3 : code % 3);
    return c == 1 ? T(0) : (c > 1 ? bl : m_bl);
  }

  // LOCAL::update_ghost_pos<false> (src/environ.cpp:488-522)
  void update_ghost_pos() {
    for (int k = 0; k < nghost; ++k)
      for (int d = 0; d < 3; ++d) x[3 * (natom + k) + d] = x[3 * i_ghost[k] + d] + image_shift(g_code[k], d);
  }

  // NEIGHLIST::build<full> (src/neighlist.cpp:190-331) for the single owner thread.
  void build_list(bool full) {
    list_full = full;
    ++nbuild;
    nl_start.assign(natom + 1, 0); nl_num.assign(natom, 0); nl_j.clear(); nl_sp.clear();
    x_prev.assign(x.begin(), x.begin() + size_t(natom) * 3);
    for (int i = 0; i < natom; ++i) {
      nl_start[i] = int(nl_j.size());
      const T xi = x[3 * i], yi = x[3 * i + 1], zi = x[3 * i + 2];
      const int mi = mol_id[i], mt = mol_type[mi], ia = intra_id[i];
      auto push = [&](int j) {
        nl_sp.push_back(mi != mol_id[j] ? 0 : special_of(mt, ia, intra_id[j]));
        nl_j.push_back(j);
      };
      const std::vector<int>& ib = bin_list_pair[x_grd[i]];
      for (int cb : ib) {  // real atoms
        bool nocheck = cb & mask_no_check; int c = cb & mask_bin;
        for (int j = grid_start[c]; j < grid_end[c]; ++j) {
          if (full) { if (j == i) continue; } else { if (j <= i) continue; }
          if (nocheck) push(j);
          else { const T* yy = &x[3 * j]; if (sqr(xi - yy[0]) + sqr(yi - yy[1]) + sqr(zi - yy[2]) < cutsq_list) push(j); }
        }
      }
      for (int cb : ib) {  // ghosts
        bool nocheck = cb & mask_no_check; int c = cb & mask_bin;
        for (int j = grid_start_ghost[c]; j < grid_end_ghost[c]; ++j) {
          if (nocheck) push(j);
          else { const T* yy = &x[3 * j]; if (sqr(xi - yy[0]) + sqr(yi - yy[1]) + sqr(zi - yy[2]) < cutsq_list) push(j); }
        }
      }
      nl_num[i] = int(nl_j.size()) - nl_start[i];
    }
    nl_start[natom] = int(nl_j.size());
  }

  // NEIGHLIST::isvalid (src/neighlist.cpp:178-187)
  bool isvalid() const {
    if (x_prev.size() < size_t(natom) * 3) return false;  // no build yet (src/neighlist.cpp:29)
    for (int i = 0; i < natom; ++i)
      if (sqr(x[3 * i] - x_prev[3 * i]) + sqr(x[3 * i + 1] - x_prev[3 * i + 1]) + sqr(x[3 * i + 2] - x_prev[3 * i + 2]) > dx2) return false;
    return true;
  }

  // full rebuild sequence of the MD loop (src/run.cpp:249-276)
  void rebuild(bool full) {
    refold_and_tag();
    sort();
    make_ghosts();
    build_list(full);
  }

  // ---- LJTableCoulDSF (src/pair.cpp:171-419) ------------------------------------------------
  T pair_cut = 0, pair_cutsq = 0, pair_dr = 0, pair_rdr = 0, vc = 0, fc = 0;
  int pair_nbin = 0;
  std::vector<T> tbl_f, tbl_u;  // [npair][nbin+1]
  void pair_set(T cutoff, int nbin, const float* force, const float* pot) {
    pair_cutsq = sqr(pair_cut = cutoff);
    pair_nbin = nbin;
    pair_dr = cutoff / T(nbin); pair_rdr = T(nbin) / cutoff;
    vc = T(-1) / cutoff; fc = vc / cutoff;
    tbl_f.assign(force, force + size_t(npair) * (nbin + 1));
    tbl_u.assign(pot, pot + size_t(npair) * (nbin + 1));
  }

  // compute<ev,false>; energy/virial accumulated in double.
  void pair_compute(bool ev, double* erg, double* vir) {
    const T K = T(kElectricK);
    for (int i = 0; i < natom; ++i) {
      const T xi = x[3 * i], yi = x[3 * i + 1], zi = x[3 * i + 2], qi = q[i];
      T fx = 0, fy = 0, fz = 0;
      for (int e = nl_start[i]; e < nl_start[i] + nl_num[i]; ++e) {
        int jj = nl_j[e], sp = nl_sp[e];
        const T* xj = &x[3 * jj];
        T dx = xi - xj[0], dy = yi - xj[1], dz = zi - xj[2];
        T rsq = sqr(dx) + sqr(dy) + sqr(dz);
        if (!(rsq < pair_cutsq)) continue;
        T r = std::sqrt(rsq);
        int ijp = pairtype[typ[i] * ntype + typ[jj]];
        T ipw1 = r * pair_rdr - T(0.5F);
        int tbr1 = int(ipw1), tbr2 = tbr1 + 1;
        T ipw2 = ipw1 - T(tbr1); ipw1 = T(1) - ipw2;
        const T* ijf = &tbl_f[size_t(ijp) * (pair_nbin + 1)];
        T slj = T(kSpecialLj[sp]), scl = T(kSpecialCoul[sp]);
        T fpair = (ijf[tbr1] * ipw1 + ijf[tbr2] * ipw2) * slj;
        fpair += qi * q[jj] * K * (T(1) / (r * rsq) + fc / r) * scl;
        fx += (dx *= fpair); fy += (dy *= fpair); fz += (dz *= fpair);
        bool own = jj < natom;
        if (own) { T* fj = &f[3 * jj]; fj[0] -= dx; fj[1] -= dy; fj[2] -= dz; }
        if (ev) {
          const T* ijpo = &tbl_u[size_t(ijp) * (pair_nbin + 1)];
          double u = double((ijpo[tbr1] * ipw1 + ijpo[tbr2] * ipw2) * slj) +
                     double(qi * q[jj] * K * (T(1) / r + vc + fc * (pair_cut - r)) * scl);
          double w = double(fpair * rsq);
          if (own || list_full) { *erg += list_full ? 0.5 * u : u; *vir += list_full ? 0.5 * w : w; }
          else { *erg += 0.5 * u; *vir += 0.5 * w; }
        }
      }
      f[3 * i] += fx; f[3 * i + 1] += fy; f[3 * i + 2] += fz;
    }
  }

  // ---- BOND / ANGLE / DIHEDRAL (src/bond.cpp:87-116, src/angle.cpp:87-136,
  //      src/dihedral.cpp:90-147); x_mol / f_mol indirection by atom id -----------------------
  const T* xm(int aid) const { return &x[3 * xmol_row[aid]]; }
  T* fm(int aid) { return &f[3 * fmol_row[aid]]; }

  void bond_compute(int nb, const int* bonds, const float* coef, bool ev, double* erg, double* vir) {
    for (int b = 0; b < nb; ++b) {
      const int* p = bonds + 3 * b;
      T icf[4]; for (int k = 0; k < 4; ++k) icf[k] = T(coef[4 * p[0] + k]);
      const T *xi = xm(p[1]), *xj = xm(p[2]);
      T *fi = fm(p[1]), *fj = fm(p[2]);
      T dx = xi[0] - xj[0], dy = xi[1] - xj[1], dz = xi[2] - xj[2];
      T rr = std::sqrt(sqr(dx) + sqr(dy) + sqr(dz));
      T dr = rr - icf[0], bdr = icf[2] - dr;
      T ff = dr * (icf[3] + icf[1] * bdr * (bdr - dr)) / rr;
      dx *= ff; dy *= ff; dz *= ff;
      fi[0] -= dx; fi[1] -= dy; fi[2] -= dz;
      fj[0] += dx; fj[1] += dy; fj[2] += dz;
      if (ev) { *erg += double(T(.5F) * sqr(dr) * (icf[1] * sqr(bdr) + icf[3])); *vir -= double(ff * rr * rr); }
    }
  }

  void angle_compute(int na, const int* angles, const float* coef, bool ev, double* erg) {
    for (int a = 0; a < na; ++a) {
      const int* p = angles + 4 * a;
      T icf[4]; for (int k = 0; k < 4; ++k) icf[k] = T(coef[4 * p[0] + k]);
      const T *xi = xm(p[1]), *xj = xm(p[2]), *xk = xm(p[3]);
      T *fi = fm(p[1]), *fj = fm(p[2]), *fk = fm(p[3]);
      T dx1 = xi[0] - xj[0], dy1 = xi[1] - xj[1], dz1 = xi[2] - xj[2];
      T dx2 = xk[0] - xj[0], dy2 = xk[1] - xj[1], dz2 = xk[2] - xj[2];
      T r1 = T(1) / std::sqrt(sqr(dx1) + sqr(dy1) + sqr(dz1));
      T r2 = T(1) / std::sqrt(sqr(dx2) + sqr(dy2) + sqr(dz2));
      dx1 *= r1; dy1 *= r1; dz1 *= r1; dx2 *= r2; dy2 *= r2; dz2 *= r2;
      T cq = dx1 * dx2 + dy1 * dy2 + dz1 * dz2;
      T dcq = cq - icf[0], b_dcq = icf[2] - dcq;
      T fe = dcq * (icf[3] + icf[1] * b_dcq * (b_dcq - dcq));
      T ff = fe * r1;
      T fx = ff * (cq * dx1 - dx2), fy = ff * (cq * dy1 - dy2), fz = ff * (cq * dz1 - dz2);
      fi[0] += fx; fi[1] += fy; fi[2] += fz; fj[0] -= fx; fj[1] -= fy; fj[2] -= fz;
      ff = fe * r2;
      fx = ff * (dx1 - cq * dx2); fy = ff * (dy1 - cq * dy2); fz = ff * (dz1 - cq * dz2);
      fk[0] -= fx; fk[1] -= fy; fk[2] -= fz; fj[0] += fx; fj[1] += fy; fj[2] += fz;
      if (ev) *erg += double(T(.5F) * sqr(dcq) * (icf[1] * sqr(b_dcq) + icf[3]));
    }
  }

  // poly = false: the reference's stored cubic (src/dihedral.cpp:27-35);
  // poly = true: coef [ntyp][6] = C0..C5 of U = sum_n C_n cos^n(phi) (Ryckaert-Bellemans, no
  // counterpart in the reference), evaluated on the same geometry with dU/dcos = sum n C_n cos^(n-1)
  void dihedral_compute(int nd, const int* dih, const float* coef, bool ev, double* erg, bool poly = false) {
    for (int a = 0; a < nd; ++a) {
      const int* p = dih + 5 * a;
      T icf[4] = {0, 0, 0, 0}, C[6] = {0, 0, 0, 0, 0, 0};
      if (poly) for (int k = 0; k < 6; ++k) C[k] = T(coef[6 * p[0] + k]);
      else for (int k = 0; k < 4; ++k) icf[k] = T(coef[4 * p[0] + k]);
      const T *xi = xm(p[1]), *xj = xm(p[2]), *xk = xm(p[3]), *xl = xm(p[4]);
      T *fi = fm(p[1]), *fj = fm(p[2]), *fk = fm(p[3]), *fl = fm(p[4]);
      T dx1 = xi[0] - xj[0], dy1 = xi[1] - xj[1], dz1 = xi[2] - xj[2];
      T dx2 = xk[0] - xj[0], dy2 = xk[1] - xj[1], dz2 = xk[2] - xj[2];
      T dx3 = xl[0] - xk[0], dy3 = xl[1] - xk[1], dz3 = xl[2] - xk[2];
      T r_rsq2 = T(1) / (sqr(dx2) + sqr(dy2) + sqr(dz2));
      T dot12 = (dx1 * dx2 + dy1 * dy2 + dz1 * dz2) * r_rsq2;
      T dot32 = (dx3 * dx2 + dy3 * dy2 + dz3 * dz2) * r_rsq2;
      dx1 -= dx2 * dot12; dy1 -= dy2 * dot12; dz1 -= dz2 * dot12;
      dx3 -= dx2 * dot32; dy3 -= dy2 * dot32; dz3 -= dz2 * dot32;
      T r1 = T(1) / std::sqrt(sqr(dx1) + sqr(dy1) + sqr(dz1));
      T r3 = T(1) / std::sqrt(sqr(dx3) + sqr(dy3) + sqr(dz3));
      dx1 *= r1; dy1 *= r1; dz1 *= r1; dx3 *= r3; dy3 *= r3; dz3 *= r3;
      T cq = dx1 * dx3 + dy1 * dy3 + dz1 * dz3;
      T ff = poly ? (((T(5) * C[5] * cq + T(4) * C[4]) * cq + T(3) * C[3]) * cq + T(2) * C[2]) * cq + C[1]
                  : (icf[3] * cq + icf[2]) * cq + icf[1];
      r1 *= ff; r3 *= ff;
      T f1x = r1 * (cq * dx1 - dx3), f1y = r1 * (cq * dy1 - dy3), f1z = r1 * (cq * dz1 - dz3);
      T f3x = r3 * (cq * dx3 - dx1), f3y = r3 * (cq * dy3 - dy1), f3z = r3 * (cq * dz3 - dz1);
      T f2x = f1x * dot12 + f3x * dot32, f2y = f1y * dot12 + f3y * dot32, f2z = f1z * dot12 + f3z * dot32;
      fi[0] += f1x; fi[1] += f1y; fi[2] += f1z;
      fl[0] += f3x; fl[1] += f3y; fl[2] += f3z;
      fj[0] += f2x - f1x; fj[1] += f2y - f1y; fj[2] += f2z - f1z;
      fk[0] -= f2x + f3x; fk[1] -= f2y + f3y; fk[2] -= f2z + f3z;
      if (ev && poly) *erg += double(((((C[5] * cq + C[4]) * cq + C[3]) * cq + C[2]) * cq + C[1]) * cq + C[0]);
      else if (ev) *erg += double(((icf[3] / T(3) * cq + icf[2] / T(2)) * cq + icf[1]) * cq + icf[0]);
    }
  }

  // ---- VERLET (src/verlet.cpp:29-97) --------------------------------------------------------
  T dt = 1, hdt = T(.5);
  void set_dt(T d) { dt = d; hdt = dt * T(.5F); }
  void update_x() {
    for (int i = 0; i < natom; ++i)
      for (int d = 0; d < 3; ++d) x[3 * i + d] += (v[3 * i + d] + f[3 * i + d] * hdt) * dt;
    std::copy(f.begin(), f.begin() + size_t(natom) * 3, f_prev.begin());
  }
  void update_v() {
    for (int i = 0; i < natom; ++i) {
      T rm = r_m[i] * T(kAccelF);
      for (int d = 0; d < 3; ++d) v[3 * i + d] += (f_prev[3 * i + d] + (f[3 * i + d] *= rm)) * hdt;
    }
  }
  double compute_ke() const {
    double ke = 0;
    for (int i = 0; i < natom; ++i) ke += double((sqr(v[3 * i]) + sqr(v[3 * i + 1]) + sqr(v[3 * i + 2])) / r_m[i]);
    return ke * kEkinF;
  }
  void clear_f() { std::fill(f.begin(), f.end(), T(0)); }

  // ---- GR (src/grsq.cpp:28-219) -------------------------------------------------------------
  int gr_nbin = 0, gr_nbin_inner = 0;
  T gr_dr, gr_rdr, gr_rmaxsq, gr_rmaxsq_inner;
  std::vector<int64_t> gr_count;  // [npair][gr_nbin]
  std::vector<double> gr_h;
  void gr_init(int nbi, T r_max, T r_max_inner) {
    if (r_max_inner <= 0) {
      r_max = r_max <= 0 ? h_bl : r_max;
      gr_nbin_inner = gr_nbin = nbi;
      gr_rmaxsq_inner = gr_rmaxsq = sqr(r_max);
      gr_dr = r_max / T(nbi); gr_rdr = T(nbi) / r_max;
    } else {
      gr_nbin_inner = nbi;
      gr_dr = r_max_inner / T(nbi); gr_rdr = T(nbi) / r_max_inner;
      gr_rmaxsq_inner = sqr(r_max_inner);
      r_max = r_max <= 0 ? h_bl : r_max;
      gr_nbin = int(r_max * gr_rdr);
      gr_rmaxsq = sqr(T(gr_nbin) * gr_dr);
    }
    gr_count.assign(size_t(npair) * gr_nbin, 0);
  }
  // tally<true>(list): counts are 2 per unordered pair (src/grsq.cpp:94-149). Bins >= nbin_inner
  // keep their previous values, as the reference's per-thread histograms do.
  void gr_tally_inner() {
    for (int p = 0; p < npair; ++p) for (int b = 0; b < gr_nbin_inner; ++b) gr_count[size_t(p) * gr_nbin + b] = 0;
    for (int i = 0; i < natom; ++i) {
      const T xi = x[3 * i], yi = x[3 * i + 1], zi = x[3 * i + 2];
      for (int e = nl_start[i]; e < nl_start[i] + nl_num[i]; ++e) {
        int jj = nl_j[e]; const T* xj = &x[3 * jj];
        T rsq = sqr(xi - xj[0]) + sqr(yi - xj[1]) + sqr(zi - xj[2]);
        int bin;
        if (rsq < gr_rmaxsq_inner && (bin = int(std::sqrt(rsq) * gr_rdr)) < gr_nbin_inner)
          gr_count[size_t(pairtype[typ[i] * ntype + typ[jj]]) * gr_nbin + bin] += (list_full ? 1 : (jj < natom ? 2 : 1));
      }
    }
  }
  // tally(bin_list, local): full stencil walk to r_max (src/grsq.cpp:151-205)
  void gr_tally_full() {
    std::fill(gr_count.begin(), gr_count.end(), 0);
    for (int i = 0; i < natom; ++i) {
      const T xi = x[3 * i], yi = x[3 * i + 1], zi = x[3 * i + 2];
      const int* ip = &pairtype[typ[i] * ntype];
      const std::vector<int>& ib = bin_list_gr[x_grd[i]];
      for (int cb : ib) {
        int c = cb & mask_bin;
        for (int j = std::max(i + 1, grid_start[c]); j < grid_end[c]; ++j) {
          const T* yy = &x[3 * j]; T rsq = sqr(xi - yy[0]) + sqr(yi - yy[1]) + sqr(zi - yy[2]); int bin;
          if (rsq < gr_rmaxsq && (bin = int(std::sqrt(rsq) * gr_rdr)) < gr_nbin) gr_count[size_t(ip[typ[j]]) * gr_nbin + bin] += 2;
        }
      }
      for (int cb : ib) {
        int c = cb & mask_bin;
        for (int j = grid_start_ghost[c]; j < grid_end_ghost[c]; ++j) {
          const T* yy = &x[3 * j]; T rsq = sqr(xi - yy[0]) + sqr(yi - yy[1]) + sqr(zi - yy[2]); int bin;
          if (rsq < gr_rmaxsq && (bin = int(std::sqrt(rsq) * gr_rdr)) < gr_nbin) gr_count[size_t(ip[typ[j]]) * gr_nbin + bin] += 1;
        }
      }
    }
  }
  // reduce_local + ctor normalisation: h = count * norm - 1 (src/grsq.cpp:56-70, :207-219)
  void gr_reduce() {
    gr_h.assign(size_t(npair) * gr_nbin, 0);
    const double pi4 = 4 * 3.141592653589793;
    double rc_dr3 = cub(double(bl) * double(gr_rdr)) / pi4;
    int ip = 0;
    for (int i = 0; i < ntype; ++i) for (int j = i; j < ntype; ++j, ++ip) {
      double normg = rc_dr3 / (double(typecount[i]) * (i == j ? double(typecount[j] - 1) : 2.0 * typecount[j]));
      for (int k = 0; k < gr_nbin; ++k) gr_h[size_t(ip) * gr_nbin + k] = double(gr_count[size_t(ip) * gr_nbin + k]) * (normg / (double(k) * (k + 1) + 1.0 / 3)) - 1;
    }
  }
  // SQ::compute with S(q) accumulated in double (src/grsq.cpp:258-322); q_i = q_min + (i+1/2) dq
  std::vector<double> sq;
  void sq_compute(int nq, double q_min, double q_max) {
    sq.assign(size_t(npair) * nq, 0);
    const double rho = natom / cub(double(bl));
    for (int p = 0; p < npair; ++p)
      sq_transform(gr_nbin, double(gr_dr), rho, &gr_h[size_t(p) * gr_nbin], nq, q_min, q_max, &sq[size_t(p) * nq]);
  }
  // S(q_i) = sum_r 4 pi rho dr r sin(q_i r)/q_i h(r), r = (k+1/2) dr (src/grsq.cpp:270-279, :305-320)
  static void sq_transform(int nbin, double dr, double rho, const double* h, int nq, double q_min, double q_max, double* out) {
    const double dq = (q_max - q_min) / nq, rho4pidr = 4 * 3.141592653589793 * rho * dr;
    for (int iq = 0; iq < nq; ++iq) {
      double qi = q_min + (iq + 0.5) * dq, acc = 0;
      for (int ir = 0; ir < nbin; ++ir) {
        double r = (ir + 0.5) * dr;
        acc += std::sin(qi * r) * rho4pidr * r / qi * h[ir];
      }
      out[iq] = acc;
    }
  }

  // ---- RMDF (src/grsq.cpp:899-1377) ----------------------------------------------------------
  // Refines the pair table against experimental S(q): datasets n with rows
  // (q, S_exp, weight, form factor per element) on the S(q) grid q_i = (i+1/2) q_max / nq.
  int rm_nsq = 0, rm_nq = 0, rm_nr = 0, rm_target = 0;
  T rm_gamma = 0;
  std::vector<T> rm_sqex, rm_wt, rm_sqwt, rm_fwt, rm_ftmx, rm_potmx, rm_deltasq, rm_sq, rm_forceq;
  std::vector<T> rm_fbase, rm_ubase;  // base tables (pair: [npair][nb1]; NNP: [npair][4][nb1])
  std::vector<T> sq_T;                // partial S(q) of SQ::compute in T: [npair][nq]
  void rmdf_init(int nsq, int nq, double q_max, const float* data, T strength, T gamma, int target) {
    rm_nsq = nsq; rm_nq = nq; rm_gamma = gamma; rm_target = target;
    rm_nr = gr_nbin_inner;
    const int nt = ntype;
    rm_sqex.assign(size_t(nsq) * nq, 0); rm_wt.assign(size_t(nsq) * nq, 0);
    rm_sqwt.assign(size_t(nsq) * npair * nq, 0); rm_fwt.assign(size_t(nsq) * npair * nq, 0);
    rm_deltasq.assign(size_t(nsq) * nq, 0); rm_sq.assign(size_t(nsq) * nq, 0);
    rm_forceq.assign(size_t(npair) * nq, 0);
    const T natsq = sqr(T(natom));
    std::vector<T> ff(nt);
    for (int n = 0; n < nsq; ++n)
      for (int i = 0; i < nq; ++i) {
        const float* row = data + (size_t(n) * nq + i) * (3 + nt);
        rm_sqex[size_t(n) * nq + i] = T(row[1]);
        T wti = T(row[2]) * strength;
        rm_wt[size_t(n) * nq + i] = wti;
        T cifi2 = 0;
        for (int j = 0; j < nt; ++j) { ff[j] = T(row[3 + j]); cifi2 += ff[j] * T(typecount[j]); }
        cifi2 = sqr(cifi2 / T(natom));
        int ipr = 0;
        for (int j = 0; j < nt; ++j) {
          T fi = ff[j], ni = T(typecount[j]) / natsq;
          for (int k = j; k < nt; ++k, ++ipr) {
            T fw = fi * ff[k] / cifi2;
            rm_sqwt[(size_t(n) * npair + ipr) * nq + i] = fw * T(typecount[k]) * ni * T(k == j ? 1 : 2);
            rm_fwt[(size_t(n) * npair + ipr) * nq + i] = fw * wti;
          }
        }
      }
    rm_ftmx.assign(size_t(rm_nr) * nq, 0); rm_potmx.assign(size_t(rm_nr) * nq, 0);
    const T dq = T(q_max) / T(nq);
    for (int r = 0; r < rm_nr; ++r) {
      T rri = (T(r) + T(0.5F)) * gr_dr, fi = T(1) / sqr(rri);
      for (int j = 0; j < nq; ++j) {
        T qi = (T(j) + T(.5F)) * dq * rri;
        T j0 = std::sin(qi) / qi;
        rm_potmx[size_t(r) * nq + j] = j0;
        rm_ftmx[size_t(r) * nq + j] = fi * (j0 - std::cos(qi));
      }
    }
    if (target == 0) { rm_fbase = tbl_f; rm_ubase = tbl_u; }
    else { rm_fbase = tfpair; rm_ubase = tpair; }
  }
  T rmdf_qspace() {  // compute_Qspace (src/grsq.cpp:1079-1118) on the current partial S(q)
    const int nq = rm_nq;
    T erg = 0;
    std::fill(rm_forceq.begin(), rm_forceq.end(), T(0));
    for (int n = 0; n < rm_nsq; ++n) {
      T* sqi = &rm_sq[size_t(n) * nq];
      for (int i = 0; i < nq; ++i) sqi[i] = 0;
      for (int p = 0; p < npair; ++p)
        for (int i = 0; i < nq; ++i) sqi[i] += sq_T[size_t(p) * nq + i] * rm_sqwt[(size_t(n) * npair + p) * nq + i];
      for (int i = 0; i < nq; ++i) {
        T d = sqi[i] - rm_sqex[size_t(n) * nq + i];
        erg += sqr(d) * rm_wt[size_t(n) * nq + i];
        T& ds = rm_deltasq[size_t(n) * nq + i];
        ds += rm_gamma * (d - ds);
      }
      for (int p = 0; p < npair; ++p)
        for (int i = 0; i < nq; ++i)
          rm_forceq[size_t(p) * nq + i] += rm_deltasq[size_t(n) * nq + i] * rm_fwt[(size_t(n) * npair + p) * nq + i];
    }
    return erg * T(natom) * T(.25F);
  }
  // compute_Rspace / compute_Potential (src/grsq.cpp:1120-1210) and their NNP-table twins
  // (src/grsq.cpp:1213-1377): table = base + force_q . kernel, special class 3 scaled by 0.5
  void rmdf_tables(bool potential) {
    const int nb1 = (rm_target == 0 ? pair_nbin : nnp_nbin) + 1, nq = rm_nq;
    const std::vector<T>& K = potential ? rm_potmx : rm_ftmx;
    const std::vector<T>& base = potential ? rm_ubase : rm_fbase;
    std::vector<T>& tab = rm_target == 0 ? (potential ? tbl_u : tbl_f) : (potential ? tpair : tfpair);
    for (int p = 0; p < npair; ++p)
      for (int r = 0; r < rm_nr; ++r) {
        T temp = 0;
        for (int q = 0; q < nq; ++q) temp += rm_forceq[size_t(p) * nq + q] * K[size_t(r) * nq + q];
        if (rm_target == 0) {
          tab[size_t(p) * nb1 + r] = base[size_t(p) * nb1 + r] + temp;
        } else {
          tab[(size_t(p) * 4 + 0) * nb1 + r] = temp + base[(size_t(p) * 4 + 0) * nb1 + r];
          tab[(size_t(p) * 4 + 3) * nb1 + r] = temp * T(0.5) + base[(size_t(p) * 4 + 3) * nb1 + r];
        }
      }
  }

  // ---- NNPOTENTIAL_TBL (src/pair_nnp.cpp:433-793) -------------------------------------------
  static constexpr int ND = 20;
  int nnp_nbin = 0;
  T nnp_cut = 0, nnp_cutsq = 0, nnp_dr = 0, nnp_rdr = 0;
  std::vector<T> tdesc, tfdesc, tpair, tfpair;  // [npair][80][nbin+1], [npair][4][nbin+1]
  std::vector<T> desc, dedg;                    // [natom][20] by atom id
  // Table construction from (coef_p[npair][4][6], coef_w/d/r[npair][80]) (src/pair_nnp.cpp:521-590)
  void nnp_tbl_init(T cutoff, int nbin, const float* cp_in, const float* cw, const float* cd, const float* cr) {
    nnp_cutsq = sqr(nnp_cut = cutoff); nnp_nbin = nbin;
    const int nb1 = nbin + 1, szd = ND * 4;
    tdesc.assign(size_t(npair) * szd * nb1, 0); tfdesc.assign(size_t(npair) * szd * nb1, 0);
    tpair.assign(size_t(npair) * 4 * nb1, 0); tfpair.assign(size_t(npair) * 4 * nb1, 0);
    nnp_dr = cutoff / T(nbin); nnp_rdr = T(nbin) / cutoff;
    std::vector<T> cp(cp_in, cp_in + size_t(npair) * 24);
    for (int p = 0; p < npair; ++p) {
      for (int k = 0; k < 4; ++k) {
        T* c = &cp[size_t(p) * 24 + k * 6];
        T src2 = sqr(c[1] / cutoff), src6 = cub(src2);
        T elim = T(-2) * c[5] / cutoff - c[0] * src6 * (T(-7) + src2 * (T(9) * c[2] + src2 * (T(11) * c[3] + src2 * T(13) * c[4])));
        T flim = (c[5] / cutoff + c[0] * src6 * (T(-6) + src2 * ((c[2] *= T(8)) + src2 * ((c[3] *= T(10)) + src2 * (c[4] *= T(12)))))) / cutoff;
        c[1] *= c[1];
        T* pu = &tpair[(size_t(p) * 4 + k) * nb1];
        T* pf = &tfpair[(size_t(p) * 4 + k) * nb1];
        for (int i = 0; i < nbin; ++i) {
          T rr = T(double(nnp_dr) * (i + 0.5)), rsq = sqr(rr);
          T s2 = c[1] / rsq, s6 = cub(s2);
          pf[i] = (c[5] / rr + c[0] * s6 * (T(-6) + s2 * (c[2] + s2 * (c[3] + s2 * c[4])))) / rsq - flim / rr;
          pu[i] = c[5] / rr + c[0] * s6 * (T(-1) + s2 * (c[2] * T(.125F) + s2 * (c[3] * T(.1F) + s2 * c[4] * T(0.08333333333333333333F)))) + flim * rr + elim;
        }
        pu[nbin] = pf[nbin] = 0;
      }
      for (int i = 0; i < nbin; ++i) {
        T rr = T(double(nnp_dr) * (i + 0.5));
        T cf2 = rr - cutoff, rrcsq_1 = T(1) + sqr(cf2);
        T cff = T(1) - T(1) / rrcsq_1;
        T cfr = (T(-2) * cff) / rr;
        cf2 /= (T(-.5f) * rr * sqr(rrcsq_1));
        for (int k = 0; k < szd; ++k) {
          T rr0 = rr - T(cr[size_t(p) * szd + k]);
          T drr0 = T(cd[size_t(p) * szd + k]) * rr0;
          T wexp = std::exp(drr0 * rr0) * T(cw[size_t(p) * szd + k]);
          tfdesc[(size_t(p) * szd + k) * nb1 + i] = wexp * (cf2 + drr0 * cfr);
          tdesc[(size_t(p) * szd + k) * nb1 + i] = wexp * cff;
        }
      }
    }
  }
  void nnp_tbl_descriptor() {  // computeDescriptor (src/pair_nnp.cpp:625-688)
    desc.assign(size_t(natom) * ND, 0);
    const int nb1 = nnp_nbin + 1;
    for (int i = 0; i < natom; ++i) {
      const T xi = x[3 * i], yi = x[3 * i + 1], zi = x[3 * i + 2];
      T* id = &desc[size_t(atom_id[i]) * ND];
      for (int e = nl_start[i]; e < nl_start[i] + nl_num[i]; ++e) {
        int jj = nl_j[e]; const T* xj = &x[3 * jj];
        T rsq = sqr(xi - xj[0]) + sqr(yi - xj[1]) + sqr(zi - xj[2]);
        if (!(rsq < nnp_cutsq)) continue;
        T r = std::sqrt(rsq), ipw1 = r * nnp_rdr - T(0.5F);
        int t1 = int(ipw1), t2 = t1 + 1; T ipw2 = ipw1 - T(t1); ipw1 = T(1) - ipw2;
        int p = pairtype[typ[i] * ntype + typ[jj]], s = ND * nl_sp[e];
        bool own = jj < natom && !list_full;
        T* jd = &desc[size_t(atom_id[jj]) * ND];
        for (int k = 0; k < ND; ++k) {
          const T* t = &tdesc[(size_t(p) * ND * 4 + s + k) * nb1];
          T dk = t[t1] * ipw1 + t[t2] * ipw2;
          id[k] += dk;
          if (own) jd[k] += dk;
        }
      }
    }
  }
  void nnp_tbl_force(bool ev, double* erg, double* vir) {  // compute<ev> (src/pair_nnp.cpp:710-793)
    const int nb1 = nnp_nbin + 1;
    for (int i = 0; i < natom; ++i) {
      const T xi = x[3 * i], yi = x[3 * i + 1], zi = x[3 * i + 2];
      const T* ig = &dedg[size_t(atom_id[i]) * ND];
      T fx = 0, fy = 0, fz = 0;
      for (int e = nl_start[i]; e < nl_start[i] + nl_num[i]; ++e) {
        int jj = nl_j[e]; const T* xj = &x[3 * jj];
        T dx = xi - xj[0], dy = yi - xj[1], dz = zi - xj[2];
        T rsq = sqr(dx) + sqr(dy) + sqr(dz);
        if (!(rsq < nnp_cutsq)) continue;
        T r = std::sqrt(rsq), ipw1 = r * nnp_rdr - T(0.5F);
        int t1 = int(ipw1), t2 = t1 + 1; T ipw2 = ipw1 - T(t1); ipw1 = T(1) - ipw2;
        int p = pairtype[typ[i] * ntype + typ[jj]], sp = nl_sp[e];
        const T* fp = &tfpair[(size_t(p) * 4 + sp) * nb1];
        T fpair = fp[t1] * ipw1 + fp[t2] * ipw2;
        const T* jg = &dedg[size_t(atom_id[jj]) * ND];
        for (int k = 0; k < ND; ++k) {
          const T* t = &tfdesc[(size_t(p) * ND * 4 + ND * sp + k) * nb1];
          fpair += (ig[k] + jg[k]) * (t[t1] * ipw1 + t[t2] * ipw2);
        }
        fx += (dx *= fpair); fy += (dy *= fpair); fz += (dz *= fpair);
        bool own = jj < natom && !list_full;
        if (own) { T* fj = &f[3 * jj]; fj[0] -= dx; fj[1] -= dy; fj[2] -= dz; }
        if (ev) {
          const T* pu = &tpair[(size_t(p) * 4 + sp) * nb1];
          double u = double(pu[t1] * ipw1 + pu[t2] * ipw2), w = double(fpair * rsq);
          double s = own ? 1.0 : 0.5;
          *erg += s * u; *vir += s * w;
        }
      }
      f[3 * i] += fx; f[3 * i + 1] += fy; f[3 * i + 2] += fz;
    }
  }

  // ---- per-element MLP (replaces TFCALL::evaluate, src/pair_nnp.cpp:92-97) ------------------
  // E_i = out_scale * (W_L tanh(... tanh(W_1 G_i + b_1) ...) + b_L); weights row-major [in][out].
  int mlp_nl = 0;  // number of weight layers
  std::vector<int> mlp_w;  // widths [nl+1]
  std::vector<double> mlp_par;
  std::vector<size_t> mlp_off;  // per (type, layer): offset of W; b follows W
  double mlp_scale = 1;
  std::vector<double> mlp_shift, mlp_inscale;  // input standardisation x = (G - shift) * scale, [ntype][D]
  void mlp_set(int nl, const int* widths, const float* par, float out_scale, const float* shift, const float* scale) {
    mlp_nl = nl; mlp_w.assign(widths, widths + nl + 1); mlp_scale = out_scale;
    size_t o = 0; mlp_off.clear();
    for (int t = 0; t < ntype; ++t) for (int l = 0; l < nl; ++l) { mlp_off.push_back(o); o += size_t(mlp_w[l]) * mlp_w[l + 1] + mlp_w[l + 1]; }
    mlp_par.assign(par, par + o);
    const size_t nd = size_t(ntype) * widths[0];
    mlp_shift.assign(nd, 0.0); mlp_inscale.assign(nd, 1.0);
    if (shift) for (size_t k = 0; k < nd; ++k) mlp_shift[k] = shift[k];
    if (scale) for (size_t k = 0; k < nd; ++k) mlp_inscale[k] = scale[k];
  }
  // forward + input-gradient backward in double for one descriptor row
  double mlp_eval_row(int t, const double* g, double* dg) const {
    std::vector<std::vector<double>> act(mlp_nl + 1);
    const int D0 = mlp_w[0];
    act[0].resize(D0);
    for (int k = 0; k < D0; ++k) act[0][k] = (g[k] - mlp_shift[size_t(t) * D0 + k]) * mlp_inscale[size_t(t) * D0 + k];
    for (int l = 0; l < mlp_nl; ++l) {
      const double* W = &mlp_par[mlp_off[size_t(t) * mlp_nl + l]];
      const double* b = W + size_t(mlp_w[l]) * mlp_w[l + 1];
      act[l + 1].assign(mlp_w[l + 1], 0);
      for (int o = 0; o < mlp_w[l + 1]; ++o) {
        double z = b[o];
        for (int k = 0; k < mlp_w[l]; ++k) z += act[l][k] * W[size_t(k) * mlp_w[l + 1] + o];
        act[l + 1][o] = (l + 1 < mlp_nl) ? std::tanh(z) : z;
      }
    }
    double e = act[mlp_nl][0] * mlp_scale;
    std::vector<double> delta(1, mlp_scale);
    for (int l = mlp_nl - 1; l >= 0; --l) {
      const double* W = &mlp_par[mlp_off[size_t(t) * mlp_nl + l]];
      std::vector<double> prev(mlp_w[l], 0);
      for (int k = 0; k < mlp_w[l]; ++k) {
        double s = 0;
        for (int o = 0; o < mlp_w[l + 1]; ++o) s += W[size_t(k) * mlp_w[l + 1] + o] * delta[o];
        prev[k] = (l > 0) ? s * (1 - act[l][k] * act[l][k]) : s;
      }
      delta.swap(prev);
    }
    for (int k = 0; k < D0; ++k) dg[k] = delta[k] * mlp_inscale[size_t(t) * D0 + k];
    return e;
  }
  // descriptor rows are indexed by atom id; returns total NN energy
  double mlp_eval(const std::vector<T>& G, std::vector<T>& dG, int D, std::vector<double>* e_atom) {
    dG.assign(size_t(natom) * D, 0);
    if (e_atom) e_atom->assign(natom, 0);
    std::vector<int> typ_by_id(natom);
    for (int i = 0; i < natom; ++i) typ_by_id[atom_id[i]] = typ[i];
    double etot = 0;
    std::vector<double> g(D), dg(D);
    for (int a = 0; a < natom; ++a) {
      for (int k = 0; k < D; ++k) g[k] = double(G[size_t(a) * D + k]);
      double e = mlp_eval_row(typ_by_id[a], g.data(), dg.data());
      for (int k = 0; k < D; ++k) dG[size_t(a) * D + k] = T(dg[k]);
      if (e_atom) (*e_atom)[a] = e;
      etot += e;
    }
    return etot;
  }

  // ---- symmetry-function NNP (radial + separable angular), used for the benchmark workloads --
  // Radial:  G^R_{i,a,n} = sum_{j in a} exp(-eta_r (r_ij - mu_n)^2) fc(r_ij)
  // Angular: G^A_{i,(a<=b),m,(zeta,lambda)} = 2^(1-zeta) sum_{j in a, k in b, j != k}
  //            (1 + lambda cos t_jik)^zeta w_m(r_ij) w_m(r_ik),  w_m(r) = exp(-eta_m r^2) fc(r)
  //          for zeta in {1,2}, lambda in {+1,-1}, evaluated through the moment sums
  //          A0 = sum w, A1 = sum w r^, A2 = sum w r^ r^, S = sum w^2 per neighbour element.
  // fc(r) = (cos(pi r / rc) + 1) / 2.
  int sf_nrad = 0, sf_neta = 0, sf_D = 0;
  T sf_rc = 0;
  std::vector<T> sf_mu, sf_eta;
  T sf_eta_r = 0;
  std::vector<T> sf_G, sf_dG;  // by atom id
  void sf_set(T rc, int nrad, const float* mu, float eta_r, int neta, const float* eta) {
    sf_rc = rc; sf_nrad = nrad; sf_neta = neta; sf_eta_r = eta_r;
    sf_mu.assign(mu, mu + nrad); sf_eta.assign(eta, eta + neta);
    sf_D = nrad * ntype + 4 * neta * npair;
  }
  void sf_fc(T r, T& fcv, T& dfc) const {
    const T pi = T(3.141592653589793);
    fcv = T(0.5) * (std::cos(pi * r / sf_rc) + T(1));
    dfc = T(-0.5) * pi / sf_rc * std::sin(pi * r / sf_rc);
  }
  int nmom() const { return 11; }  // A0, A1[3], A2[6] (xx,yy,zz,xy,xz,yz), S
  // moment accumulators of one atom: [ntype][neta][11] then radial [ntype][nrad]
  void sf_moments(int i, std::vector<T>& M) const {
    M.assign(size_t(ntype) * sf_neta * 11 + size_t(ntype) * sf_nrad, 0);
    T* R = &M[size_t(ntype) * sf_neta * 11];
    const T xi = x[3 * i], yi = x[3 * i + 1], zi = x[3 * i + 2];
    for (int e = nl_start[i]; e < nl_start[i] + nl_num[i]; ++e) {
      int jj = nl_j[e]; const T* xj = &x[3 * jj];
      T ux = xj[0] - xi, uy = xj[1] - yi, uz = xj[2] - zi;
      T rsq = ux * ux + uy * uy + uz * uz;
      if (!(rsq < sf_rc * sf_rc)) continue;
      T r = std::sqrt(rsq), fcv, dfc; sf_fc(r, fcv, dfc);
      T h[3] = {ux / r, uy / r, uz / r};
      int a = typ[jj];
      for (int m = 0; m < sf_neta; ++m) {
        T w = std::exp(-sf_eta[m] * rsq) * fcv;
        T* A = &M[(size_t(a) * sf_neta + m) * 11];
        A[0] += w;
        A[1] += w * h[0]; A[2] += w * h[1]; A[3] += w * h[2];
        A[4] += w * h[0] * h[0]; A[5] += w * h[1] * h[1]; A[6] += w * h[2] * h[2];
        A[7] += w * h[0] * h[1]; A[8] += w * h[0] * h[2]; A[9] += w * h[1] * h[2];
        A[10] += w * w;
      }
      for (int n = 0; n < sf_nrad; ++n) R[size_t(a) * sf_nrad + n] += std::exp(-sf_eta_r * sqr(r - sf_mu[n])) * fcv;
    }
  }
  void sf_descriptor() {
    sf_G.assign(size_t(natom) * sf_D, 0);
    std::vector<T> M;
    for (int i = 0; i < natom; ++i) {
      sf_moments(i, M);
      T* G = &sf_G[size_t(atom_id[i]) * sf_D];
      const T* R = &M[size_t(ntype) * sf_neta * 11];
      for (int k = 0; k < ntype * sf_nrad; ++k) G[k] = R[k];
      int off = ntype * sf_nrad, pp = 0;
      for (int a = 0; a < ntype; ++a) for (int b = a; b < ntype; ++b, ++pp)
        for (int m = 0; m < sf_neta; ++m) {
          const T* A = &M[(size_t(a) * sf_neta + m) * 11];
          const T* B = &M[(size_t(b) * sf_neta + m) * 11];
          T P0 = A[0] * B[0], P1 = A[1] * B[1] + A[2] * B[2] + A[3] * B[3];
          T P2 = A[4] * B[4] + A[5] * B[5] + A[6] * B[6] + T(2) * (A[7] * B[7] + A[8] * B[8] + A[9] * B[9]);
          if (a == b) { P0 -= A[10]; P1 -= A[10]; P2 -= A[10]; }
          T* g = &G[off + (pp * sf_neta + m) * 4];
          g[0] = P0 + P1; g[1] = P0 - P1;
          g[2] = T(0.5) * (P0 + T(2) * P1 + P2); g[3] = T(0.5) * (P0 - T(2) * P1 + P2);
        }
    }
  }
  // adjoint of the moments of atom i given dE/dG_i
  void sf_adjoint(int i, const T* dG, std::vector<T>& adj) const {
    std::vector<T> M; sf_moments(i, M);
    adj.assign(M.size(), 0);
    T* aR = &adj[size_t(ntype) * sf_neta * 11];
    for (int k = 0; k < ntype * sf_nrad; ++k) aR[k] = dG[k];
    int off = ntype * sf_nrad, pp = 0;
    for (int a = 0; a < ntype; ++a) for (int b = a; b < ntype; ++b, ++pp)
      for (int m = 0; m < sf_neta; ++m) {
        const T* g = &dG[off + (pp * sf_neta + m) * 4];
        T u0 = g[0] + g[1] + T(0.5) * (g[2] + g[3]);
        T u1 = g[0] - g[1] + g[2] - g[3];
        T u2 = T(0.5) * (g[2] + g[3]);
        const T* A = &M[(size_t(a) * sf_neta + m) * 11];
        const T* B = &M[(size_t(b) * sf_neta + m) * 11];
        T* dA = &adj[(size_t(a) * sf_neta + m) * 11];
        T* dB = &adj[(size_t(b) * sf_neta + m) * 11];
        dA[0] += u0 * B[0]; dB[0] += u0 * A[0];
        for (int c = 1; c <= 3; ++c) { dA[c] += u1 * B[c]; dB[c] += u1 * A[c]; }
        for (int c = 4; c <= 6; ++c) { dA[c] += u2 * B[c]; dB[c] += u2 * A[c]; }
        for (int c = 7; c <= 9; ++c) { dA[c] += T(2) * u2 * B[c]; dB[c] += T(2) * u2 * A[c]; }
        if (a == b) dA[10] -= u0 + u1 + u2;
      }
  }
  // dE_i/du for one neighbour at u = x_j - x_i, element a, from atom i's adjoint
  void sf_grad(const T* adj, int a, const T u[3], T r, T g[3]) const {
    T fcv, dfc; sf_fc(r, fcv, dfc);
    T h[3] = {u[0] / r, u[1] / r, u[2] / r};
    g[0] = g[1] = g[2] = 0;
    T radial = 0;
    const T* aR = &adj[size_t(ntype) * sf_neta * 11 + size_t(a) * sf_nrad];
    for (int n = 0; n < sf_nrad; ++n) {
      T ex = std::exp(-sf_eta_r * sqr(r - sf_mu[n]));
      radial += aR[n] * (ex * dfc - T(2) * sf_eta_r * (r - sf_mu[n]) * ex * fcv);
    }
    T along = radial;
    for (int m = 0; m < sf_neta; ++m) {
      const T* d = &adj[(size_t(a) * sf_neta + m) * 11];
      T ex = std::exp(-sf_eta[m] * r * r);
      T w = ex * fcv, dw = ex * dfc - T(2) * sf_eta[m] * r * ex * fcv;
      T p1 = d[1] * h[0] + d[2] * h[1] + d[3] * h[2];
      T Mf[3][3] = {{d[4], d[7] / 2, d[8] / 2}, {d[7] / 2, d[5], d[9] / 2}, {d[8] / 2, d[9] / 2, d[6]}};
      T Mr[3]; for (int c = 0; c < 3; ++c) Mr[c] = Mf[c][0] * h[0] + Mf[c][1] * h[1] + Mf[c][2] * h[2];
      T qq = Mr[0] * h[0] + Mr[1] * h[1] + Mr[2] * h[2];
      along += dw * (d[0] + T(2) * w * d[10] + p1 + qq) - (w / r) * (p1 + T(2) * qq);
      for (int c = 0; c < 3; ++c) g[c] += (w / r) * (d[1 + c] + T(2) * Mr[c]);
    }
    for (int c = 0; c < 3; ++c) g[c] += along * h[c];
  }
  // forces (added to f) and energy/virial given dE/dG by atom id in sf_dG
  void sf_force(bool ev, double* vir) {
    std::vector<std::vector<T>> adj(natom);
    for (int i = 0; i < natom; ++i) sf_adjoint(i, &sf_dG[size_t(atom_id[i]) * sf_D], adj[i]);
    for (int i = 0; i < natom; ++i) {
      const T xi = x[3 * i], yi = x[3 * i + 1], zi = x[3 * i + 2];
      for (int e = nl_start[i]; e < nl_start[i] + nl_num[i]; ++e) {
        int jj = nl_j[e]; const T* xj = &x[3 * jj];
        T u[3] = {xj[0] - xi, xj[1] - yi, xj[2] - zi};
        T rsq = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
        if (!(rsq < sf_rc * sf_rc)) continue;
        T r = std::sqrt(rsq);
        int j = jj < natom ? jj : i_ghost[jj - natom];
        T gi[3], gj[3], mu[3] = {-u[0], -u[1], -u[2]};
        sf_grad(adj[i].data(), typ[jj], u, r, gi);    // dE_i/du
        sf_grad(adj[j].data(), typ[i], mu, r, gj);    // dE_j/d(-u)
        // F_i = dE_i/du - dE_j/du'
        for (int c = 0; c < 3; ++c) f[3 * i + c] += gi[c] - gj[c];
        bool own = jj < natom && !list_full;
        if (own) for (int c = 0; c < 3; ++c) f[3 * j + c] += gj[c] - gi[c];
        if (ev) {
          double w = -double(gi[0] * u[0] + gi[1] * u[1] + gi[2] * u[2]);
          double wj = -double(gj[0] * mu[0] + gj[1] * mu[1] + gj[2] * mu[2]);
          *vir += w + (own ? wj : 0.0);
        }
      }
    }
  }
};

}  // namespace

// ============================================================================================
// SQD: direct k-space S(q) (src/grsq.cpp:362-606).
//  * The vector set, |G| and the bins follow the reference's float arithmetic (numtype = float):
//    g = two_pi / bl, kcutoff = int(q_max / g), cuti2 = int(sqr(q_max / g)) (:370-376);
//    the three loops over the half space (:378-423); gscl = sqrt(float(n^2)) * g; the vectors
//    sorted by gscl (:424, any order among equal values); vbegin/vend from gx = i * dq,
//    dq = q_max / nq (:436-455).  qq_ref is the reference's float mean in sorted order
//    (:440-455); qq is the FP64 mean of the same float gscl values.
//  * SQD::compute (:571-606) in FP64: G = 2 pi (h, k, l) / L, kr = G . x, S_a / C_a = sums of
//    sin / cos over the atoms of element a; per bin avg(S_a S_b + C_a C_b) * natom / (N_a N_b)
//    - [a == b] natom / N_a (:476-489).  Empty bins get 0 (the reference divides by zero there
//    and SQD::dump skips them, :629).
void sqd_restate(int natom, int ntype, float bl, const float* x, const int* typ, int nq, float q_max, double* qq,
                 double* qq_ref, double* sq, long long* count) {
  const float pi_f = 3.14159265358979323846f, two_pi_f = pi_f * 2;
  const float g = two_pi_f / bl;
  const int kc = (int)(q_max / g);
  const int cuti2 = (int)sqr(q_max / g);
  struct V { int h, k, l; float s; };
  std::vector<V> vs;
  for (int i = -kc; i <= kc; ++i)
    for (int j = -kc; j <= kc; ++j)
      for (int k = 1; k <= kc; ++k) {
        const int n2 = i * i + j * j + k * k;
        if (n2 > cuti2) continue;
        vs.push_back({i, j, k, std::sqrt((float)n2) * g});
      }
  for (int i = -kc; i <= kc; ++i)
    for (int j = 1; j <= kc; ++j) {
      const int n2 = i * i + j * j;
      if (n2 > cuti2) continue;
      vs.push_back({i, j, 0, std::sqrt((float)n2) * g});
    }
  for (int i = 1; i <= kc; ++i) {
    if (i * i > cuti2) continue;
    vs.push_back({i, 0, 0, std::sqrt((float)(i * i)) * g});
  }
  std::stable_sort(vs.begin(), vs.end(), [](const V& a, const V& b) { return a.s < b.s; });
  const int ngv = (int)vs.size();
  const float dq = q_max / nq;
  std::vector<int> vb(nq), ve(nq);
  int i2 = 0;
  vb[0] = 0;
  for (int i = 1; i < nq; ++i) {
    const float gx = i * dq;
    float gy = 0;
    const int b0 = i2;
    while (i2 < ngv && vs[i2].s < gx) { gy += vs[i2].s; ++i2; }
    vb[i] = ve[i - 1] = i2;
    qq_ref[i - 1] = ve[i - 1] > b0 ? double(gy / (ve[i - 1] - b0)) : 0.0;
  }
  {
    float gy = 0;
    const int b0 = i2;
    while (i2 < ngv) { gy += vs[i2].s; ++i2; }
    ve[nq - 1] = ngv;
    qq_ref[nq - 1] = ngv > b0 ? double(gy / (ngv - b0)) : 0.0;
  }
  std::vector<int> tc(ntype, 0);
  for (int a = 0; a < natom; ++a) ++tc[typ[a]];
  const int npair = ntype * (ntype + 1) / 2;
  const double two_pi_L = 2 * 3.14159265358979323846 / double(bl);
  std::vector<double> ss(ntype), cc(ntype);
  for (int b = 0; b < nq; ++b) {
    count[b] = ve[b] - vb[b];
    double qs = 0;
    std::vector<double> acc(npair, 0.0);
    for (int v = vb[b]; v < ve[b]; ++v) {
      qs += vs[v].s;
      const double qx = two_pi_L * vs[v].h, qy = two_pi_L * vs[v].k, qz = two_pi_L * vs[v].l;
      std::fill(ss.begin(), ss.end(), 0.0);
      std::fill(cc.begin(), cc.end(), 0.0);
      for (int a = 0; a < natom; ++a) {
        const double kr = qx * double(x[3 * a]) + qy * double(x[3 * a + 1]) + qz * double(x[3 * a + 2]);
        ss[typ[a]] += std::sin(kr);
        cc[typ[a]] += std::cos(kr);
      }
      int p = 0;
      for (int t = 0; t < ntype; ++t)
        for (int u = t; u < ntype; ++u, ++p) acc[p] += ss[t] * ss[u] + cc[t] * cc[u];
    }
    qq[b] = count[b] ? qs / double(count[b]) : 0.0;
    int p = 0;
    for (int t = 0; t < ntype; ++t)
      for (int u = t; u < ntype; ++u, ++p) {
        const double factor = double(natom) / (double(tc[t]) * tc[u]);
        const double base = t == u ? double(natom) / tc[u] : 0.0;
        sq[size_t(b) * npair + p] = count[b] ? acc[p] / double(count[b]) * factor - base : 0.0;
      }
  }
}

// ============================================================================================
// C ABI (ctypes).  Handles are Eng<float> (prec 0) or Eng<double> (prec 1).
extern "C" {

struct OrcHandle { int prec; void* e; };

#define ORC_DISPATCH(h, call) ((h)->prec ? static_cast<Eng<double>*>((h)->e)->call : static_cast<Eng<float>*>((h)->e)->call)

OrcHandle* orc_create(int prec) {
  OrcHandle* h = new OrcHandle;
  h->prec = prec;
  h->e = prec ? static_cast<void*>(new Eng<double>) : static_cast<void*>(new Eng<float>);
  return h;
}
void orc_destroy(OrcHandle* h) {
  if (!h) return;
  if (h->prec) delete static_cast<Eng<double>*>(h->e); else delete static_cast<Eng<float>*>(h->e);
  delete h;
}

int orc_init(OrcHandle* h, int natom, int ntype, int nmol, int nmoltype, float bl, const float* x, const float* v,
             const float* q, const float* mass, const int* typ, const int* mol_id, const int* intra_id,
             const int* mol_type, const int* mt_natom, const int* mt_special_off, const int* mt_special) {
  SysIn s{natom, ntype, nmol, nmoltype, bl, x, v, q, mass, typ, mol_id, intra_id, mol_type, mt_natom, mt_special_off, mt_special};
  ORC_DISPATCH(h, init(s));
  return 0;
}

// run.cpp:72-82: comm_cutoff = skin + max(gr_cutoff, pair_cutoff); grid = (skin + pair_cutoff)/2
int orc_neigh_init(OrcHandle* h, float pair_cutoff, float skin, float gr_cutoff) {
  float comm = skin + std::max(gr_cutoff, pair_cutoff);
  float gl = float(0.5 * (skin + pair_cutoff));
  if (h->prec) {
    auto* e = static_cast<Eng<double>*>(h->e);
    e->init_grid(gl, comm);
    e->bin_list_pair = e->build_bin_list(double(pair_cutoff + skin));
    e->bin_list_gr = e->build_bin_list(double(gr_cutoff + skin));
    e->cutsq_list = sqr(double(pair_cutoff + skin)); e->dx2 = sqr(double(skin) / 2);
  } else {
    auto* e = static_cast<Eng<float>*>(h->e);
    e->init_grid(gl, comm);
    e->bin_list_pair = e->build_bin_list(pair_cutoff + skin);
    e->bin_list_gr = e->build_bin_list(gr_cutoff + skin);
    e->cutsq_list = sqr(pair_cutoff + skin); e->dx2 = sqr(skin / 2);
  }
  return 0;
}

int orc_num_gx(OrcHandle* h) { return ORC_DISPATCH(h, num_gx); }
int orc_rebuild(OrcHandle* h, int full) { ORC_DISPATCH(h, rebuild(full != 0)); return ORC_DISPATCH(h, nghost); }
int orc_update_ghost_pos(OrcHandle* h) { ORC_DISPATCH(h, update_ghost_pos()); return 0; }
int orc_isvalid(OrcHandle* h) { return ORC_DISPATCH(h, isvalid()) ? 1 : 0; }
long long orc_list_size(OrcHandle* h) { return (long long)ORC_DISPATCH(h, nl_j.size()); }

// Directed list entries as (atom id i, atom id j, image code, special class).
long long orc_list_export(OrcHandle* h, int* ai, int* aj, int* code, int* sp) {
  long long n = 0;
  auto run = [&](auto* e) {
    for (int i = 0; i < e->natom; ++i)
      for (int k = e->nl_start[i]; k < e->nl_start[i] + e->nl_num[i]; ++k) {
        int jj = e->nl_j[k];
        ai[n] = e->atom_id[i]; aj[n] = e->atom_id[jj];
        code[n] = jj < e->natom ? 13 : e->g_code[jj - e->natom];
        sp[n] = e->nl_sp[k];
        ++n;
      }
  };
  if (h->prec) run(static_cast<Eng<double>*>(h->e)); else run(static_cast<Eng<float>*>(h->e));
  return n;
}

// per-atom state by atom id: what 0=x 1=v 2=f 3=f_prev
int orc_get(OrcHandle* h, int what, double* out) {
  auto run = [&](auto* e) {
    for (int i = 0; i < e->natom; ++i) {
      int a = e->atom_id[i];
      for (int d = 0; d < 3; ++d) {
        double val = what == 0 ? double(e->x[3 * i + d]) : what == 1 ? double(e->v[3 * i + d]) : what == 2 ? double(e->f[3 * i + d]) : double(e->f_prev[3 * i + d]);
        out[3 * a + d] = val;
      }
    }
  };
  if (h->prec) run(static_cast<Eng<double>*>(h->e)); else run(static_cast<Eng<float>*>(h->e));
  return 0;
}
int orc_set(OrcHandle* h, int what, const double* in) {
  auto run = [&](auto* e) {
    using TT = typename std::remove_reference<decltype(e->x[0])>::type;
    for (int i = 0; i < e->natom; ++i) {
      int a = e->atom_id[i];
      for (int d = 0; d < 3; ++d) {
        TT val = TT(in[3 * a + d]);
        if (what == 0) e->x[3 * i + d] = val; else if (what == 1) e->v[3 * i + d] = val; else if (what == 2) e->f[3 * i + d] = val; else e->f_prev[3 * i + d] = val;
      }
    }
  };
  if (h->prec) run(static_cast<Eng<double>*>(h->e)); else run(static_cast<Eng<float>*>(h->e));
  return 0;
}
int orc_get_cell_ids(OrcHandle* h, int* cell_by_id) {
  auto run = [&](auto* e) { for (int i = 0; i < e->natom; ++i) cell_by_id[e->atom_id[i]] = e->x_grd[i]; };
  if (h->prec) run(static_cast<Eng<double>*>(h->e)); else run(static_cast<Eng<float>*>(h->e));
  return 0;
}
// x_mol image codes by atom id (13 = the real atom)
int orc_get_xmol_codes(OrcHandle* h, int* code_by_id) {
  auto run = [&](auto* e) {
    for (int a = 0; a < e->natom; ++a) { int row = e->xmol_row[a]; code_by_id[a] = row < e->natom ? 13 : e->g_code[row - e->natom]; }
  };
  if (h->prec) run(static_cast<Eng<double>*>(h->e)); else run(static_cast<Eng<float>*>(h->e));
  return 0;
}

int orc_clear_f(OrcHandle* h) { ORC_DISPATCH(h, clear_f()); return 0; }
int orc_pair_set(OrcHandle* h, float cutoff, int nbin, const float* force, const float* pot) {
  if (h->prec) static_cast<Eng<double>*>(h->e)->pair_set(double(cutoff), nbin, force, pot);
  else static_cast<Eng<float>*>(h->e)->pair_set(cutoff, nbin, force, pot);
  return 0;
}
int orc_pair_compute(OrcHandle* h, int ev, double* erg, double* vir) { ORC_DISPATCH(h, pair_compute(ev != 0, erg, vir)); return 0; }
int orc_bond_compute(OrcHandle* h, int n, const int* t, const float* c, int ev, double* erg, double* vir) { ORC_DISPATCH(h, bond_compute(n, t, c, ev != 0, erg, vir)); return 0; }
int orc_angle_compute(OrcHandle* h, int n, const int* t, const float* c, int ev, double* erg) { ORC_DISPATCH(h, angle_compute(n, t, c, ev != 0, erg)); return 0; }
int orc_dihedral_compute(OrcHandle* h, int n, const int* t, const float* c, int ev, double* erg) { ORC_DISPATCH(h, dihedral_compute(n, t, c, ev != 0, erg)); return 0; }
int orc_dihedral_poly_compute(OrcHandle* h, int n, const int* t, const float* c, int ev, double* erg) { ORC_DISPATCH(h, dihedral_compute(n, t, c, ev != 0, erg, true)); return 0; }
int orc_set_dt(OrcHandle* h, double dt) {
  if (h->prec) static_cast<Eng<double>*>(h->e)->set_dt(dt); else static_cast<Eng<float>*>(h->e)->set_dt(float(dt));
  return 0;
}
int orc_update_x(OrcHandle* h) { ORC_DISPATCH(h, update_x()); return 0; }
int orc_update_v(OrcHandle* h) { ORC_DISPATCH(h, update_v()); return 0; }
double orc_compute_ke(OrcHandle* h) { return ORC_DISPATCH(h, compute_ke()); }

int orc_gr_init(OrcHandle* h, int nbi, float r_max, float r_max_inner) {
  if (h->prec) static_cast<Eng<double>*>(h->e)->gr_init(nbi, double(r_max), double(r_max_inner));
  else static_cast<Eng<float>*>(h->e)->gr_init(nbi, r_max, r_max_inner);
  return ORC_DISPATCH(h, gr_nbin);
}
int orc_gr_tally(OrcHandle* h, int full) { if (full) ORC_DISPATCH(h, gr_tally_full()); else ORC_DISPATCH(h, gr_tally_inner()); return 0; }
int orc_gr_get(OrcHandle* h, long long* counts, double* hr) {
  auto run = [&](auto* e) {
    e->gr_reduce();
    for (size_t k = 0; k < e->gr_count.size(); ++k) { counts[k] = e->gr_count[k]; if (hr) hr[k] = e->gr_h[k]; }
  };
  if (h->prec) run(static_cast<Eng<double>*>(h->e)); else run(static_cast<Eng<float>*>(h->e));
  return 0;
}
int orc_sq(OrcHandle* h, int nq, double q_min, double q_max, double* out) {
  auto run = [&](auto* e) { e->gr_reduce(); e->sq_compute(nq, q_min, q_max); std::copy(e->sq.begin(), e->sq.end(), out); };
  if (h->prec) run(static_cast<Eng<double>*>(h->e)); else run(static_cast<Eng<float>*>(h->e));
  return 0;
}

int orc_rmdf_init(OrcHandle* h, int nsq, int nq, double q_max, const float* data, float strength, float gamma, int target) {
  if (h->prec) static_cast<Eng<double>*>(h->e)->rmdf_init(nsq, nq, q_max, data, strength, gamma, target);
  else static_cast<Eng<float>*>(h->e)->rmdf_init(nsq, nq, q_max, data, strength, gamma, target);
  return 0;
}
// one refinement step on the current g(r): S(q) (q_min = 0), compute_Qspace, compute_Rspace and,
// when potential != 0, compute_Potential; returns the Q-space energy
double orc_rmdf_step(OrcHandle* h, double q_max, int potential, double* sq_out, double* deltasq_out, double* forceq_out) {
  double erg = 0;
  auto run = [&](auto* e) {
    using TT = typename std::remove_reference<decltype(e->x[0])>::type;
    e->gr_reduce();
    e->sq_compute(e->rm_nq, 0.0, q_max);
    e->sq_T.assign(e->sq.begin(), e->sq.end());
    erg = double(e->rmdf_qspace());
    e->rmdf_tables(false);
    if (potential) e->rmdf_tables(true);
    for (size_t k = 0; k < e->rm_sq.size(); ++k) { if (sq_out) sq_out[k] = double(e->rm_sq[k]); if (deltasq_out) deltasq_out[k] = double(e->rm_deltasq[k]); }
    if (forceq_out) for (size_t k = 0; k < e->rm_forceq.size(); ++k) forceq_out[k] = double(e->rm_forceq[k]);
    (void)sizeof(TT);
  };
  if (h->prec) run(static_cast<Eng<double>*>(h->e)); else run(static_cast<Eng<float>*>(h->e));
  return erg;
}
// current pair tables (target 0: LJTableCoulDSF force/potential [npair][nb1]; 1: NNP pair tables [npair][4][nb1])
int orc_rmdf_tables(OrcHandle* h, double* f_out, double* u_out) {
  auto run = [&](auto* e) {
    const auto& F = e->rm_target == 0 ? e->tbl_f : e->tfpair;
    const auto& U = e->rm_target == 0 ? e->tbl_u : e->tpair;
    for (size_t k = 0; k < F.size(); ++k) { f_out[k] = double(F[k]); u_out[k] = double(U[k]); }
  };
  if (h->prec) run(static_cast<Eng<double>*>(h->e)); else run(static_cast<Eng<float>*>(h->e));
  return 0;
}

int orc_sq_transform(int nbin, double dr, double rho, const double* hr, int nq, double q_min, double q_max, double* out) {
  Eng<double>::sq_transform(nbin, dr, rho, hr, nq, q_min, q_max, out);
  return 0;
}

int orc_nnp_tbl_init(OrcHandle* h, float cutoff, int nbin, const float* cp, const float* cw, const float* cd, const float* cr) {
  if (h->prec) static_cast<Eng<double>*>(h->e)->nnp_tbl_init(double(cutoff), nbin, cp, cw, cd, cr);
  else static_cast<Eng<float>*>(h->e)->nnp_tbl_init(cutoff, nbin, cp, cw, cd, cr);
  return 0;
}
int orc_nnp_tbl_tables(OrcHandle* h, double* tdesc, double* tfdesc, double* tpair, double* tfpair) {
  auto run = [&](auto* e) {
    std::copy(e->tdesc.begin(), e->tdesc.end(), tdesc); std::copy(e->tfdesc.begin(), e->tfdesc.end(), tfdesc);
    std::copy(e->tpair.begin(), e->tpair.end(), tpair); std::copy(e->tfpair.begin(), e->tfpair.end(), tfpair);
  };
  if (h->prec) run(static_cast<Eng<double>*>(h->e)); else run(static_cast<Eng<float>*>(h->e));
  return 0;
}
int orc_nnp_tbl_descriptor(OrcHandle* h, double* out) {
  auto run = [&](auto* e) { e->nnp_tbl_descriptor(); for (size_t k = 0; k < e->desc.size(); ++k) out[k] = double(e->desc[k]); };
  if (h->prec) run(static_cast<Eng<double>*>(h->e)); else run(static_cast<Eng<float>*>(h->e));
  return 0;
}
int orc_nnp_tbl_set_dedg(OrcHandle* h, const double* dedg) {
  auto run = [&](auto* e) {
    using TT = typename std::remove_reference<decltype(e->x[0])>::type;
    e->dedg.resize(size_t(e->natom) * 20);
    for (size_t k = 0; k < e->dedg.size(); ++k) e->dedg[k] = TT(dedg[k]);
  };
  if (h->prec) run(static_cast<Eng<double>*>(h->e)); else run(static_cast<Eng<float>*>(h->e));
  return 0;
}
int orc_nnp_tbl_force(OrcHandle* h, int ev, double* erg, double* vir) { ORC_DISPATCH(h, nnp_tbl_force(ev != 0, erg, vir)); return 0; }

int orc_mlp_set(OrcHandle* h, int nl, const int* widths, const float* par, float out_scale, const float* shift,
                const float* scale) {
  ORC_DISPATCH(h, mlp_set(nl, widths, par, out_scale, shift, scale));
  return 0;
}
// evaluate the MLP on descriptor rows G[natom][D] (by atom id); writes dE/dG and per-atom E
double orc_mlp_eval(OrcHandle* h, int D, const double* G, double* dG, double* e_atom) {
  double etot = 0;
  auto run = [&](auto* e) {
    using TT = typename std::remove_reference<decltype(e->x[0])>::type;
    std::vector<TT> g(G, G + size_t(e->natom) * D), dg;
    std::vector<double> ea;
    etot = e->mlp_eval(g, dg, D, &ea);
    for (size_t k = 0; k < dg.size(); ++k) dG[k] = double(dg[k]);
    if (e_atom) std::copy(ea.begin(), ea.end(), e_atom);
  };
  if (h->prec) run(static_cast<Eng<double>*>(h->e)); else run(static_cast<Eng<float>*>(h->e));
  return etot;
}

int orc_sf_set(OrcHandle* h, float rc, int nrad, const float* mu, float eta_r, int neta, const float* eta) {
  if (h->prec) static_cast<Eng<double>*>(h->e)->sf_set(double(rc), nrad, mu, eta_r, neta, eta);
  else static_cast<Eng<float>*>(h->e)->sf_set(rc, nrad, mu, eta_r, neta, eta);
  return ORC_DISPATCH(h, sf_D);
}
int orc_sf_descriptor(OrcHandle* h, double* out) {
  auto run = [&](auto* e) { e->sf_descriptor(); for (size_t k = 0; k < e->sf_G.size(); ++k) out[k] = double(e->sf_G[k]); };
  if (h->prec) run(static_cast<Eng<double>*>(h->e)); else run(static_cast<Eng<float>*>(h->e));
  return 0;
}
int orc_sf_force(OrcHandle* h, const double* dG, int ev, double* vir) {
  auto run = [&](auto* e) {
    using TT = typename std::remove_reference<decltype(e->x[0])>::type;
    e->sf_dG.assign(dG, dG + size_t(e->natom) * e->sf_D);
    (void)sizeof(TT);
    e->sf_force(ev != 0, vir);
  };
  if (h->prec) run(static_cast<Eng<double>*>(h->e)); else run(static_cast<Eng<float>*>(h->e));
  return 0;
}

// SQD on positions x [natom][3] (any atom order) and elements typ [natom]; qq, qq_ref [nq],
// sq [nq][npair], count [nq]
int orc_sqd(int natom, int ntype, float bl, const float* x, const int* typ, int nq, float q_max, double* qq,
            double* qq_ref, double* sq, long long* count) {
  sqd_restate(natom, ntype, bl, x, typ, nq, q_max, qq, qq_ref, sq, count);
  return 0;
}

}  // extern "C"
