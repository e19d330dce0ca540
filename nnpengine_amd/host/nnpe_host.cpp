// nnpe_host.cpp -- host classes over the nnpe C ABI (see nnpe_host.h).
#include "nnpe_host.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <random>

namespace nnpe_host {

static nnpe_ctx* g_ctx = nullptr;

nnpe_ctx* ctx() {
  if (!g_ctx) check(nnpe_create(&g_ctx, 0), "nnpe_create");
  return g_ctx;
}

void fail(const char* file, int line, const char* msg) {
  std::fprintf(stderr, "%s:%d -> %s\n", file, line, msg);
  std::exit(EXIT_FAILURE);
}

void check(int rc, const char* where) {
  if (rc == NNPE_OK) return;
  std::fprintf(stderr, "%s -> %s: %s\n", where, nnpe_status_string(rc), g_ctx ? nnpe_last_error(g_ctx) : "");
  std::exit(EXIT_FAILURE);
}

static bool blank(const char* s) {
  for (; *s; ++s)
    if (*s != ' ' && *s != '\n' && *s != '\r' && *s != '\t') return false;
  return true;
}

bool next_line(FILE* fp, char* buf, int n) {
  while (std::fgets(buf, n, fp))
    if (!blank(buf)) return true;
  return false;
}

void expand_env(char* buf, int n) {
  std::string out;
  for (const char* p = buf; *p;) {
    if (p[0] == '$' && p[1] == '{') {
      const char* e = std::strchr(p, '}');
      if (!e) NNPE_END_PROGRAM("getenv failed");
      std::string name(p + 2, e);
      const char* v = std::getenv(name.c_str());
      if (!v) NNPE_END_PROGRAM("getenv failed");
      out += v;
      p = e + 1;
    } else {
      out += *p++;
    }
  }
  if (int(out.size()) >= n) NNPE_END_PROGRAM("string overflow");
  std::memcpy(buf, out.c_str(), out.size() + 1);
}

static std::mt19937& rng() {
  static std::mt19937 g([] {
    const char* s = std::getenv("NNPE_SEED");
    return s ? unsigned(std::strtoul(s, nullptr, 10)) : std::random_device{}();
  }());
  return g;
}

}  // namespace nnpe_host

using nnpe_host::check;

// ---- MOLECULE (src/molecule.cpp:25-131) --------------------------------------------------------
void MOLECULE::from_file(const char* fname) {
  FILE* fp = std::fopen(fname, "r");
  if (!fp) NNPE_END_PROGRAM("read error");
  enum { HEADER, BONDS, ANGLES, DIHEDRALS, OTHER } st = HEADER;
  char buf[512], word[256];
  while (std::fgets(buf, sizeof buf, fp)) {
    const char* s = buf;
    while (*s == ' ' || *s == '\t') ++s;
    if (!*s || *s == '\n' || *s == '\r') continue;
    if (!std::strncmp(s, "Coords", 6) || !std::strncmp(s, "Types", 5) || !std::strncmp(s, "Charges", 7)) { st = OTHER; continue; }
    if (!std::strncmp(s, "Bonds", 5)) { st = BONDS; continue; }
    if (!std::strncmp(s, "Angles", 6)) { st = ANGLES; continue; }
    if (!std::strncmp(s, "Dihedrals", 9)) { st = DIHEDRALS; continue; }
    int a[6];
    switch (st) {
      case HEADER: {
        int v;
        if (std::sscanf(s, "%d %255s", &v, word) != 2) NNPE_END_PROGRAM("read error");
        if (!std::strncmp(word, "atoms", 5)) n_atom = v;
        else if (!std::strncmp(word, "bonds", 5)) n_bond = v;
        else if (!std::strncmp(word, "angles", 6)) n_angle = v;
        else if (!std::strncmp(word, "dihedrals", 9)) n_dihedral = v;
        else NNPE_END_PROGRAM("read error");
        break;
      }
      case BONDS:
        if (std::sscanf(s, "%d %d %d %d", a, a + 1, a + 2, a + 3) != 4) NNPE_END_PROGRAM("read error");
        bond.insert(bond.end(), {a[1] - 1, a[2] - 1, a[3] - 1});
        break;
      case ANGLES:
        if (std::sscanf(s, "%d %d %d %d %d", a, a + 1, a + 2, a + 3, a + 4) != 5) NNPE_END_PROGRAM("read error");
        angle.insert(angle.end(), {a[1] - 1, a[2] - 1, a[3] - 1, a[4] - 1});
        break;
      case DIHEDRALS:
        if (std::sscanf(s, "%d %d %d %d %d %d", a, a + 1, a + 2, a + 3, a + 4, a + 5) != 6) NNPE_END_PROGRAM("read error");
        dihedral.insert(dihedral.end(), {a[1] - 1, a[2] - 1, a[3] - 1, a[4] - 1, a[5] - 1});
        break;
      default:
        break;
    }
  }
  std::fclose(fp);
  if (int(bond.size()) != 3 * n_bond || int(angle.size()) != 4 * n_angle || int(dihedral.size()) != 5 * n_dihedral)
    NNPE_END_PROGRAM("read error");
  special.assign(size_t(n_atom) * n_atom, 0);
  for (int b = 0; b < n_bond; ++b) special[bond[3 * b + 1] * n_atom + bond[3 * b + 2]] = special[bond[3 * b + 2] * n_atom + bond[3 * b + 1]] = 1;
  for (int b = 0; b < n_angle; ++b) special[angle[4 * b + 1] * n_atom + angle[4 * b + 3]] = special[angle[4 * b + 3] * n_atom + angle[4 * b + 1]] = 2;
  for (int b = 0; b < n_dihedral; ++b)
    special[dihedral[5 * b + 1] * n_atom + dihedral[5 * b + 4]] = special[dihedral[5 * b + 4] * n_atom + dihedral[5 * b + 1]] = 3;
}

// ---- ENVIRON --------------------------------------------------------------------------------------
int ENVIRON::natom, ENVIRON::ntype, ENVIRON::npair, ENVIRON::nmol, ENVIRON::nmoltype;
numtype ENVIRON::bl, ENVIRON::comm_cutoff, ENVIRON::grid_length;
std::vector<MOLECULE> ENVIRON::moltype;
std::vector<numtype> ENVIRON::x, ENVIRON::q, ENVIRON::mass;
std::vector<int> ENVIRON::typ, ENVIRON::mol_id, ENVIRON::intra_id, ENVIRON::mol_type, ENVIRON::typecount;

void ENVIRON::initialize(const char* input_file) {
  FILE* fp = std::fopen(input_file, "r");
  if (!fp) NNPE_END_PROGRAM("read error");
  char buf[512], name[256];
  if (!nnpe_host::next_line(fp, buf, sizeof buf) || std::sscanf(buf, "%d %d", &nmoltype, &nmol) != 2) NNPE_END_PROGRAM("read error");
  moltype.assign(nmoltype, MOLECULE());
  for (int i = 0; i < nmoltype; ++i) {
    if (!nnpe_host::next_line(fp, buf, sizeof buf) || std::sscanf(buf, "%255s", name) != 1) NNPE_END_PROGRAM("read error");
    moltype[i].from_file(name);
  }
  if (!nnpe_host::next_line(fp, buf, sizeof buf) || std::sscanf(buf, "%d %d %f", &ntype, &natom, &bl) != 3) NNPE_END_PROGRAM("read error");
  x.resize(size_t(natom) * 3); q.resize(natom); mass.resize(natom);
  typ.resize(natom); mol_id.resize(natom); intra_id.resize(natom); mol_type.clear();
  int intra = 0, imol = 0, mt = 0, mt_prev = 0;
  for (int i = 0; i < natom; ++i) {
    if (!nnpe_host::next_line(fp, buf, sizeof buf) ||
        std::sscanf(buf, "%d %d %f %f %f %f %f", &mt, &typ[i], &mass[i], &q[i], &x[3 * i], &x[3 * i + 1], &x[3 * i + 2]) != 7)
      NNPE_END_PROGRAM("read error");
    if (typ[i] < 0 || typ[i] >= ntype || mt < 0 || mt >= nmoltype) NNPE_END_PROGRAM("read error");
    intra_id[i] = intra;
    mol_id[i] = imol;
    if (!intra) mol_type.push_back(mt);
    else if (mt != mt_prev) NNPE_END_PROGRAM("read error");
    if (moltype[mt].n_atom == ++intra) { ++imol; intra = 0; }
    mt_prev = mt;
  }
  std::fclose(fp);
  if (imol != nmol) NNPE_END_PROGRAM("read error");
  npair = ntype * (ntype + 1) / 2;
  typecount.assign(ntype, 0);
  for (int t : typ) ++typecount[t];
  std::vector<int> mt_natom(nmoltype), mt_off(nmoltype), mt_sp;
  for (int t = 0; t < nmoltype; ++t) {
    mt_natom[t] = moltype[t].n_atom;
    mt_off[t] = int(mt_sp.size());
    mt_sp.insert(mt_sp.end(), moltype[t].special.begin(), moltype[t].special.end());
  }
  check(nnpe_system_set(nnpe_host::ctx(), natom, ntype, bl, x.data(), nullptr, q.data(), mass.data(), typ.data(),
                        mol_id.data(), intra_id.data(), nmol, mol_type.data(), nmoltype, mt_natom.data(), mt_off.data(),
                        mt_sp.data()),
        "ENVIRON::initialize");
}

void ENVIRON::init_grid(numtype gl, numtype ghost_cutoff) {
  grid_length = gl;
  comm_cutoff = ghost_cutoff;
}

// Maxwell-Boltzmann velocities for a fraction of the atoms (Andersen), src/environ.cpp:190-205
void ENVIRON::initV(numtype tempera_rel_300, numtype fraction) {
  static const double kT_m_sqrt = 0.0009118367518929329;
  std::vector<double> v(size_t(natom) * 3);
  check(nnpe_state_get(nnpe_host::ctx(), NNPE_V, v.data()), "ENVIRON::initV");
  std::uniform_real_distribution<double> uni(0, 1);
  std::normal_distribution<double> nrm(0, 1);
  const double t = tempera_rel_300 * 300;
  for (int i = 0; i < natom; ++i) {
    if (uni(nnpe_host::rng()) < fraction) {
      const double fc = std::sqrt(t / mass[i]) * kT_m_sqrt;
      for (int d = 0; d < 3; ++d) v[3 * i + d] = nrm(nnpe_host::rng()) * fc;
    }
  }
  check(nnpe_state_set(nnpe_host::ctx(), NNPE_V, v.data()), "ENVIRON::initV");
}

void ENVIRON::dump(const char* xyzfile, bool append) {
  std::vector<double> xx(size_t(natom) * 3);
  check(nnpe_state_get(nnpe_host::ctx(), NNPE_X, xx.data()), "ENVIRON::dump");
  FILE* fp = std::fopen(xyzfile, append ? "a" : "w");
  if (!fp) NNPE_END_PROGRAM("write error");
  std::fprintf(fp, "%d\nx y z w %.15e %.15e\n", natom, -0.5 * bl, 0.5 * bl);
  for (int i = 0; i < natom; ++i) std::fprintf(fp, "T%d %.15e %.15e %.15e\n", typ[i], xx[3 * i], xx[3 * i + 1], xx[3 * i + 2]);
  std::fclose(fp);
}

// ---- NEIGHLIST ------------------------------------------------------------------------------------
numtype NEIGHLIST::short_cutoff = 0;
static numtype g_pair_cutoff = 0, g_skin = 0;

int** NEIGHLIST::build_bin_list(numtype, int, numtype, numtype) { return nullptr; }  // the stencil lives on the device

NEIGHLIST::NEIGHLIST(int, numtype cutoff, numtype skin, int, int**) {
  g_pair_cutoff = cutoff;
  g_skin = skin;
  const numtype gr_cutoff = ENVIRON::comm_cutoff - skin;
  check(nnpe_neigh_init(nnpe_host::ctx(), cutoff, skin, gr_cutoff, short_cutoff), "NEIGHLIST");
}

template <bool full> int NEIGHLIST::build() {
  check(nnpe_neigh_build(nnpe_host::ctx()), "NEIGHLIST::build");
  ++nbuild;
  long long n = 0;
  check(nnpe_neigh_size(nnpe_host::ctx(), 0, &n, nullptr), "NEIGHLIST::build");
  return int(n);
}
template int NEIGHLIST::build<true>();
template int NEIGHLIST::build<false>();

bool NEIGHLIST::isvalid() {
  int v = 0;
  check(nnpe_neigh_isvalid(nnpe_host::ctx(), &v), "NEIGHLIST::isvalid");
  return v != 0;
}

// ---- energies: the ev compute calls accumulate per term in FP64 on the device -------------------
static void take_energy(int term, numtype* erg, numtype* virial) {
  double ev[2 * NNPE_NTERM];
  check(nnpe_energy_get(nnpe_host::ctx(), ev), "energy");
  if (erg) *erg += numtype(ev[2 * term]);
  if (virial) *virial += numtype(ev[2 * term + 1]);
}
static void clear_energy() { check(nnpe_energy_clear(nnpe_host::ctx()), "energy"); }

// ---- LJTableCoulDSF ---------------------------------------------------------------------------------
LJTableCoulDSF::LJTableCoulDSF(numtype cf, int nbin) : nbin_r(nbin), cutoff(cf), dr(cf / nbin) {
  force.assign(size_t(ENVIRON::npair) * (nbin + 1), 0);
  potential.assign(force.size(), 0);
}

void LJTableCoulDSF::loadPotential(int kind, const char* potfile) {
  if (kind == 2) NNPE_END_PROGRAM("special-ignoring tables (RMDF) are not supported");
  if (potfile) {
    FILE* fp = std::fopen(potfile, "r");
    if (!fp) NNPE_END_PROGRAM("read error");
    for (int i = 0; i < nbin_r; ++i) {
      float r;
      if (std::fscanf(fp, "%f ", &r) != 1) NNPE_END_PROGRAM("read error");
      if (std::fabs(r - (i + .5F) * dr) > 1e-5F) std::fprintf(stderr, "warning: large diff: %g %g\n", r, (i + .5F) * dr);
      for (int p = 0; p < ENVIRON::npair; ++p)
        if (std::fscanf(fp, "%f %f ", &force[size_t(p) * (nbin_r + 1) + i], &potential[size_t(p) * (nbin_r + 1) + i]) != 2)
          NNPE_END_PROGRAM("read error");
    }
    std::fclose(fp);
  }
  check(nnpe_pair_table(nnpe_host::ctx(), cutoff, nbin_r, force.data(), potential.data()), "LJTableCoulDSF::loadPotential");
}

template <bool ev, bool nsp> void LJTableCoulDSF::compute(NEIGHLIST*, numtype* erg, numtype* virial) {
  if (ev) clear_energy();
  check(nnpe_pair_compute(nnpe_host::ctx(), ev), "LJTableCoulDSF::compute");
  if (ev) take_energy(NNPE_TERM_PAIR, erg, virial);
}
template void LJTableCoulDSF::compute<true, false>(NEIGHLIST*, numtype*, numtype*);
template void LJTableCoulDSF::compute<false, false>(NEIGHLIST*, numtype*, numtype*);

void LJTableCoulDSF::dumpForce(const char* fname, bool, bool append) {
  // the refined table as it is now on the device (RMDF rewrites it in place, src/grsq.cpp:1246-1300)
  if (refined) check(nnpe_rmdf_tables(nnpe_host::ctx(), force.data(), potential.data()), "LJTableCoulDSF::dumpForce");
  FILE* fp = std::fopen(fname, append ? "a+" : "w");
  if (!fp) NNPE_END_PROGRAM("write error");
  for (int i = 0; i < nbin_r; ++i) {
    std::fprintf(fp, "%.15e", (i + .5F) * dr);
    for (int p = 0; p < ENVIRON::npair; ++p)
      std::fprintf(fp, " %.15e %.15e", force[size_t(p) * (nbin_r + 1) + i], potential[size_t(p) * (nbin_r + 1) + i]);
    std::fprintf(fp, "\n");
  }
  std::fclose(fp);
}

// ---- bonded: autoGenerate (src/bond.cpp:45-82) expands the molecule templates -----------------
static std::vector<int> expand_terms(int kind) {
  std::vector<int> out;
  const int w = 3 + kind;
  for (int i = 0; i < ENVIRON::natom; ++i) {
    if (ENVIRON::intra_id[i]) continue;
    const MOLECULE& m = ENVIRON::moltype[ENVIRON::mol_type[ENVIRON::mol_id[i]]];
    const std::vector<int>& t = kind == 0 ? m.bond : kind == 1 ? m.angle : m.dihedral;
    for (size_t k = 0; k < t.size(); k += w) {
      out.push_back(t[k]);
      for (int a = 1; a < w; ++a) out.push_back(t[k + a] + i);
    }
  }
  return out;
}

BOND::BOND(int n, int ns) : nbtyp(n), nbonds(ns), coef(size_t(n) * 4, 0) {}
void BOND::setParam(int i, numtype b0, numtype ka, numtype kb, numtype kc) {
  if (i >= nbtyp) NNPE_END_PROGRAM("bad i");
  coef[4 * i] = b0; coef[4 * i + 1] = ka * 2; coef[4 * i + 2] = kb; coef[4 * i + 3] = kc * 2;
}
void BOND::setParam() { for (int i = 0; i < nbtyp; ++i) { coef[4 * i + 1] *= 2; coef[4 * i + 3] *= 2; } }
void BOND::autoGenerate() {
  std::vector<int> t = expand_terms(0);
  if (int(t.size()) != 3 * nbonds) NNPE_END_PROGRAM("auto generate error");
  check(nnpe_bonded_set(nnpe_host::ctx(), 0, nbtyp, coef.data(), nbonds, t.data()), "BOND::autoGenerate");
}
// the GPU bonded kernel evaluates bonds, angles and dihedrals together: BOND::compute runs it,
// ANGLE/DIHEDRAL::compute report their energies from the same launch
static double g_bonded_e[3] = {0, 0, 0}, g_bonded_w = 0;
template <bool ev> void BOND::compute(int, numtype* erg, numtype* virial) {
  if (ev) clear_energy();
  check(nnpe_bonded_compute(nnpe_host::ctx(), ev), "BOND::compute");
  if (ev) {
    double e[2 * NNPE_NTERM];
    check(nnpe_energy_get(nnpe_host::ctx(), e), "BOND::compute");
    g_bonded_e[0] = e[2 * NNPE_TERM_BOND]; g_bonded_e[1] = e[2 * NNPE_TERM_ANGLE]; g_bonded_e[2] = e[2 * NNPE_TERM_DIHEDRAL];
    g_bonded_w = e[2 * NNPE_TERM_BOND + 1];
    *erg += numtype(g_bonded_e[0] + g_bonded_e[1] + g_bonded_e[2]);
    *virial += numtype(g_bonded_w);
  }
}
template void BOND::compute<true>(int, numtype*, numtype*);
template void BOND::compute<false>(int, numtype*, numtype*);

ANGLE::ANGLE(int n, int ns) : natyp(n), nangles(ns), coef(size_t(n) * 4, 0) {}
void ANGLE::setParam(int i, numtype theta_0, numtype ka, numtype kb, numtype kc) {
  if (i >= natyp) NNPE_END_PROGRAM("bad i");
  const numtype pi = 3.141592653589793f;
  coef[4 * i] = std::cos(theta_0 * pi / 180); coef[4 * i + 1] = ka * 2; coef[4 * i + 2] = kb; coef[4 * i + 3] = kc * 2;
}
void ANGLE::setParam() { for (int i = 0; i < natyp; ++i) { coef[4 * i + 1] *= 2; coef[4 * i + 3] *= 2; } }
void ANGLE::autoGenerate() {
  std::vector<int> t = expand_terms(1);
  if (int(t.size()) != 4 * nangles) NNPE_END_PROGRAM("auto generate error");
  check(nnpe_bonded_set(nnpe_host::ctx(), 1, natyp, coef.data(), nangles, t.data()), "ANGLE::autoGenerate");
}
template <bool ev> void ANGLE::compute(int, numtype*, numtype*) {}  // evaluated with BOND::compute
template void ANGLE::compute<true>(int, numtype*, numtype*);
template void ANGLE::compute<false>(int, numtype*, numtype*);

DIHEDRAL::DIHEDRAL(int n, int ns) : ndtyp(n), ndihedrals(ns), coef(size_t(n) * 4, 0) {}
void DIHEDRAL::setParam(int i, numtype c0, numtype c1, numtype c2, numtype c3) {
  if (i >= ndtyp) NNPE_END_PROGRAM("bad i");
  coef[4 * i] = c0 - c2; coef[4 * i + 1] = c1 - 3 * c3; coef[4 * i + 2] = c2 * 2 * 2; coef[4 * i + 3] = c3 * 4 * 3;
}
void DIHEDRAL::setParam() { for (int i = 0; i < ndtyp; ++i) { coef[4 * i + 2] *= 2; coef[4 * i + 3] *= 3; } }
void DIHEDRAL::autoGenerate() {
  std::vector<int> t = expand_terms(2);
  if (int(t.size()) != 5 * ndihedrals) NNPE_END_PROGRAM("auto generate error");
  if (ndihedrals) check(nnpe_bonded_set(nnpe_host::ctx(), 2, ndtyp, coef.data(), ndihedrals, t.data()), "DIHEDRAL::autoGenerate");
}
template <bool ev> void DIHEDRAL::compute(int, numtype*, numtype*) {}  // evaluated with BOND::compute
template void DIHEDRAL::compute<true>(int, numtype*, numtype*);
template void DIHEDRAL::compute<false>(int, numtype*, numtype*);

// ---- NNPOTENTIAL_TBL (src/pair_nnp.cpp:433-793) ------------------------------------------------------
// Network file (first token of the param file): text
//   nlayer w0 w1 ... w_nlayer out_scale precision(fp32|bf16|fp16)
//   then per element, per layer: W [in][out] row-major, b [out]
NNPOTENTIAL_TBL::NNPOTENTIAL_TBL(numtype cf, int nbr, const char* paramfile, const char* wtfile, BOND* bnd, ANGLE* agl,
                                 DIHEDRAL* dih)
    : nbin_r(nbr), cutoff(cf) {
  if (ENVIRON::nmoltype != 1) NNPE_END_PROGRAM("NNP does not support multiple molecules");
  char path_nn[512];
  FILE* fp = std::fopen(paramfile, "r");
  if (!fp || std::fscanf(fp, "%511s", path_nn) != 1) NNPE_END_PROGRAM("read error");
  std::fclose(fp);
  // network
  fp = std::fopen(path_nn, "r");
  if (!fp) NNPE_END_PROGRAM("read error");
  int nl = 0;
  if (std::fscanf(fp, "%d", &nl) != 1 || nl < 2) NNPE_END_PROGRAM("read error");
  std::vector<int> widths(nl + 1);
  for (int& w : widths)
    if (std::fscanf(fp, "%d", &w) != 1) NNPE_END_PROGRAM("read error");
  float out_scale;
  char prec[32];
  if (std::fscanf(fp, "%f %31s", &out_scale, prec) != 2) NNPE_END_PROGRAM("read error");
  if (widths[0] != n_descriptor) NNPE_END_PROGRAM("network input width must be 20");
  size_t npar = 0;
  for (int l = 0; l < nl; ++l) npar += size_t(widths[l]) * widths[l + 1] + widths[l + 1];
  std::vector<float> par(npar * ENVIRON::ntype);
  for (float& v : par)
    if (std::fscanf(fp, "%f", &v) != 1) NNPE_END_PROGRAM("read error");
  std::fclose(fp);
  const int p = !std::strcmp(prec, "fp32") ? 0 : !std::strcmp(prec, "bf16") ? 1 : !std::strcmp(prec, "fp16") ? 2 : !std::strcmp(prec, "bf16x1") ? 3 : -1;
  if (p < 0) NNPE_END_PROGRAM("precision must be fp32, bf16, fp16 or bf16x1");
  // weight file: bonded coefficients then the descriptor/pair coefficients (src/pair_nnp.cpp:505-517)
  fp = std::fopen(wtfile, "rb");
  if (!fp) NNPE_END_PROGRAM("read error");
  const int np = ENVIRON::npair;
  auto rd = [&](float* d, size_t n) { if (n && std::fread(d, sizeof(float), n, fp) != n) NNPE_END_PROGRAM("read error"); };
  rd(bnd->coef.data(), size_t(bnd->nbtyp) * 4);
  rd(agl->coef.data(), size_t(agl->natyp) * 4);
  rd(dih->coef.data(), size_t(dih->ndtyp) * 4);
  bnd->setParam(); agl->setParam(); dih->setParam();
  std::vector<float> cp(size_t(np) * 24), cw(size_t(np) * 80), cd(cw.size()), cr(cw.size());
  rd(cp.data(), cp.size()); rd(cw.data(), cw.size()); rd(cd.data(), cd.size()); rd(cr.data(), cr.size());
  std::fclose(fp);
  bnd->autoGenerate(); agl->autoGenerate(); dih->autoGenerate();
  check(nnpe_nnp_tbl_set(nnpe_host::ctx(), cf, nbr, cp.data(), cw.data(), cd.data(), cr.data()), "NNPOTENTIAL_TBL");
  check(nnpe_mlp_set(nnpe_host::ctx(), nl, widths.data(), par.data(), out_scale, p, nullptr, nullptr), "NNPOTENTIAL_TBL");
}

void NNPOTENTIAL_TBL::computeDescriptor(NEIGHLIST*) { check(nnpe_nnp_tbl_descriptor(nnpe_host::ctx()), "computeDescriptor"); }

void NNPOTENTIAL_TBL::nnEval() {
  clear_energy();
  check(nnpe_mlp_eval(nnpe_host::ctx(), 1), "nnEval");
  double e[2 * NNPE_NTERM];
  check(nnpe_energy_get(nnpe_host::ctx(), e), "nnEval");
  last_nn_energy = numtype(e[2 * NNPE_TERM_NN]);
}

template <bool ev> void NNPOTENTIAL_TBL::compute(NEIGHLIST*, numtype* erg, numtype* virial) {
  if (ev) clear_energy();
  check(nnpe_nnp_tbl_force(nnpe_host::ctx(), ev), "NNPOTENTIAL_TBL::compute");
  if (ev) take_energy(NNPE_TERM_NNP_PAIR, erg, virial);
}
template void NNPOTENTIAL_TBL::compute<true>(NEIGHLIST*, numtype*, numtype*);
template void NNPOTENTIAL_TBL::compute<false>(NEIGHLIST*, numtype*, numtype*);

// ---- VERLET (src/verlet.cpp) --------------------------------------------------------------------------
numtype VERLET::dt = 1;
void VERLET::setDt(numtype ddt) {
  dt = ddt;
  check(nnpe_verlet_set(nnpe_host::ctx(), ddt, 0), "VERLET::setDt");
}
void VERLET::updateX() { check(nnpe_verlet_update_x(nnpe_host::ctx()), "VERLET::updateX"); }
void VERLET::updateV() { check(nnpe_verlet_update_v(nnpe_host::ctx()), "VERLET::updateV"); }
void VERLET::scaleV(numtype sc) { check(nnpe_verlet_scale_v(nnpe_host::ctx(), sc), "VERLET::scaleV"); }
numtype VERLET::computeKE() {
  double ke = 0;
  check(nnpe_verlet_compute_ke(nnpe_host::ctx(), &ke), "VERLET::computeKE");
  return numtype(ke);
}

// ---- GR / SQ with the running-average dump header (src/grsq.cpp:221-256, :324-359) --------------------
GR::GR(int nbi, numtype r_max, numtype r_max_inner) : nbin_r_inner(nbi) {
  int nb = 0;
  check(nnpe_gr_init(nnpe_host::ctx(), nbi, r_max, r_max_inner, &nb), "GR");
  nbin_r = nb;
  dr = (r_max_inner > 0 ? r_max_inner : (r_max > 0 ? r_max : ENVIRON::bl / 2)) / nbi;
  gr.assign(size_t(ENVIRON::npair) * nb, 0);
}
template <bool inner> void GR::tally(NEIGHLIST*) { check(nnpe_gr_tally(nnpe_host::ctx(), 0), "GR::tally"); }
template void GR::tally<true>(NEIGHLIST*);
template void GR::tally<false>(NEIGHLIST*);
void GR::tally(int**, LOCAL*) { check(nnpe_gr_tally(nnpe_host::ctx(), 1), "GR::tally"); }
void GR::reduce_local(int) { check(nnpe_gr_get(nnpe_host::ctx(), nullptr, gr.data()), "GR::reduce_local"); }

static void dump_table(const char* fname, bool append, int nrow, int ncol, std::vector<double>& val, double d) {
  int num = 1;
  FILE* fp;
  if (append && (fp = std::fopen(fname, "r"))) {
    if (std::fscanf(fp, "# %d ", &num) != 1) num = 0;
    num = std::min(num + 1, 100);  // MAXACC
    const double factor = 1.0 / num, cofactor = 1 - factor;
    for (int i = 0; i < nrow; ++i) {
      float v;
      if (std::fscanf(fp, "%f ", &v) != 1) break;
      for (int p = 0; p < ncol; ++p) {
        if (std::fscanf(fp, "%f ", &v) != 1) break;
        val[size_t(p) * nrow + i] = val[size_t(p) * nrow + i] * factor + cofactor * v;
      }
    }
    std::fclose(fp);
  }
  fp = std::fopen(fname, "w");
  if (!fp) NNPE_END_PROGRAM("write error");
  std::fprintf(fp, "# %d\n", num);
  for (int i = 0; i < nrow; ++i) {
    std::fprintf(fp, "%.15e", (i + 0.5) * d);
    for (int p = 0; p < ncol; ++p) std::fprintf(fp, " %.15e", val[size_t(p) * nrow + i]);
    std::fprintf(fp, "\n");
  }
  std::fclose(fp);
}

void GR::dump(const char* fname, bool append) { dump_table(fname, append, nbin_r, ENVIRON::npair, gr, dr); }

SQ::SQ(int nq, numtype qmax, GR* ggr) : nbin_q(nq), q_max(qmax), gr(ggr) {
  check(nnpe_sq_init(nnpe_host::ctx(), nq, 0.0, qmax), "SQ");
  sq.assign(size_t(ENVIRON::npair) * nq, 0);
}
void SQ::compute(int) { check(nnpe_sq_compute(nnpe_host::ctx(), sq.data()), "SQ::compute"); }
void SQ::dump(const char* fname, bool append) { dump_table(fname, append, nbin_q, ENVIRON::npair, sq, double(q_max) / nbin_q); }

// ---- SQD (src/grsq.cpp:362-644) ----------------------------------------------------------------------
SQD::SQD(int nq, numtype qmax) : q_max(qmax), nbin_q(nq) {
  check(nnpe_sqd_init(nnpe_host::ctx(), nq, qmax, &ngv), "SQD");
  std::cerr << "SQDIRECT::SQDIRECT " << ngv << " G vecs\n";
  qq.assign(nq, 0);
  sq.assign(size_t(nq) * ENVIRON::npair, 0);
  count.assign(nq, 0);
}
void SQD::compute(int) {
  check(nnpe_sqd_compute(nnpe_host::ctx()), "SQD::compute");
  check(nnpe_sqd_get(nnpe_host::ctx(), qq.data(), sq.data(), count.data()), "SQD::compute");
}
// running average over appended runs (# n header, at most 100); empty bins are not written
void SQD::dump(const char* fname, bool append) {
  int num = 1;
  FILE* fp;
  if (append && (fp = std::fopen(fname, "r"))) {
    if (std::fscanf(fp, "# %d ", &num) != 1) num = 0;
    num = std::min(num + 1, 100);  // MAXACC
    const double factor = 1.0 / num, cofactor = 1 - factor;
    std::cerr << "sq_acc = " << num << "\n";
    for (int i = 0; i < nbin_q; ++i) {
      if (!count[i]) continue;
      float v;
      if (std::fscanf(fp, "%f ", &v) != 1) break;
      if (std::fabs(v - qq[i]) > 1e-5f) std::cerr << "warning: large diff: " << v << " " << qq[i] << "\n";
      for (int p = 0; p < ENVIRON::npair; ++p) {
        if (std::fscanf(fp, "%f ", &v) != 1) break;
        double& s = sq[size_t(i) * ENVIRON::npair + p];
        s = s * factor + cofactor * v;
      }
    }
    std::fclose(fp);
  }
  fp = std::fopen(fname, "w");
  if (!fp) NNPE_END_PROGRAM("write error");
  std::fprintf(fp, "# %d\n", num);
  for (int i = 0; i < nbin_q; ++i) {
    if (!count[i]) continue;
    std::fprintf(fp, "%.15e", qq[i]);
    for (int p = 0; p < ENVIRON::npair; ++p) std::fprintf(fp, " %.15e", sq[size_t(i) * ENVIRON::npair + p]);
    std::fprintf(fp, "\n");
  }
  std::fclose(fp);
}

// ---- RMDF (src/grsq.cpp:899-1443) --------------------------------------------------------------------
RMDF::RMDF(int nexp, char** fexp, SQ* c_sq, numtype strength, numtype _gamma, LJTableCoulDSF* pair, NNPOTENTIAL_TBL*)
    : n_sq(nexp), nbin_q(c_sq->nbin_q), gamma(_gamma), q_max(c_sq->q_max) {
  if (gamma == 1) std::cerr << "DA/RMD; ";
  else if (gamma > 0 && gamma < 1) std::cerr << "FMIRL training; ";
  else if (gamma == 0) std::cerr << "FMIRL inference; ";
  else NNPE_END_PROGRAM("invalid gamma");
  std::cerr << "strength = " << strength << "\n";
  const int nt = ENVIRON::ntype, row = 3 + nt;
  std::vector<float> data(size_t(n_sq) * nbin_q * row, 0.0f);
  const double dq = double(q_max) / nbin_q;
  for (int n = 0; n < n_sq; ++n) {
    std::cerr << "reading " << fexp[n] << "\n";
    FILE* fp = std::fopen(fexp[n], "r");
    if (!fp) NNPE_END_PROGRAM("read error");
    float* d = data.data() + size_t(n) * nbin_q * row;
    for (int i = 0; i < nbin_q; ++i)
      for (int c = 0; c < row; ++c)
        if (std::fscanf(fp, "%f ", d + size_t(i) * row + c) != 1) NNPE_END_PROGRAM("read error");
    for (int i = 0; i < nbin_q; ++i)
      if (std::fabs(d[size_t(i) * row] - (i + 0.5) * dq) > 1e-5)
        std::cerr << "warning: large diff: " << d[size_t(i) * row] << " " << (i + 0.5) * dq << "\n";
    std::fclose(fp);
    std::cerr << "last entry: " << d[size_t(nbin_q - 1) * row] << " " << d[size_t(nbin_q - 1) * row + 1] << " "
              << d[size_t(nbin_q - 1) * row + 2] * strength << "\n";
  }
  sq.assign(size_t(n_sq) * nbin_q, 0.0f);
  deltasq.assign(sq.size(), 0.0f);
  deltasq_accu.assign(sq.size(), 0.0f);
  if (n_sq > 0) {
    check(nnpe_rmdf_init(nnpe_host::ctx(), n_sq, data.data(), strength, gamma, pair ? 0 : 1), "RMDF");
    if (pair) pair->refined = true;
  }
}
// the relaxed residual to resume from: column 2 (gamma != 0) or column 3 (FMIRL inference) of a
// dumpSQ file; a missing file is not an error
void RMDF::loadDeltaSQ(const char* fname) {
  std::cerr << "reading " << fname << "\n";
  FILE* fp = std::fopen(fname, "r");
  if (!fp) {
    std::cerr << "cannot open file\n";
    return;
  }
  if (std::fscanf(fp, "# %*d ") != 0) NNPE_END_PROGRAM("read error");
  const double dq = double(q_max) / nbin_q;
  float qi = 0;
  for (int i = 0; i < nbin_q; ++i) {
    if (std::fscanf(fp, "%f ", &qi) != 1) NNPE_END_PROGRAM("read error");
    if (std::fabs(qi - (i + 0.5) * dq) > 1e-5) std::cerr << "warning: large diff: " << qi << " " << (i + 0.5) * dq << "\n";
    for (int n = 0; n < n_sq; ++n) {
      float a, b, c;
      if (std::fscanf(fp, "%f %f %f ", &a, &b, &c) != 3) NNPE_END_PROGRAM("read error");
      deltasq[size_t(n) * nbin_q + i] = gamma == 0 ? c : b;
    }
  }
  std::fclose(fp);
  if (n_sq > 0) {
    std::cerr << "last entry: " << qi << " " << deltasq.back() << "\n";
    check(nnpe_rmdf_set_delta(nnpe_host::ctx(), deltasq.data()), "RMDF::loadDeltaSQ");
  }
}
numtype RMDF::step(bool potential) {
  if (n_sq == 0) return 0;
  check(nnpe_rmdf_step(nnpe_host::ctx(), potential ? 1 : 0), "RMDF");
  double e[2 * NNPE_NTERM];
  check(nnpe_energy_get(nnpe_host::ctx(), e), "RMDF");
  return numtype(e[2 * NNPE_TERM_RMDF]);
}
// format: q, then per data set S_n(q) (running average), deltasq, deltasq running average
void RMDF::dumpSQ(const char* fname, bool append) {
  if (n_sq > 0) check(nnpe_rmdf_get(nnpe_host::ctx(), sq.data(), deltasq.data(), nullptr), "RMDF::dumpSQ");
  int num = 1;
  double factor = 1, cofactor = 0;
  FILE* fp;
  const double dq = double(q_max) / nbin_q;
  if (append && (fp = std::fopen(fname, "r"))) {
    if (std::fscanf(fp, "# %d ", &num) != 1) num = 0;
    num = std::min(num + 1, 100);  // MAXACC
    cofactor = 1 - (factor = 1.0 / num);
    std::cerr << "rmdf_acc = " << num << "\n";
    for (int i = 0; i < nbin_q; ++i) {
      float v;
      if (std::fscanf(fp, "%f ", &v) != 1) break;
      if (std::fabs(v - (i + 0.5) * dq) > 1e-5) std::cerr << "warning: large diff: " << v << " " << (i + 0.5) * dq << "\n";
      for (int n = 0; n < n_sq; ++n) {
        float a, b, c;
        if (std::fscanf(fp, "%f %f %f ", &a, &b, &c) != 3) break;
        float& s = sq[size_t(n) * nbin_q + i];
        s = float(s * factor + cofactor * a);
        deltasq_accu[size_t(n) * nbin_q + i] = c;
      }
    }
    std::fclose(fp);
  }
  fp = std::fopen(fname, "w");
  if (!fp) NNPE_END_PROGRAM("write error");
  std::fprintf(fp, "# %d\n", num);
  for (int i = 0; i < nbin_q; ++i) {
    std::fprintf(fp, "%.15e", (i + 0.5) * dq);
    for (int n = 0; n < n_sq; ++n) {
      const size_t k = size_t(n) * nbin_q + i;
      std::fprintf(fp, " %.15e %.15e %.15e", double(sq[k]), double(deltasq[k]),
                   deltasq[k] * factor + cofactor * deltasq_accu[k]);
    }
    std::fprintf(fp, "\n");
  }
  std::fclose(fp);
}
