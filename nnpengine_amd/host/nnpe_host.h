// nnpe_host.h -- NNPEngine's host classes over the nnpe C ABI.
//
// The class names and method shapes are the reference's (src/environ.h, src/neighlist.h,
// src/pair.h, src/pair_nnp.h, src/bond.h, src/angle.h, src/dihedral.h, src/grsq.h,
// src/verlet.h) so a driver written against them (src/run.cpp) keeps its structure; every
// method forwards to the GPU engine through include/nnpe.h.  With one GPU there is one owner
// "thread" (NUM_THREAD = 1): the thread_id arguments are accepted and ignored, barriers,
// singles and critical sections of the reference's driver become no-ops, and all state lives
// in the process-wide context nnpe_host::ctx().
#pragma once
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/nnpe.h"

using numtype = float;

namespace nnpe_host {
nnpe_ctx* ctx();
// END_PROGRAM behaviour of the reference (src/util.cpp:117-124): message, then exit(EXIT_FAILURE)
void check(int rc, const char* where);
[[noreturn]] void fail(const char* file, int line, const char* msg);
// input-file helpers with the reference's conventions (src/util.cpp:126-211): skip blank lines,
// expand ${VAR}
bool next_line(FILE* fp, char* buf, int n);
void expand_env(char* buf, int n);
}  // namespace nnpe_host

#define NNPE_END_PROGRAM(msg) nnpe_host::fail(__FILE__, __LINE__, msg)

class MOLECULE {
 public:
  int n_atom = 0, n_bond = 0, n_angle = 0, n_dihedral = 0;
  std::vector<int> bond, angle, dihedral;  // 0-based (type, atoms...) rows of 3 / 4 / 5
  std::vector<int> special;                // [n_atom][n_atom]
  static constexpr numtype special_lj[4] = {1, 0, 0, 0.5}, special_coul[4] = {1, 0, 0, 5 / 6.};
  void from_file(const char* fname);  // src/molecule.cpp:25-131
};

class ENVIRON {
 public:
  static constexpr int NUM_THREAD = 1;
  static int natom, ntype, npair, nmol, nmoltype;
  static numtype bl;
  static std::vector<MOLECULE> moltype;
  static std::vector<numtype> x, q, mass;  // atom-id order, as read
  static std::vector<int> typ, mol_id, intra_id, mol_type, typecount;
  static numtype comm_cutoff, grid_length;
  static void initialize(const char* input_file);  // src/environ.cpp:22-141
  static void init_grid(numtype grid_length, numtype ghost_cutoff);
  static void initV(numtype tempera_rel_300, numtype fraction = 2);  // src/environ.cpp:190-205
  static void dump(const char* xyzfile, bool append);                 // src/environ.cpp:178-188
};

class LOCAL {
 public:
  explicit LOCAL(int) {}
  static void set_ghost_lim(numtype) {}
};

class NEIGHLIST {
 public:
  unsigned nbuild = 0;
  NEIGHLIST(int thr, numtype cutoff, numtype skin, int cap, int** bin_l);
  template <bool full> int build();  // full lists always (one owner)
  bool isvalid();
  static int** build_bin_list(numtype cutoff, int max_nbin, numtype cutoff2 = 0, numtype cutoff3 = 0);
  static numtype short_cutoff;  // optional NNP list (0 = none)
};

class LJTableCoulDSF {
 public:
  int nbin_r;
  numtype cutoff, dr;
  std::vector<numtype> force, potential;  // [npair][nbin+1]
  bool refined = false;                   // an RMDF refines this table on the device
  LJTableCoulDSF(numtype cf, int nbin);
  void loadPotential(int kind, const char* potfile = nullptr);  // src/pair.cpp:197-263
  template <bool ev, bool nsp> void compute(NEIGHLIST* list, numtype* erg, numtype* virial);
  void dumpForce(const char* fname, bool nsp, bool append = false);
};

class BOND {
 public:
  int nbtyp, nbonds;
  std::vector<numtype> coef;  // [nbtyp][4] (b0, 2a, b, 2c)
  BOND(int n, int ns);
  void setParam(int i, numtype b0, numtype ka, numtype kb, numtype kc);
  void setParam();
  void autoGenerate();
  template <bool ev> void compute(int thread_id, numtype* erg, numtype* virial);
};

class ANGLE {
 public:
  int natyp, nangles;
  std::vector<numtype> coef;  // [natyp][4] (cos t0, 2a, b, 2c)
  ANGLE(int n, int ns);
  void setParam(int i, numtype theta_0, numtype ka, numtype kb, numtype kc);
  void setParam();
  void autoGenerate();
  template <bool ev> void compute(int thread_id, numtype* erg, numtype* virial);
};

class DIHEDRAL {
 public:
  int ndtyp, ndihedrals;
  std::vector<numtype> coef;  // [ndtyp][4] cubic in cos
  DIHEDRAL(int n, int ns);
  void setParam(int i, numtype c0 = 0, numtype c1 = 0, numtype c2 = 0, numtype c3 = 0);
  void setParam();
  void autoGenerate();
  template <bool ev> void compute(int thread_id, numtype* erg, numtype* virial);
};

class NNPOTENTIAL_TBL {
 public:
  static constexpr int n_descriptor = 20;
  int nbin_r;
  numtype cutoff;
  // paramfile: first token = network file (replaces the TF SavedModel path); wtfile: the
  // reference's binary weight file (bond/angle/dihedral coefs, coef_p, coef_w, coef_d, coef_r)
  NNPOTENTIAL_TBL(numtype cf, int nbr, const char* paramfile, const char* wtfile, BOND* bnd, ANGLE* agl, DIHEDRAL* dih);
  void computeDescriptor(NEIGHLIST* list);
  void nnEval();
  void evalFinalize(LOCAL*) {}
  template <bool ev> void compute(NEIGHLIST* list, numtype* erg, numtype* virial);
  void allocateTable() {}
  numtype last_nn_energy = 0;
};

class VERLET {
 public:
  explicit VERLET(int) {}
  static void setDt(numtype ddt);
  void updateX();
  void updateV();
  void scaleV(numtype sc);
  numtype computeKE();
  static numtype dt;
};

class GR {
 public:
  int nbin_r, nbin_r_inner;
  numtype dr;
  std::vector<double> gr;  // [npair][nbin], h(r) = g(r) - 1
  GR(int nbi, numtype r_max = -1, numtype r_max_inner = -1);
  template <bool inner> void tally(NEIGHLIST* list);
  void tally(int** bin_list, LOCAL* local);
  void reduce_local(int thread_id);
  void dump(const char* fname, bool append = false);  // src/grsq.cpp:221-256
};

class SQ {
 public:
  int nbin_q;
  numtype q_max;
  std::vector<double> sq;  // [npair][nq]
  GR* gr;
  SQ(int nq, numtype qmax, GR* ggr);
  void compute(int thread_id);
  void dump(const char* fname, bool append = false);  // src/grsq.cpp:324-359
};

// direct k-space S(q) (src/grsq.h:40-63): the half space of reciprocal vectors up to q_max,
// binned by |G|; sq [nq][npair]
class SQD {
 public:
  numtype q_max;
  int nbin_q;
  long long ngv = 0;
  std::vector<double> qq, sq;
  std::vector<long long> count;
  SQD(int nq, numtype qmax);
  void compute(int thread_id);
  void dump(const char* fname, bool append = false);  // src/grsq.cpp:608-644
};

// S(q) refinement of the pair tables (src/grsq.h:65-112).  The experimental data files hold
// nq rows (q, S_exp, weight, one form factor per element) on the SQ grid; pair != nullptr
// refines the LJTableCoulDSF table, else the NNPOTENTIAL_TBL pair tables.
class RMDF {
 public:
  int n_sq, nbin_q;
  numtype gamma, q_max;
  std::vector<float> sq, deltasq, deltasq_accu;  // [n_sq][nq], host copies for dumpSQ
  RMDF(int nexp, char** fexp, SQ* c_sq, numtype strength, numtype gamma, LJTableCoulDSF* pair, NNPOTENTIAL_TBL* pair_nnp);
  void loadDeltaSQ(const char* fname);                   // src/grsq.cpp:1419-1443
  // compute_Qspace + compute_Rspace[_NNP] (+ compute_Potential[_NNP] when potential) on the
  // current S(q) (src/grsq.cpp:1079-1377); returns the Q-space energy
  numtype step(bool potential);
  void dumpSQ(const char* fname, bool append = false);  // src/grsq.cpp:1378-1415
};
