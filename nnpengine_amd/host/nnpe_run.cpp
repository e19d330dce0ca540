// nnpe_run.cpp -- the reference's `classical` and `nnp_run` drivers (src/run.cpp:53-457,
// :968-1403) over the GPU engine: same positional input file, same step order, same stdout log.
//
//   nnpe_run classical in.file       nnpe_run nnp_run in.file
//
// One owner thread: the reference's omp barriers/singles/criticals disappear; every class method
// is one stream-ordered call into libnnpe.so.  Differences from the reference (all explicit):
//  * the time step comes from the environment (NNPE_DT, default 1 fs as the reference);
//  * nnp_run evaluates our per-element MLP (network file = first token of the NNP param file)
//    instead of a TensorFlow SavedModel;
//  * NNPE_SEED fixes the Andersen thermostat's random stream (the reference seeds from
//    std::random_device).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iomanip>
#include <iostream>

#include "nnpe_host.h"

namespace {

struct Input {
  FILE* fp;
  char buf[512];
  explicit Input(const char* f) : fp(std::fopen(f, "r")) {
    if (!fp) NNPE_END_PROGRAM("cannot open input file");
  }
  ~Input() { std::fclose(fp); }
  char* line() {
    if (!nnpe_host::next_line(fp, buf, sizeof buf)) NNPE_END_PROGRAM("read error");
    nnpe_host::expand_env(buf, sizeof buf);
    return buf;
  }
};

void log_header() {
  std::cerr << std::setw(7) << "md_step" << std::setw(18) << "temp [K]" << std::setw(18) << "press [GPa]" << std::setw(18)
            << "erg [kcal/mol]" << std::setw(18) << "dif [kcal/mol]" << std::setw(10) << "nbuild" << "\n";
}

void log_row(int md_step, numtype ekin, numtype virial, numtype erg, numtype erg_exp, unsigned nbuild) {
  const double ek_T = 335.47968899917714, ev_GPa = 6.9476954570553735;
  const numtype temperature = ekin / ENVIRON::natom * ek_T;
  const numtype press = (virial + ekin * 2) / 3 * ev_GPa / std::pow(ENVIRON::bl, 3);
  std::cout << std::setw(7) << md_step << std::setw(18) << temperature << std::setw(18) << press << std::setw(18) << erg
            << std::setw(18) << erg_exp << std::setw(10) << nbuild << std::endl;
}

// the RMDF lines of the input (src/run.cpp:155-170): "strength gamma", "n_sq file...", the
// initial weights file (RMDF::loadDeltaSQ)
struct RmdfInput {
  float strength = 0, gamma = 1;
  int nexp = 0;
  char files[2][256];
  char* fptr[2] = {files[0], files[1]};
  char delta[256];
  void read(Input& in) {
    if (std::sscanf(in.line(), "%f %f", &strength, &gamma) != 2) NNPE_END_PROGRAM("read error");
    const char* buf = in.line();
    int used = 0, k = 0;
    if (std::sscanf(buf, "%d %n", &nexp, &used) != 1) NNPE_END_PROGRAM("read error");
    if (nexp > 2) NNPE_END_PROGRAM("max n_sq is 2");
    for (int i = 0; i < nexp; ++i) {
      if (std::sscanf(buf + used, "%255s %n", files[i], &k) != 1) NNPE_END_PROGRAM("read error");
      used += k;
    }
    if (std::sscanf(in.line(), "%255s", delta) != 1) NNPE_END_PROGRAM("read error");
  }
};

void dumps(Input& in, GR& gr, SQ& sq, RMDF& rmdf, LJTableCoulDSF* pair) {
  char name[256];
  int app, nq;
  float qmax;
  if (std::sscanf(in.line(), "%255s %d", name, &app) == 2 && *name != '*') ENVIRON::dump(name, app);
  if (std::sscanf(in.line(), "%255s %d", name, &app) == 2 && *name != '*') gr.dump(name, app);
  if (std::sscanf(in.line(), "%255s %d", name, &app) == 2 && *name != '*') sq.dump(name, app);
  // sqd: direct k-space S(q) of the final positions (src/run.cpp:421-432)
  if (std::sscanf(in.line(), "%255s %d %f %d", name, &app, &qmax, &nq) != 4) NNPE_END_PROGRAM("read error");
  if (*name != '*') {
    std::cerr << "sqd = " << name << "; append = " << bool(app) << "\n";
    SQD sqd(nq, qmax);
    sqd.compute(0);
    sqd.dump(name, app);
  }
  // sqx: the RMDF S(q) and residuals (src/run.cpp:434-440)
  if (std::sscanf(in.line(), "%255s %d", name, &app) == 2 && *name != '*') rmdf.dumpSQ(name, app);
  if (pair && std::sscanf(in.line(), "%255s %d", name, &app) == 2 && *name != '*') pair->dumpForce(name, false, app);
}

struct Run {
  int total_steps = 0, log_interval = 1, scale_interval = 1, gr_interval = -1, gr_interval_long = -1;
  numtype temp_start = 1, temp_end = 1;
  void read(Input& in) {
    if (std::sscanf(in.line(), "%f %f", &temp_start, &temp_end) != 2) NNPE_END_PROGRAM("read error");
    temp_start /= 300; temp_end /= 300;
    if (std::sscanf(in.line(), "%d %d %d %d %d", &total_steps, &log_interval, &scale_interval, &gr_interval,
                    &gr_interval_long) < 3)
      NNPE_END_PROGRAM("read error");
    gr_interval = gr_interval < 0 ? total_steps : gr_interval;
    gr_interval_long = gr_interval_long < 0 ? total_steps : gr_interval_long;
    if ((total_steps % log_interval) || (total_steps % scale_interval)) NNPE_END_PROGRAM("bad steps");
  }
};

numtype env_dt() {
  const char* s = std::getenv("NNPE_DT");
  return s ? numtype(std::atof(s)) : numtype(1);
}

// classical (src/run.cpp:53-457)
int classical(const char* fin) {
  Input in(fin);
  char name[256];
  std::sscanf(in.line(), "%255s", name);
  ENVIRON::initialize(name);
  numtype pair_cutoff = 10, skin = 1, gr_cutoff = 19;
  if (std::sscanf(in.line(), "%f %f %f", &pair_cutoff, &skin, &gr_cutoff) != 3) NNPE_END_PROGRAM("read error");
  const numtype comm_cutoff = skin + std::max(gr_cutoff, pair_cutoff);
  LOCAL::set_ghost_lim(comm_cutoff);
  ENVIRON::init_grid(0.5 * (skin + pair_cutoff), comm_cutoff);
  int nbin;
  if (std::sscanf(in.line(), "%d %255s", &nbin, name) != 2) NNPE_END_PROGRAM("read error");
  NEIGHLIST list(0, pair_cutoff, skin, 2048, nullptr);
  LJTableCoulDSF pair(pair_cutoff, nbin);
  pair.loadPotential(0, name);
  int nb = 0, na = 0, nd = 0, nt;
  for (int m = 0; m < ENVIRON::nmol; ++m) {
    const MOLECULE& mt = ENVIRON::moltype[ENVIRON::mol_type[m]];
    nb += mt.n_bond; na += mt.n_angle; nd += mt.n_dihedral;
  }
  float a, b, c, d;
  std::sscanf(in.line(), "%d", &nt);
  BOND bond(nt, nb);
  for (int i = 0; i < nt; ++i) { std::sscanf(in.line(), "%f %f", &a, &b); bond.setParam(i, a, 0, 0, b); }
  bond.autoGenerate();
  std::sscanf(in.line(), "%d", &nt);
  ANGLE angle(nt, na);
  for (int i = 0; i < nt; ++i) { std::sscanf(in.line(), "%f %f %f %f", &a, &b, &c, &d); angle.setParam(i, a, b, c, d); }
  angle.autoGenerate();
  std::sscanf(in.line(), "%d", &nt);
  DIHEDRAL dihedral(nt, nd);
  for (int i = 0; i < nt; ++i) { std::sscanf(in.line(), "%f %f %f %f", &a, &b, &c, &d); dihedral.setParam(i, a, b, c, d); }
  dihedral.autoGenerate();
  int nbin_gr, nq;
  float qmax;
  if (std::sscanf(in.line(), "%d %d %f", &nbin_gr, &nq, &qmax) != 3) NNPE_END_PROGRAM("read error");
  GR gr(nbin_gr, gr_cutoff, pair_cutoff);
  SQ sq(nq, qmax, &gr);
  RmdfInput ri;
  ri.read(in);
  RMDF rmdf(ri.nexp, ri.fptr, &sq, ri.strength, ri.gamma, &pair, nullptr);
  rmdf.loadDeltaSQ(ri.delta);
  Run R;
  R.read(in);
  ENVIRON::initV(R.temp_start);
  VERLET verlet(0);
  VERLET::setDt(env_dt());
  log_header();
  numtype erg = 0, virial = 0, ekin = 0, erg_exp = 0;
  for (int md_step = 0; md_step <= R.total_steps; ++md_step) {
    const bool f_log = !(md_step % R.log_interval);
    if (md_step > 0) verlet.updateX();
    if (md_step == 0 || !list.isvalid()) list.build<false>();
    if (!(md_step % R.gr_interval)) {
      if (md_step % R.gr_interval_long) gr.tally<true>(&list);
      else gr.tally(nullptr, nullptr);
      gr.reduce_local(0);
      sq.compute(0);
      erg_exp = rmdf.step(f_log && ri.gamma != 1);  // compute_Qspace, compute_Rspace (+ compute_Potential)
    }
    erg = virial = 0;
    if (f_log) {
      pair.compute<true, false>(&list, &erg, &virial);  // the pair kernel writes the forces ...
      bond.compute<true>(0, &erg, &virial);            // ... and the bonded terms add to them
    } else {
      pair.compute<false, false>(&list, nullptr, nullptr);
      bond.compute<false>(0, nullptr, nullptr);
    }
    verlet.updateV();
    if (!(md_step % R.scale_interval)) {
      ekin = verlet.computeKE();
      ENVIRON::initV(R.temp_start + (R.temp_end - R.temp_start) * md_step / R.total_steps, 0.1);  // THERMO_ANDERSEN_FRAC
    }
    if (f_log) log_row(md_step, ekin, virial, erg, erg_exp, list.nbuild);
  }
  dumps(in, gr, sq, rmdf, &pair);
  return 0;
}

// nnp_run (src/run.cpp:968-1403)
int nnp_run(const char* fin) {
  Input in(fin);
  char name[256], name2[256];
  std::sscanf(in.line(), "%255s", name);
  ENVIRON::initialize(name);
  numtype pair_cutoff = 8, skin = 1, gr_cutoff = 9;
  if (std::sscanf(in.line(), "%f %f %f", &pair_cutoff, &skin, &gr_cutoff) != 3) NNPE_END_PROGRAM("read error");
  const numtype comm_cutoff = skin + std::max(gr_cutoff, pair_cutoff);
  ENVIRON::init_grid(0.5 * (skin + pair_cutoff), comm_cutoff);
  NEIGHLIST list(0, pair_cutoff, skin, 2048, nullptr);
  int nb = 0, na = 0, nd = 0, nt;
  for (int m = 0; m < ENVIRON::nmol; ++m) {
    const MOLECULE& mt = ENVIRON::moltype[ENVIRON::mol_type[m]];
    nb += mt.n_bond; na += mt.n_angle; nd += mt.n_dihedral;
  }
  std::sscanf(in.line(), "%d", &nt);
  BOND bond(nt, nb);
  std::sscanf(in.line(), "%d", &nt);
  ANGLE angle(nt, na);
  std::sscanf(in.line(), "%d", &nt);
  DIHEDRAL dihedral(nt, nd);
  std::sscanf(in.line(), "%255s", name);
  std::sscanf(in.line(), "%255s", name2);
  int nbin;
  std::sscanf(in.line(), "%d", &nbin);
  NNPOTENTIAL_TBL nnp(pair_cutoff, nbin, name, name2, &bond, &angle, &dihedral);
  int nbin_gr, nq;
  float qmax;
  if (std::sscanf(in.line(), "%d %d %f", &nbin_gr, &nq, &qmax) != 3) NNPE_END_PROGRAM("read error");
  GR gr(nbin_gr, gr_cutoff, pair_cutoff);
  SQ sq(nq, qmax, &gr);
  RmdfInput ri;
  ri.read(in);
  RMDF rmdf(ri.nexp, ri.fptr, &sq, ri.strength, ri.gamma, nullptr, &nnp);
  rmdf.loadDeltaSQ(ri.delta);
  Run R;
  R.read(in);
  ENVIRON::initV(R.temp_start);
  VERLET verlet(0);
  VERLET::setDt(env_dt());
  log_header();
  numtype erg = 0, virial = 0, ekin = 0, erg_exp = 0;
  for (int md_step = 0; md_step <= R.total_steps; ++md_step) {
    const bool f_log = !(md_step % R.log_interval);
    if (md_step > 0) verlet.updateX();
    if (md_step == 0 || !list.isvalid()) list.build<false>();
    if (!(md_step % R.gr_interval)) {
      if (md_step % R.gr_interval_long) gr.tally<true>(&list);
      else gr.tally(nullptr, nullptr);
      gr.reduce_local(0);
      sq.compute(0);
      erg_exp = rmdf.step(f_log && ri.gamma != 1);  // compute_Qspace, compute_Rspace_NNP (+ compute_Potential_NNP)
    }
    // forces: the NNP force kernel and the bonded terms add to a cleared force array
    erg = virial = 0;
    nnp.computeDescriptor(&list);
    nnp.nnEval();
    nnpe_host::check(nnpe_clear_force(nnpe_host::ctx()), "clear force");
    if (f_log) {
      nnp.compute<true>(&list, &erg, &virial);
      bond.compute<true>(0, &erg, &virial);
      erg += nnp.last_nn_energy;
    } else {
      nnp.compute<false>(&list, nullptr, nullptr);
      bond.compute<false>(0, nullptr, nullptr);
    }
    verlet.updateV();
    if (!(md_step % R.scale_interval)) {
      ekin = verlet.computeKE();
      ENVIRON::initV(R.temp_start + (R.temp_end - R.temp_start) * md_step / R.total_steps, 0.1);
    }
    if (f_log) log_row(md_step, ekin, virial, erg, erg_exp, list.nbuild);
  }
  dumps(in, gr, sq, rmdf, nullptr);
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  std::cout << std::scientific;
  std::cerr << std::scientific;
  if (argc != 3) NNPE_END_PROGRAM("usage: nnpe_run classical|nnp_run input_file");
  if (!std::strcmp(argv[1], "classical")) return classical(argv[2]);
  if (!std::strcmp(argv[1], "nnp_run")) return nnp_run(argv[2]);
  NNPE_END_PROGRAM("unknown arg");
}
