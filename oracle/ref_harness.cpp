// ref_harness.cpp -- drives the reference's own classes, compiled from the reference sources
// (see oracle/Makefile), to produce fixtures and the CPU baseline.
//
// TEST INFRASTRUCTURE ONLY (test reference and the CPU-baseline leg; never the product path).
//
// It follows the classical() MD loop of src/run.cpp:204-394 with the reference's 27 OpenMP
// threads: refold_and_tag -> ENVIRON::sort -> generate_ghost -> tidyup_ghost ->
// update_ghost_pos<true> -> sort_ghost -> update_ghost_and_info -> NEIGHLIST::build<false>,
// then BOND/ANGLE/DIHEDRAL::compute and LJTableCoulDSF::compute, VERLET::updateX/updateV.
// grsq.cpp and pair_nnp.cpp are not linked: they include the TensorFlow C header, which the
// image does not have.
//
// usage: ref_harness <mode> <params-file> <outdir>
//   mode "snap": one list build + one force evaluation on the configuration as read
//   mode "md"  : <nsteps> velocity-Verlet steps (skin-triggered rebuilds), then a snapshot;
//                prints the wall time of the MD loop.
// params-file lines:
//   config_file
//   pair_cutoff skin gr_cutoff
//   nbin table_file
//   nbtyp, then nbtyp lines "b0 k"                 (BOND::setParam(i,b0,0,0,k), run.cpp:115-125)
//   natyp, then natyp lines "theta0 ka kb kc"      (ANGLE::setParam)
//   ndtyp, then ndtyp lines "c0 c1 c2 c3"          (DIHEDRAL::setParam)
//   velocity_file (raw float32 [natom][3] by atom id, or *)
//   nsteps dt

#include <omp.h>

#include <chrono>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "angle.h"
#include "bond.h"
#include "dihedral.h"
#include "environ.h"
#include "neighlist.h"
#include "pair.h"
#include "util.h"
#include "verlet.h"

static void die(const char* m) { fprintf(stderr, "ref_harness: %s\n", m); exit(2); }

int main(int argc, char** argv) {
  if (argc != 4) die("usage: ref_harness <snap|md> <params> <outdir>");
  std::string mode = argv[1], outdir = argv[3];
  FILE* fp = fopen(argv[2], "r");
  if (!fp) die("cannot open params");
  char buf[SZ_FBF], s1[SZ_FBF];
  auto line = [&]() { fgets_non_empty(fp, buf); };
  line(); sscanf(buf, "%s", s1);
  ENVIRON::initialize(s1);
  float pair_cutoff, skin, gr_cutoff;
  line(); if (sscanf(buf, "%f %f %f", &pair_cutoff, &skin, &gr_cutoff) != 3) die("cutoffs");
  const float comm_cutoff = skin + std::max(gr_cutoff, pair_cutoff);
  LOCAL::set_ghost_lim(comm_cutoff);
  ENVIRON::init_grid(0.5 * (skin + pair_cutoff), comm_cutoff);
  int** bin_list_pair = NEIGHLIST::build_bin_list(pair_cutoff + skin, 126);
  int nbin; line(); if (sscanf(buf, "%d %s", &nbin, s1) != 2) die("table");
  LJTableCoulDSF pair(pair_cutoff, nbin);
  pair.loadPotential(0, s1);
  pair.loadPotential(1);

  int nb = 0, na = 0, nd = 0;
  for (int im = 0; im < ENVIRON::nmol; ++im) {
    MOLECULE* m = ENVIRON::moltype + ENVIRON::mol_type[im];
    nb += m->n_bond; na += m->n_angle; nd += m->n_dihedral;
  }
  int nt; float a, b, c, d;
  line(); sscanf(buf, "%d", &nt);
  BOND bond(nt, nb);
  for (int i = 0; i < nt; ++i) { line(); sscanf(buf, "%f %f", &a, &b); bond.setParam(i, a, 0, 0, b); }
  bond.autoGenerate();
  line(); sscanf(buf, "%d", &nt);
  ANGLE angle(nt, na);
  for (int i = 0; i < nt; ++i) { line(); sscanf(buf, "%f %f %f %f", &a, &b, &c, &d); angle.setParam(i, a, b, c, d); }
  angle.autoGenerate();
  line(); sscanf(buf, "%d", &nt);
  DIHEDRAL dihedral(nt, nd);
  for (int i = 0; i < nt; ++i) { line(); sscanf(buf, "%f %f %f %f", &a, &b, &c, &d); dihedral.setParam(i, a, b, c, d); }
  dihedral.autoGenerate();
  line(); sscanf(buf, "%s", s1);
  const int N = ENVIRON::natom;
  if (s1[0] != '*') {
    FILE* fv = fopen(s1, "rb");
    if (!fv) die("velocity file");
    std::vector<float> vv(size_t(N) * 3);
    if (fread(vv.data(), sizeof(float), vv.size(), fv) != vv.size()) die("velocity size");
    fclose(fv);
    for (int i = 0; i < N; ++i) for (int k = 0; k < 3; ++k) ENVIRON::v[i][k] = vv[3 * i + k];
  } else {
    for (int i = 0; i < N; ++i) for (int k = 0; k < 3; ++k) ENVIRON::v[i][k] = 0;
  }
  int nsteps = 0; float dt = 1;
  line(); sscanf(buf, "%d %f", &nsteps, &dt);
  fclose(fp);
  if (mode == "snap") nsteps = 0;

  double erg = 0, virial = 0, ekin = 0;
  bool is_list_valid = true;
  double t_loop = 0;
  long long nentries = 0;
  std::vector<int> out_i, out_j, out_c, out_s;
  unsigned nbuild = 0;
  // dt = 0 for the start-up updateX: it leaves x as read and sets f_prev = f = 0, the state the
  // reference reaches with its setDt(-1)/updateX/setDt(1) start-up (src/run.cpp:193, :220-227).
  VERLET::setDt(0);
  std::chrono::steady_clock::time_point t0;
#pragma omp parallel num_threads(ENVIRON::NUM_THREAD)
  {
    if (omp_get_num_threads() != ENVIRON::NUM_THREAD) die("omp failed to start");
    int me = omp_get_thread_num();
    LOCAL local(me);
    NEIGHLIST list(me, pair_cutoff, skin, 2048, bin_list_pair);
    VERLET verlet(me);
    local.clear_vector(ENVIRON::f, 3);
#pragma omp barrier
    verlet.updateX();
#pragma omp barrier
#pragma omp single
    { VERLET::setDt(dt); }
#pragma omp single
    { t0 = std::chrono::steady_clock::now(); }
    for (int step = 0; step <= nsteps; ++step) {
      const bool last = step == nsteps;
      if (step > 0) verlet.updateX();  // step 0 force evaluation is on the configuration as read
#pragma omp single
      { is_list_valid = true; }
      bool lv = list.isvalid();
#pragma omp critical
      { is_list_valid = is_list_valid && lv; }
#pragma omp barrier
      if (!is_list_valid) {
        local.refold_and_tag();
#pragma omp barrier
#pragma omp single
        { ENVIRON::sort(); }
        local.generate_ghost();
#pragma omp barrier
#pragma omp single
        { ENVIRON::tidyup_ghost(); }
        local.update_ghost_pos<true>();
#pragma omp barrier
#pragma omp single
        { ENVIRON::sort_ghost(); }
        local.update_ghost_and_info();
#pragma omp barrier
        list.build<false>();
      } else {
        local.update_ghost_pos<false>();
      }
      local.clear_vector_cs(*ENVIRON::f, 3);
#pragma omp barrier
      numtype e_l = 0, v_l = 0;
      bond.compute<true>(me, &e_l, &v_l);
#pragma omp barrier
      angle.compute<true>(me, &e_l, &v_l);
#pragma omp barrier
      dihedral.compute<true>(me, &e_l, &v_l);
#pragma omp barrier
      pair.compute<true, false>(&list, &e_l, &v_l);
#pragma omp single
      { erg = 0; virial = 0; }
#pragma omp critical
      { erg += e_l; virial += v_l; }
#pragma omp barrier
      if (last) {
        // dump this thread's list as directed entries (atom id i, atom id j, image code, special)
#pragma omp critical
        {
          for (int i = 0; i < list.n_local; ++i) {
            int ii = list.local_list[i];
            for (int k = 0; k < list.num_neigh[i]; ++k) {
              int jj = list.nei_list[i][k];
              int code = 13;
              if (jj >= N) {
                int src = ENVIRON::i_ghost[jj - N];
                int cc[3];
                for (int dd = 0; dd < 3; ++dd) {
                  float sh = ENVIRON::x[jj][dd] - ENVIRON::x[src][dd];
                  cc[dd] = sh > ENVIRON::h_bl ? 2 : (sh < -ENVIRON::h_bl ? 0 : 1);
                }
                code = cc[0] * 9 + cc[1] * 3 + cc[2];
              }
              out_i.push_back(ENVIRON::atom_id[ii]); out_j.push_back(ENVIRON::atom_id[jj]);
              out_c.push_back(code); out_s.push_back(list.special[i][k]);
              ++nentries;
            }
          }
          if (me == 0) nbuild = list.nbuild;
        }
#pragma omp barrier
      }
      if (step < nsteps) {
        verlet.updateV();
      }
    }
#pragma omp single
    { t_loop = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
    double ke_l = verlet.computeKE();
#pragma omp critical
    { ekin += ke_l; }
  }

  // outputs: per-atom arrays by atom id as float64
  std::vector<double> xo(size_t(N) * 3), vo(size_t(N) * 3), fo(size_t(N) * 3);
  for (int i = 0; i < N; ++i) {
    int id = ENVIRON::atom_id[i];
    for (int k = 0; k < 3; ++k) { xo[3 * id + k] = ENVIRON::x[i][k]; vo[3 * id + k] = ENVIRON::v[i][k]; fo[3 * id + k] = ENVIRON::f[i][k]; }
  }
  auto wr = [&](const char* name, const void* p, size_t bytes) {
    std::string fn = outdir + "/" + name;
    FILE* fo2 = fopen(fn.c_str(), "wb");
    if (!fo2) die("cannot write output");
    fwrite(p, 1, bytes, fo2);
    fclose(fo2);
  };
  wr("x.bin", xo.data(), xo.size() * 8);
  wr("v.bin", vo.data(), vo.size() * 8);
  wr("f.bin", fo.data(), fo.size() * 8);
  std::vector<int> lst(size_t(nentries) * 4);
  for (long long k = 0; k < nentries; ++k) { lst[4 * k] = out_i[k]; lst[4 * k + 1] = out_j[k]; lst[4 * k + 2] = out_c[k]; lst[4 * k + 3] = out_s[k]; }
  wr("list.bin", lst.data(), lst.size() * 4);
  std::string fn = outdir + "/summary.txt";
  FILE* fs = fopen(fn.c_str(), "w");
  fprintf(fs, "natom %d\nnghost %d\nnentries %lld\nerg %.17g\nvirial %.17g\nekin %.17g\nnbuild %u\nnsteps %d\nt_loop %.9g\nthreads %d\n",
          N, ENVIRON::nghost, nentries, erg, virial, ekin, nbuild, nsteps, t_loop, ENVIRON::NUM_THREAD);
  fclose(fs);
  printf("ref_harness %s natom=%d nsteps=%d t_loop=%.6f s\n", mode.c_str(), N, nsteps, t_loop);
  return 0;
}
