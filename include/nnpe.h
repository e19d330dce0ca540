/* nnpe.h -- C ABI of the MI355X (gfx950) engine for NNPEngine's molecular-liquid MD hot path.
 *
 * Plain C: an opaque context, host pointers + sizes, int status codes (0 = NNPE_OK).  All
 * device state (positions, lists, tables, MLP weights) lives in the context's HBM buffers; the
 * host passes and receives per-atom arrays in atom-id order (the reference's input order).
 * Every call is ordered on the context's HIP stream; getters synchronise it.
 *
 * Each entry point names the reference interface it replaces (paths in the reference repo).
 */
#ifndef NNPE_H
#define NNPE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct nnpe_ctx nnpe_ctx;

enum nnpe_status {
  NNPE_OK = 0,
  NNPE_ERR_ARG = 1,             /* bad argument / shape */
  NNPE_ERR_HIP = 2,             /* HIP runtime error (message in nnpe_last_error) */
  NNPE_ERR_STATE = 3,           /* call out of order (e.g. compute before a list build) */
  NNPE_ERR_NEIGH_OVERFLOW = 4,  /* per-atom list capacity exceeded ("nei_list overflow", src/neighlist.cpp:327) */
  NNPE_ERR_GHOST_OVERFLOW = 5,  /* ghost capacity exceeded ("ghost overflow", src/environ.cpp:440) */
  NNPE_ERR_CELL_RANGE = 6,      /* atom outside the cell grid (the CHECK_BINNING case, src/environ.cpp:245-259) */
  NNPE_ERR_NO_DEVICE = 7,       /* no HIP device */
  NNPE_ERR_UNSUPPORTED = 8      /* configuration the library does not support */
};

/* per-atom state selectors for nnpe_state_get / nnpe_state_set */
enum nnpe_state_kind { NNPE_X = 0, NNPE_V = 1, NNPE_F = 2, NNPE_FPREV = 3 };

/* energy/virial terms accumulated by the compute calls (nnpe_energy_get) */
enum nnpe_term { NNPE_TERM_PAIR = 0, NNPE_TERM_BOND = 1, NNPE_TERM_ANGLE = 2, NNPE_TERM_DIHEDRAL = 3,
                 NNPE_TERM_NNP_PAIR = 4, NNPE_TERM_NN = 5, NNPE_TERM_SF = 6, NNPE_TERM_KE = 7,
                 NNPE_TERM_RMDF = 8, NNPE_NTERM = 9 };

const char* nnpe_status_string(int status);
const char* nnpe_last_error(const nnpe_ctx* ctx);

/* Lifecycle.  One context per device and system. */
int nnpe_create(nnpe_ctx** out, int device);
int nnpe_destroy(nnpe_ctx* ctx);
int nnpe_synchronize(nnpe_ctx* ctx);

/* ENVIRON::initialize (src/environ.cpp:22-141) after the host has parsed the configuration and
 * molecule files (MOLECULE::from_file, src/molecule.cpp:25-131).
 * x, v: [natom][3]; q, mass, typ, mol_id, intra_id: [natom]; mol_type: [nmol];
 * mt_special: concatenated [mt_natom[t]][mt_natom[t]] special classes at mt_special_off[t].
 * v may be NULL (zero velocities).  natom < 8 388 608 (a neighbour-list entry holds a 23-bit
 * row); larger systems return NNPE_ERR_ARG. */
int nnpe_system_set(nnpe_ctx* ctx, int natom, int ntype, float bl,
                    const float* x, const float* v, const float* q, const float* mass,
                    const int* typ, const int* mol_id, const int* intra_id,
                    int nmol, const int* mol_type,
                    int nmoltype, const int* mt_natom, const int* mt_special_off, const int* mt_special);

/* Per-atom state in atom-id order as double [natom][3] (what = nnpe_state_kind).
 * Replaces the host reads/writes of ENVIRON::x / v / f / f_prev (src/environ.h:31-32). */
int nnpe_state_get(nnpe_ctx* ctx, int what, double* out);
int nnpe_state_set(nnpe_ctx* ctx, int what, const double* in);

/* ENVIRON::init_grid + NEIGHLIST::build_bin_list + LOCAL::set_ghost_lim
 * (src/environ.cpp:143-155, src/neighlist.cpp:35-168, src/run.cpp:72-82).
 * comm cutoff = skin + max(gr_cutoff, pair_cutoff); cell size = (pair_cutoff + skin) / 2.
 * short_cutoff > 0 also builds short lists to short_cutoff + skin, one per neighbour element (the
 * symmetry-function NNP environments need them). */
int nnpe_neigh_init(nnpe_ctx* ctx, float pair_cutoff, float skin, float gr_cutoff, float short_cutoff);
/* Rebuild: LOCAL::refold_and_tag, ENVIRON::sort, generate_ghost, tidyup_ghost,
 * update_ghost_pos<true>, sort_ghost, update_ghost_and_info, NEIGHLIST::build<true>
 * (src/run.cpp:249-276).  Lists are full (each atom lists every neighbour and image). */
int nnpe_neigh_build(nnpe_ctx* ctx);
/* NEIGHLIST::isvalid (src/neighlist.cpp:178-187): *valid = 1 if no atom moved more than skin/2. */
int nnpe_neigh_isvalid(nnpe_ctx* ctx, int* valid);
/* Number of directed entries (which = 0 pair list, 1 short list) and the ghost count. */
int nnpe_neigh_size(nnpe_ctx* ctx, int which, long long* n_entries, int* n_ghost);
/* Export directed entries as atom ids + image code ((sx+1)*9+(sy+1)*3+(sz+1), 13 = same image)
 * + special class; cap = array capacity. */
int nnpe_neigh_export(nnpe_ctx* ctx, int which, int* ai, int* aj, int* code, int* special,
                      long long cap, long long* n_out);

/* LJTableCoulDSF (src/pair.cpp:171-263): tables force[npair][nbin+1] (F/r) and
 * potential[npair][nbin+1] at r = (i+1/2) dr, last entry 0. */
int nnpe_pair_table(nnpe_ctx* ctx, float cutoff, int nbin, const float* force, const float* potential);
/* LJTableCoulDSF::compute<ev,false> (src/pair.cpp:289-419).  Overwrites the force array. */
int nnpe_pair_compute(nnpe_ctx* ctx, int ev);

/* BOND / ANGLE / DIHEDRAL setParam + autoGenerate (src/bond.cpp:27-82, src/angle.cpp:26-82,
 * src/dihedral.cpp:27-85).  kind 0 bond (terms [n][3]), 1 angle ([n][4]), 2 dihedral ([n][5]);
 * terms are (type, atom ids...); coef [ntyp][4] in the classes' stored form.  kind 3 (no
 * counterpart in the reference, whose dihedral is cubic in cos phi): dihedral terms [n][5] with
 * coef [ntyp][6] = C0..C5 of U = sum_n C_n cos^n(phi) (Ryckaert-Bellemans; phi = 0 cis, as the
 * reference measures it); it replaces the kind-2 terms and their energy goes to the same slot. */
int nnpe_bonded_set(nnpe_ctx* ctx, int kind, int ntyp, const float* coef, int nterm, const int* terms);
/* BOND/ANGLE/DIHEDRAL::compute<ev> (src/bond.cpp:87-116, src/angle.cpp:87-136,
 * src/dihedral.cpp:90-147) on unwrapped molecule positions.  Adds to the force array. */
int nnpe_bonded_compute(nnpe_ctx* ctx, int ev);
/* No counterpart in the reference: nnpe_bonded_compute for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same forces and energy/virial sums, bit for bit; the per-workgroup partial sums live in a
 * fixed device array (768 KiB per device, held until the process ends) instead of the
 * context's buffer, so the call allocates nothing.  At most 1,048,576 atoms
 * (NNPE_ERR_UNSUPPORTED beyond).  Calls that overlap on one device share that array and race. */
int nnpe_bonded_compute_noalloc(nnpe_ctx* ctx, int ev);

/* Energy and virial sums (FP64) of the ev compute calls since the last clear:
 * out[2*term] = energy, out[2*term+1] = virial, term = nnpe_term. */
int nnpe_energy_get(nnpe_ctx* ctx, double* out);
int nnpe_energy_clear(nnpe_ctx* ctx);

/* VERLET (src/verlet.cpp:29-97).  fp64 = 0: the reference's FP32 update; fp64 = 1: FP64
 * positions/velocities/accelerations with FP32 forces. */
int nnpe_verlet_set(nnpe_ctx* ctx, double dt, int fp64);
int nnpe_verlet_update_x(nnpe_ctx* ctx);
int nnpe_verlet_update_v(nnpe_ctx* ctx);
int nnpe_verlet_compute_ke(nnpe_ctx* ctx, double* ke);
int nnpe_verlet_scale_v(nnpe_ctx* ctx, double s);
/* LOCAL::clear_vector(ENVIRON::f) (src/environ.cpp:526-534). */
int nnpe_clear_force(nnpe_ctx* ctx);

/* GR (src/grsq.cpp:28-80): nbin_inner bins to r_max_inner, extended to r_max. */
int nnpe_gr_init(nnpe_ctx* ctx, int nbin_inner, float r_max, float r_max_inner, int* nbin_total);
/* GR::tally<true>(list) (full = 0, src/grsq.cpp:94-149) or GR::tally(bin_list, local)
 * (full = 1, src/grsq.cpp:151-205).  Counts are 2 per unordered pair. */
int nnpe_gr_tally(nnpe_ctx* ctx, int full);
/* GR::reduce_local (src/grsq.cpp:207-219): counts [npair][nbin] and h(r) = g(r) - 1. */
int nnpe_gr_get(nnpe_ctx* ctx, long long* counts, double* h);
/* SQ (src/grsq.cpp:258-322) with q_i = q_min + (i+1/2)(q_max-q_min)/nq, accumulated in FP64. */
int nnpe_sq_init(nnpe_ctx* ctx, int nq, double q_min, double q_max);
int nnpe_sq_compute(nnpe_ctx* ctx, double* out);

/* SQD ctor (src/grsq.cpp:362-491): the half space of reciprocal vectors G = 2 pi (h, k, l) / L
 * with |h|,|k|,|l| <= int(q_max / g) and h^2 + k^2 + l^2 <= int((q_max / g)^2), g = 2 pi / L,
 * binned by |G| into nq bins of width q_max / nq (the reference's float arithmetic decides the
 * limits and the bins).  *n_gvec receives the number of vectors.  One box, at most 4 elements. */
int nnpe_sqd_init(nnpe_ctx* ctx, int nq, float q_max, long long* n_gvec);
/* SQD::compute (src/grsq.cpp:571-606) on the current positions: per bin the average over its
 * vectors of S_a S_b + C_a C_b (S_a, C_a the sin / cos sums over the atoms of element a),
 * times natom / (N_a N_b), minus natom / N_a for a == b.  Exact integer phases, FP64 sums. */
int nnpe_sqd_compute(nnpe_ctx* ctx);
/* Results of the last nnpe_sqd_compute: qq [nq] mean |G| of each bin, sq [nq][npair], count [nq]
 * vectors per bin (SQD::dump, src/grsq.cpp:608-644, skips the empty bins); any may be NULL. */
int nnpe_sqd_get(nnpe_ctx* ctx, double* qq, double* sq, long long* count);

/* RMDF ctor (src/grsq.cpp:899-1060): refine the pair tables against nsq measured S(q) sets.
 * data [nsq][nq][3 + ntype] rows (q, S_exp(q), weight, form factor per element) on the S(q)
 * grid of nnpe_sq_init, which must have q_min = 0; the refined table's bins must be the inner
 * g(r) bins.  gamma = 1 is direct refinement (DA/RMD), 0 <= gamma < 1 FMIRL.  target 0: the
 * LJTableCoulDSF table (nnpe_pair_table); target 1: the NNPOTENTIAL_TBL pair tables, special
 * classes 0 and 3 (the latter scaled by 0.5).  The tables as they are now are kept as the base. */
int nnpe_rmdf_init(nnpe_ctx* ctx, int nsq, const float* data, float strength, float gamma, int target);
/* RMDF::loadDeltaSQ (src/grsq.cpp:1419-1443): relaxed residual [nsq][nq] to resume from. */
int nnpe_rmdf_set_delta(nnpe_ctx* ctx, const float* deltasq);
/* compute_Qspace then compute_Rspace[_NNP] (+ compute_Potential[_NNP] when potential) on the
 * current S(q) (src/grsq.cpp:1079-1377); residual energy into NNPE_TERM_RMDF (stored). */
int nnpe_rmdf_step(nnpe_ctx* ctx, int potential);
/* Last step's S_n(q) and relaxed residual [nsq][nq] and force_q [npair][nq]; any may be NULL. */
int nnpe_rmdf_get(nnpe_ctx* ctx, float* sq, float* deltasq, float* force_q);
/* The refined tables, force then potential, in the target's layout. */
int nnpe_rmdf_tables(nnpe_ctx* ctx, float* force, float* potential);

/* NNPOTENTIAL_TBL ctor (src/pair_nnp.cpp:433-598): tables built from the weight-file arrays
 * coef_p [npair][4][6] (eps, sigma, c8, c10, c12, qij per special class) and
 * coef_w / coef_d / coef_r [npair][80] (20 descriptors x 4 special classes). */
int nnpe_nnp_tbl_set(nnpe_ctx* ctx, float cutoff, int nbin, const float* coef_p, const float* coef_w,
                     const float* coef_d, const float* coef_r);
/* NNPOTENTIAL_TBL::computeDescriptor (src/pair_nnp.cpp:625-688): 20 descriptors per atom. */
int nnpe_nnp_tbl_descriptor(nnpe_ctx* ctx);
/* NNPOTENTIAL_TBL::compute<ev> (src/pair_nnp.cpp:710-793) with the current dE/dG; adds to the force. */
int nnpe_nnp_tbl_force(nnpe_ctx* ctx, int ev);

/* Radial + angular symmetry-function descriptors (16 radial centres mu[16] with exponent eta_r
 * per neighbour element; 4 angular exponents eta[4] x (zeta, lambda) in {1,2} x {+1,-1} per
 * element pair), cutoff rc with fc = (cos(pi r/rc)+1)/2.  D = 16*ntype + 16*npair. */
int nnpe_sf_set(nnpe_ctx* ctx, float rc, int nrad, const float* mu, float eta_r, int neta, const float* eta);
/* Descriptors of the current positions; also records the pairs within rc for nnpe_sf_force. */
int nnpe_sf_descriptor(nnpe_ctx* ctx);
/* Symmetry-function forces from the current dE/dG over the pairs recorded by the last
 * nnpe_sf_descriptor (the reference's computeDescriptor -> nnEval -> compute order); adds to
 * the force, virial in NNPE_TERM_SF. */
int nnpe_sf_force(nnpe_ctx* ctx, int ev);

/* TFCALL buffers (src/pair_nnp.cpp:490-496, :697-706) in atom-id order: what 0 = descriptors
 * (inputs[0]), 1 = dE/dG (outputs[1]), 2 = per-atom network energy; *D = row width. */
int nnpe_nnp_get(nnpe_ctx* ctx, int what, float* out, int* D);
/* Provide dE/dG [natom][D] from the host (the reference's external-network path). */
int nnpe_nnp_set_dedg(nnpe_ctx* ctx, const float* dG, int D);

/* Per-element MLP replacing TFCALL::evaluate (src/pair_nnp.cpp:92-97).  widths[nlayer+1]
 * (input D, hidden widths multiples of 32, output 1); params per element, per layer:
 * W [in][out] row-major then b [out].  E_i = out_scale * network((G_i - in_shift) * in_scale)
 * with the input standardisation [ntype][D] applied in FP32 before any 16-bit rounding
 * (NULL = identity).  precision 0 = FP32, 1 = BF16 with weights (and backward gradients) as
 * hi + lo BF16 pairs, 2 = FP16, 3 = BF16 with one term per value (MFMA, FP32 accumulation,
 * per-layer power-of-two scaling in the FP16 form). */
int nnpe_mlp_set(nnpe_ctx* ctx, int nlayer, const int* widths, const float* params, float out_scale,
                 int precision, const float* in_shift, const float* in_scale);
/* nnEval: descriptors -> per-atom energy and dE/dG; energy sum in NNPE_TERM_NN when ev. */
int nnpe_mlp_eval(nnpe_ctx* ctx, int ev);

/* MD loop of classical / nnp_run (src/run.cpp:230-393, :1127-1353) without the thermostat:
 * updateX -> rebuild -> g(r)/S(q) -> forces -> updateV -> log. */
typedef struct nnpe_md_params {
  double dt;          /* fs (the reference hard-codes 1 fs, src/run.cpp:226) */
  int fp64;           /* 1: FP64 positions/velocities/accelerations */
  int rebuild_every;  /* > 0: rebuild every n steps, steps replayed from captured hipGraphs when
                         use_graph; 0: NEIGHLIST::isvalid decides each step (host read-back) */
  int pair;           /* tabulated LJ + shifted-force Coulomb */
  int bonded;         /* bonds / angles / dihedrals */
  int nnp;            /* 0 none, 1 NNPOTENTIAL_TBL descriptors, 2 symmetry functions; needs nnpe_mlp_set */
  int gr_every;       /* g(r) list tally every n steps (0 = never) */
  int gr_long_every;  /* full stencil tally every n steps (multiple of gr_every, 0 = never) */
  int sq;             /* S(q) after each tally */
  int log_every;      /* energy/virial record every n steps (0 = never) */
  int use_graph;      /* replay captured step graphs (fixed-interval rebuilds only) */
  int rmdf;           /* RMDF refinement after each S(q) (needs nnpe_rmdf_init, sq, gr_every) */
} nnpe_md_params;
/* Run nsteps; log_out receives [nlog][2*NNPE_NTERM] records (energy, virial per term; the
 * NNPE_TERM_KE slot holds the kinetic energy) for steps s with s % log_every == 0. */
int nnpe_md_run(nnpe_ctx* ctx, const nnpe_md_params* params, int nsteps, double* log_out, int log_cap, int* nlog);
/* out[8]: steps done, list builds, late fixed-interval rebuilds (skin violations), max entries
 * per atom (pair list, short list), ghosts, list capacities (pair, short). */
int nnpe_md_stats(nnpe_ctx* ctx, long long* out);
/* Restart the step counter (the next step evaluates forces on the positions as they are). */
int nnpe_md_reset_step(nnpe_ctx* ctx);
/* Continue at step `step` (> 0: the next step starts with the position update, as inside a
 * run).  With nnpe_state_set of X, V and FPREV (the accelerations) saved at that step, a run
 * that stopped on a list or ghost overflow (capacities grown) is repeated from a checkpoint.
 * No reference counterpart (the reference ends the program on an overflow). */
int nnpe_md_set_step(nnpe_ctx* ctx, long long step);

/* ---- Serving: a request loop over independent boxes -----------------------------------------
 * No reference counterpart (the reference runs one system per process).  Queued boxes are
 * admitted first-in first-out while the arena's atom capacity (its preallocated HBM footprint)
 * allows; all resident boxes advance together, in chunks of rebuild_every steps, in the same
 * fused launches (one arena of rows, cells and ghosts with per-box L and grids), the steps of a
 * chunk replayed from captured hipGraphs.  Tabulated LJ/Coulomb, bonded terms and the
 * symmetry-function NNP, as nnpe_md_run with pair = bonded = 1, nnp = 2. */
typedef struct nnpe_serve nnpe_serve;
typedef struct nnpe_serve_config {
  int device;
  int max_atoms;          /* arena capacity in atoms (sets the HBM footprint) */
  int max_boxes;          /* boxes resident at once */
  int ntype;
  float pair_cutoff, skin, nnp_cutoff;
  int rebuild_every;      /* chunk length: lists are rebuilt at every chunk start */
  double dt;
  int fp64;
  int pair_nbin;          /* tabulated pair, [npair][nbin+1] each (nnpe_pair_table) */
  const float* pair_force;
  const float* pair_potential;
  int nbond_types, nangle_types, ndihedral_types;       /* nnpe_bonded_set coefficients */
  const float* bond_coef;
  const float* angle_coef;
  const float* dihedral_coef;
  int nmoltype;           /* molecule types, as nnpe_system_set */
  const int* mt_natom;
  const int* mt_special_off;
  const int* mt_special;
  const float* sf_mu;     /* symmetry functions, as nnpe_sf_set (16 radial, 4 angular) */
  float sf_eta_r;
  const float* sf_eta;
  int nlayer;             /* network, as nnpe_mlp_set */
  const int* widths;
  const float* params;
  float out_scale;
  int precision;
  const float* in_shift;
  const float* in_scale;
} nnpe_serve_config;
int nnpe_serve_create(const nnpe_serve_config* cfg, nnpe_serve** out);
int nnpe_serve_destroy(nnpe_serve* s);
/* Queue one box (topology as nnpe_system_set with ids local to the box; bonded terms as
 * nnpe_bonded_set, local atom ids, grouped by molecule); nsteps a multiple of rebuild_every. */
int nnpe_serve_submit(nnpe_serve* s, int natom, float bl, const float* x, const float* v, const float* q,
                      const float* mass, const int* typ, const int* mol_id, const int* intra_id, int nmol,
                      const int* mol_type, int nbond, const int* bonds, int nangle, const int* angles, int ndih,
                      const int* dihedrals, int nsteps, long long* id);
/* Admit what fits, advance every resident box by one chunk, retire the finished boxes: their
 * ids go to done_ids (up to cap), their count to ndone. */
int nnpe_serve_advance(nnpe_serve* s, long long* done_ids, int cap, int* ndone);
/* Final positions and velocities [natom][3] of a finished request (local atom order); the
 * request is released. */
int nnpe_serve_result(nnpe_serve* s, long long id, double* x, double* v);
/* out[8]: resident boxes, resident atoms, queued, finished unread, chunks, relayouts, arena bytes,
 * atom capacity */
int nnpe_serve_stats(nnpe_serve* s, long long* out);
const char* nnpe_serve_last_error(nnpe_serve* s);

/* Per-stage timing (no reference counterpart: the reference has no timers).  Average ms of one
 * launch of `stage` over `reps` back-to-back launches, HIP events on the context stream. */
enum nnpe_stage { NNPE_STAGE_PAIR = 0, NNPE_STAGE_BONDED = 1, NNPE_STAGE_SF_DESC = 2, NNPE_STAGE_MLP = 3,
                  NNPE_STAGE_SF_ADJOINT = 4, NNPE_STAGE_SF_FORCE = 5, NNPE_STAGE_REBUILD = 6,
                  NNPE_STAGE_GR_FULL = 7, NNPE_STAGE_GR_LIST = 8, NNPE_STAGE_VERLET = 9, NNPE_STAGE_SQ = 10,
                  NNPE_STAGE_TBL_DESC = 11, NNPE_STAGE_TBL_FORCE = 12,
                  NNPE_STAGE_COPY = 13 /* 64 MiB read + 64 MiB write streaming copy */,
                  NNPE_STAGE_REBUILD_GR = 14 /* rebuild with the full g(r) tally folded in */,
                  NNPE_STAGE_MFMA = 15 /* MFMA issue-rate probe: 2048 waves x 2048 rounds x 4
                                          v_mfma_f32_32x32x16_bf16 = 5.50e11 flop */ };
int nnpe_profile_stage(nnpe_ctx* ctx, int stage, int reps, double* avg_ms);

#ifdef __cplusplus
}
#endif
#endif /* NNPE_H */
