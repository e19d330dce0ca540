"""Python handle on one nnpe context (a system resident in HBM), calling the C ABI.

The method names follow the reference classes they stand in for (ENVIRON, NEIGHLIST,
LJTableCoulDSF, BOND/ANGLE/DIHEDRAL, VERLET, GR, SQ); every call goes to the HIP library.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import NnpeError


def _f32(a):
    return np.ascontiguousarray(a, np.float32)


def _i32(a):
    return np.ascontiguousarray(a, np.int32)


def _p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


class Engine:
    def __init__(self, sysm, device=0):
        self.L = _lib.load()
        self.sys = sysm
        h = C.c_void_p()
        self._check(self.L.nnpe_create(C.byref(h), device), ctx=None)
        self.h = h
        self.natom = sysm.natom
        self.ntype = sysm.ntype
        self.npair = sysm.npair
        offs, flat = sysm.special_flat()
        k = [_f32(sysm.x), _f32(sysm.v), _f32(sysm.q), _f32(sysm.mass), _i32(sysm.typ), _i32(sysm.mol_id),
             _i32(sysm.intra_id), _i32(sysm.mol_type), _i32([m.n_atom for m in sysm.moltypes]), _i32(offs), _i32(flat)]
        self._check(self.L.nnpe_system_set(self.h, sysm.natom, sysm.ntype, C.c_float(sysm.bl), _p(k[0], C.c_float),
                                           _p(k[1], C.c_float), _p(k[2], C.c_float), _p(k[3], C.c_float),
                                           _p(k[4], C.c_int), _p(k[5], C.c_int), _p(k[6], C.c_int), sysm.nmol,
                                           _p(k[7], C.c_int), len(sysm.moltypes), _p(k[8], C.c_int),
                                           _p(k[9], C.c_int), _p(k[10], C.c_int)))
        self.gr_nbin = 0
        self.sq_nq = 0

    def _check(self, rc, ctx=True):
        if rc != 0:
            msg = self.L.nnpe_last_error(self.h).decode() if ctx else ""
            raise NnpeError(rc, msg)

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.L.nnpe_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        self._check(self.L.nnpe_synchronize(self.h))

    # ENVIRON state -----------------------------------------------------------------------
    def get(self, what):
        out = np.zeros((self.natom, 3), np.float64)
        self._check(self.L.nnpe_state_get(self.h, {"x": 0, "v": 1, "f": 2, "f_prev": 3}[what], _p(out, C.c_double)))
        return out

    def set(self, what, arr):
        a = np.ascontiguousarray(arr, np.float64)
        self._check(self.L.nnpe_state_set(self.h, {"x": 0, "v": 1, "f": 2, "f_prev": 3}[what], _p(a, C.c_double)))

    # NEIGHLIST ---------------------------------------------------------------------------
    def neigh_init(self, pair_cutoff, skin, gr_cutoff, short_cutoff=0.0):
        self._check(self.L.nnpe_neigh_init(self.h, C.c_float(pair_cutoff), C.c_float(skin), C.c_float(gr_cutoff),
                                           C.c_float(short_cutoff)))

    def build(self):
        self._check(self.L.nnpe_neigh_build(self.h))

    def isvalid(self):
        v = C.c_int(0)
        self._check(self.L.nnpe_neigh_isvalid(self.h, C.byref(v)))
        return bool(v.value)

    def list_size(self, which=0):
        n, g = C.c_longlong(0), C.c_int(0)
        self._check(self.L.nnpe_neigh_size(self.h, which, C.byref(n), C.byref(g)))
        return n.value, g.value

    def list_entries(self, which=0):
        n, _ = self.list_size(which)
        arrs = [np.zeros(max(n, 1), np.int32) for _ in range(4)]
        got = C.c_longlong(0)
        self._check(self.L.nnpe_neigh_export(self.h, which, *[_p(a, C.c_int) for a in arrs], C.c_longlong(n), C.byref(got)))
        return np.stack([a[:got.value] for a in arrs], 1)

    # forces ------------------------------------------------------------------------------
    def pair_table(self, cutoff, force_tbl, pot_tbl):
        self._tbl = (_f32(force_tbl), _f32(pot_tbl))
        self._check(self.L.nnpe_pair_table(self.h, C.c_float(cutoff), force_tbl.shape[1] - 1,
                                           _p(self._tbl[0], C.c_float), _p(self._tbl[1], C.c_float)))

    def pair_compute(self, ev=True):
        self._check(self.L.nnpe_pair_compute(self.h, int(ev)))

    def bonded_set_from_system(self):
        s = self.sys
        rb = getattr(s, "dihedral_rb", None)
        dih = (2, "dihedral", s.dihedral_coef) if rb is None else (3, "dihedral", rb)
        for kind, name, coef in ((0, "bond", s.bond_coef), (1, "angle", s.angle_coef), dih):
            terms = s.expand(name)
            coef = np.zeros((0, 6 if kind == 3 else 4), np.float32) if coef is None else _f32(coef)
            if len(terms):
                self._check(self.L.nnpe_bonded_set(self.h, kind, len(coef), _p(coef, C.c_float), len(terms),
                                                   _p(_i32(terms), C.c_int)))

    def bonded_compute(self, ev=True):
        self._check(self.L.nnpe_bonded_compute(self.h, int(ev)))

    def bonded_compute_noalloc(self, ev=True):
        """nnpe_bonded_compute with the partial sums in the library's fixed device array."""
        self._check(self.L.nnpe_bonded_compute_noalloc(self.h, int(ev)))

    def clear_force(self):
        self._check(self.L.nnpe_clear_force(self.h))

    def energy(self):
        out = np.zeros(2 * _lib.NTERM, np.float64)
        self._check(self.L.nnpe_energy_get(self.h, _p(out, C.c_double)))
        return {t: (out[2 * k], out[2 * k + 1]) for k, t in enumerate(_lib.TERMS)}

    def energy_clear(self):
        self._check(self.L.nnpe_energy_clear(self.h))

    # VERLET ------------------------------------------------------------------------------
    def verlet_set(self, dt, fp64=False):
        self._check(self.L.nnpe_verlet_set(self.h, C.c_double(dt), int(fp64)))

    def update_x(self):
        self._check(self.L.nnpe_verlet_update_x(self.h))

    def update_v(self):
        self._check(self.L.nnpe_verlet_update_v(self.h))

    def compute_ke(self):
        ke = C.c_double(0)
        self._check(self.L.nnpe_verlet_compute_ke(self.h, C.byref(ke)))
        return ke.value

    def scale_v(self, s):
        self._check(self.L.nnpe_verlet_scale_v(self.h, C.c_double(s)))

    # GR / SQ -----------------------------------------------------------------------------
    def gr_init(self, nbin_inner, r_max, r_max_inner):
        nb = C.c_int(0)
        self._check(self.L.nnpe_gr_init(self.h, nbin_inner, C.c_float(r_max), C.c_float(r_max_inner), C.byref(nb)))
        self.gr_nbin = nb.value
        return nb.value

    def gr_tally(self, full=False):
        self._check(self.L.nnpe_gr_tally(self.h, int(full)))

    def gr_get(self):
        counts = np.zeros((self.npair, self.gr_nbin), np.int64)
        h = np.zeros((self.npair, self.gr_nbin), np.float64)
        self._check(self.L.nnpe_gr_get(self.h, _p(counts, C.c_longlong), _p(h, C.c_double)))
        return counts, h

    def sq_init(self, nq, q_min, q_max):
        self.sq_nq = nq
        self._check(self.L.nnpe_sq_init(self.h, nq, C.c_double(q_min), C.c_double(q_max)))

    def sq_compute(self):
        out = np.zeros((self.npair, self.sq_nq), np.float64)
        self._check(self.L.nnpe_sq_compute(self.h, _p(out, C.c_double)))
        return out

    def sqd_init(self, nq, q_max):
        """SQD (direct k-space S(q)) vector set and bins; returns the number of vectors."""
        n = C.c_longlong(0)
        self.sqd_nq = nq
        self._check(self.L.nnpe_sqd_init(self.h, nq, C.c_float(q_max), C.byref(n)))
        return n.value

    def sqd_compute(self):
        self._check(self.L.nnpe_sqd_compute(self.h))

    def sqd_get(self):
        """qq [nq] (mean |G| per bin), sq [nq][npair], count [nq] of the last sqd_compute."""
        qq = np.zeros(self.sqd_nq)
        sq = np.zeros((self.sqd_nq, self.npair))
        cnt = np.zeros(self.sqd_nq, np.int64)
        self._check(self.L.nnpe_sqd_get(self.h, _p(qq, C.c_double), _p(sq, C.c_double), _p(cnt, C.c_longlong)))
        return qq, sq, cnt

    # RMDF ----------------------------------------------------------------------------------
    def rmdf_init(self, data, strength, gamma, target=0):
        """data [nsq][nq][3 + ntype] (q, S_exp, weight, form factors) on the S(q) grid."""
        d = _f32(data)
        if d.ndim != 3 or d.shape[1] != self.sq_nq or d.shape[2] != 3 + self.ntype:
            raise ValueError(f"RMDF data must be [nsq][{self.sq_nq}][{3 + self.ntype}], got {d.shape}")
        self._rmdf = dict(nsq=d.shape[0], target=target)
        self._check(self.L.nnpe_rmdf_init(self.h, d.shape[0], _p(d, C.c_float), C.c_float(strength),
                                          C.c_float(gamma), int(target)))

    def rmdf_set_delta(self, deltasq):
        d = _f32(deltasq).reshape(self._rmdf["nsq"], self.sq_nq)
        self._check(self.L.nnpe_rmdf_set_delta(self.h, _p(d, C.c_float)))

    def rmdf_step(self, potential=False):
        self._check(self.L.nnpe_rmdf_step(self.h, int(potential)))

    def rmdf_get(self):
        nsq = self._rmdf["nsq"]
        sq = np.zeros((nsq, self.sq_nq), np.float32)
        ds = np.zeros((nsq, self.sq_nq), np.float32)
        fq = np.zeros((self.npair, self.sq_nq), np.float32)
        self._check(self.L.nnpe_rmdf_get(self.h, _p(sq, C.c_float), _p(ds, C.c_float), _p(fq, C.c_float)))
        return sq, ds, fq

    def rmdf_tables(self, nbin):
        """Refined (force, potential) tables; nbin = the refined table's bin count."""
        shape = (self.npair, nbin + 1) if self._rmdf["target"] == 0 else (self.npair, 4, nbin + 1)
        f = np.zeros(shape, np.float32)
        u = np.zeros(shape, np.float32)
        self._check(self.L.nnpe_rmdf_tables(self.h, _p(f, C.c_float), _p(u, C.c_float)))
        return f, u

    # NNPOTENTIAL_TBL ---------------------------------------------------------------------
    def nnp_tbl_set(self, cutoff, nbin, coef_p, coef_w, coef_d, coef_r):
        self._nnp = [_f32(coef_p), _f32(coef_w), _f32(coef_d), _f32(coef_r)]
        self._check(self.L.nnpe_nnp_tbl_set(self.h, C.c_float(cutoff), nbin, *[_p(a, C.c_float) for a in self._nnp]))

    def nnp_tbl_descriptor(self):
        self._check(self.L.nnpe_nnp_tbl_descriptor(self.h))

    def nnp_tbl_force(self, ev=True):
        self._check(self.L.nnpe_nnp_tbl_force(self.h, int(ev)))

    # symmetry functions ------------------------------------------------------------------
    def sf_set(self, rc, mu, eta_r, eta_a):
        self._sf = (_f32(mu), _f32(eta_a))
        self._check(self.L.nnpe_sf_set(self.h, C.c_float(rc), len(mu), _p(self._sf[0], C.c_float), C.c_float(eta_r),
                                       len(eta_a), _p(self._sf[1], C.c_float)))

    def sf_descriptor(self):
        self._check(self.L.nnpe_sf_descriptor(self.h))

    def sf_force(self, ev=True):
        self._check(self.L.nnpe_sf_force(self.h, int(ev)))

    # TFCALL buffers / MLP ----------------------------------------------------------------
    def nnp_get(self, what):
        D = C.c_int(0)
        k = {"G": 0, "dG": 1, "e_atom": 2}[what]
        self._check(self.L.nnpe_nnp_get(self.h, k, None, C.byref(D)))
        out = np.zeros((self.natom, D.value), np.float32)
        self._check(self.L.nnpe_nnp_get(self.h, k, _p(out, C.c_float), C.byref(D)))
        return out[:, 0] if what == "e_atom" else out

    def nnp_set_dedg(self, dG):
        d = _f32(dG)
        self._check(self.L.nnpe_nnp_set_dedg(self.h, _p(d, C.c_float), d.shape[1]))

    def mlp_set(self, widths, params, out_scale=1.0, precision="fp32", shift=None, scale=None):
        """shift, scale: [ntype][D] input standardisation x = (G - shift) * scale (None: identity)."""
        prec = {"fp32": 0, "bf16": 1, "fp16": 2, "bf16x1": 3}[precision]
        sh = None if shift is None else _f32(shift)
        sc = None if scale is None else _f32(scale)
        self._mlp = (_i32(widths), _f32(params), sh, sc)
        self._check(self.L.nnpe_mlp_set(self.h, len(widths) - 1, _p(self._mlp[0], C.c_int), _p(self._mlp[1], C.c_float),
                                        C.c_float(out_scale), prec, None if sh is None else _p(sh, C.c_float),
                                        None if sc is None else _p(sc, C.c_float)))

    def mlp_eval(self, ev=True):
        self._check(self.L.nnpe_mlp_eval(self.h, int(ev)))

    # MD loop -----------------------------------------------------------------------------
    def md_run(self, nsteps, dt=0.5, fp64=False, rebuild_every=10, pair=True, bonded=True, nnp=0, gr_every=0,
               gr_long_every=0, sq=False, log_every=0, use_graph=True, rmdf=False):
        p = _lib.MdParams(dt, int(fp64), rebuild_every, int(pair), int(bonded), int(nnp), gr_every, gr_long_every,
                          int(sq), log_every, int(use_graph), int(rmdf))
        cap = nsteps // log_every + 2 if log_every > 0 else 1
        log = np.zeros((cap, 2 * _lib.NTERM), np.float64)
        n = C.c_int(0)
        self._check(self.L.nnpe_md_run(self.h, C.byref(p), nsteps, _p(log, C.c_double), cap, C.byref(n)))
        return log[:n.value]

    def md_stats(self):
        out = np.zeros(8, np.int64)
        self._check(self.L.nnpe_md_stats(self.h, _p(out, C.c_longlong)))
        keys = ("steps", "nbuild", "skin_violations", "max_long", "max_short", "nghost", "cap_long", "cap_short")
        return dict(zip(keys, out.tolist()))

    def md_reset_step(self):
        self._check(self.L.nnpe_md_reset_step(self.h))

    def md_set_step(self, step):
        self._check(self.L.nnpe_md_set_step(self.h, int(step)))

    def md_run_checkpointed(self, nsteps, chunk=1000, retries=3, **kw):
        """md_run in chunks of `chunk` steps (a multiple of the rebuild interval); a chunk that
        ends on a list or ghost overflow (the engine grows its capacities) is repeated from the
        positions, velocities and accelerations saved at its start.  Returns the concatenated
        log and the number of repeated chunks."""
        logs, redo, done = [], 0, 0
        while done < nsteps:
            n = min(chunk, nsteps - done)
            step0 = self.md_stats()["steps"]
            saved = [self.get(w) for w in ("x", "v", "f", "f_prev")]
            for attempt in range(retries + 1):
                try:
                    logs.append(self.md_run(n, **kw))
                    break
                except _lib.NnpeError as err:
                    if err.code not in (4, 5) or attempt == retries:
                        raise
                    redo += 1
                    for w, a in zip(("x", "v", "f", "f_prev"), saved):
                        self.set(w, a)
                    self.md_set_step(step0)
            done += n
        return (np.concatenate(logs) if logs else np.zeros((0, 2 * _lib.NTERM))), redo

    def profile(self, stage, reps=20):
        """Average ms of one launch of a stage (HIP events on the engine's stream)."""
        ms = C.c_double(0)
        self._check(self.L.nnpe_profile_stage(self.h, _lib.STAGES[stage], reps, C.byref(ms)))
        return ms.value


class Serve:
    """Request loop over independent boxes (nnpe_serve_*): boxes are queued with submit(), admitted
    by the arena's atom capacity, advanced together one chunk per advance(), and collected with
    result()."""

    def __init__(self, sysm_ref, widths, params, shift, scale, precision="bf16", max_atoms=200000, max_boxes=256,
                 pair_cutoff=9.0, skin=2.0, nnp_cutoff=6.0, rebuild_every=10, dt=0.5, fp64=True, pair_nbin=900,
                 lj=None, device=0):
        """sysm_ref: a System of the same kind as the requests (molecule types, bonded coefficients,
        element count)."""
        from nnpengine_amd import model as M
        from nnpengine_amd import system as S
        self.L = _lib.load()
        self.rebuild_every = rebuild_every
        ft, ut = S.lj_table(sysm_ref, pair_cutoff, pair_nbin, lj if lj is not None else S.WATER_LJ)
        sfp = M.sf_params(nnp_cutoff)
        offs, flat = sysm_ref.special_flat()
        coefs = []
        for c in (sysm_ref.bond_coef, sysm_ref.angle_coef, sysm_ref.dihedral_coef):
            coefs.append(np.zeros((0, 4), np.float32) if c is None else _f32(c))
        keep = dict(ft=_f32(ft), ut=_f32(ut), mt_natom=_i32([m.n_atom for m in sysm_ref.moltypes]), offs=_i32(offs),
                    flat=_i32(flat), mu=_f32(sfp["mu"]), eta=_f32(sfp["eta_a"]), widths=_i32(widths), params=_f32(params),
                    shift=_f32(shift), scale=_f32(scale), c0=coefs[0], c1=coefs[1], c2=coefs[2])
        self._keep = keep
        prec = {"fp32": 0, "bf16": 1, "fp16": 2, "bf16x1": 3}[precision]
        cfg = _lib.ServeConfig(device, max_atoms, max_boxes, sysm_ref.ntype, pair_cutoff, skin, nnp_cutoff, rebuild_every,
                               dt, int(fp64), pair_nbin, _p(keep["ft"], C.c_float), _p(keep["ut"], C.c_float),
                               len(coefs[0]), len(coefs[1]), len(coefs[2]), _p(keep["c0"], C.c_float),
                               _p(keep["c1"], C.c_float), _p(keep["c2"], C.c_float), len(sysm_ref.moltypes),
                               _p(keep["mt_natom"], C.c_int), _p(keep["offs"], C.c_int), _p(keep["flat"], C.c_int),
                               _p(keep["mu"], C.c_float), float(sfp["eta_r"]), _p(keep["eta"], C.c_float),
                               len(widths) - 1, _p(keep["widths"], C.c_int), _p(keep["params"], C.c_float), 0.5, prec,
                               _p(keep["shift"], C.c_float), _p(keep["scale"], C.c_float))
        h = C.c_void_p()
        rc = self.L.nnpe_serve_create(C.byref(cfg), C.byref(h))
        if rc:
            raise _lib.NnpeError(rc, "nnpe_serve_create")
        self.h = h
        self.natom = {}

    def _check(self, rc):
        if rc:
            raise _lib.NnpeError(rc, self.L.nnpe_serve_last_error(self.h).decode())

    def submit(self, sysm, nsteps):
        """Queue a box (a System); returns the request id."""
        k = [_f32(sysm.x), _f32(sysm.v), _f32(sysm.q), _f32(sysm.mass), _i32(sysm.typ), _i32(sysm.mol_id),
             _i32(sysm.intra_id), _i32(sysm.mol_type)]
        terms = [_i32(sysm.expand(n)) for n in ("bond", "angle", "dihedral")]
        tp = [_p(t, C.c_int) if len(t) else None for t in terms]
        rid = C.c_longlong(0)
        self._check(self.L.nnpe_serve_submit(self.h, sysm.natom, C.c_float(sysm.bl), _p(k[0], C.c_float),
                                             _p(k[1], C.c_float), _p(k[2], C.c_float), _p(k[3], C.c_float),
                                             _p(k[4], C.c_int), _p(k[5], C.c_int), _p(k[6], C.c_int), sysm.nmol,
                                             _p(k[7], C.c_int), len(terms[0]), tp[0], len(terms[1]), tp[1],
                                             len(terms[2]), tp[2], nsteps, C.byref(rid)))
        self.natom[rid.value] = sysm.natom
        return rid.value

    def advance(self):
        """One chunk for every resident box; returns the ids that finished."""
        cap = 4096
        ids = np.zeros(cap, np.int64)
        n = C.c_int(0)
        self._check(self.L.nnpe_serve_advance(self.h, _p(ids, C.c_longlong), cap, C.byref(n)))
        return ids[:min(n.value, cap)].tolist()

    def result(self, rid):
        n = self.natom.pop(rid)
        x = np.zeros((n, 3), np.float64)
        v = np.zeros((n, 3), np.float64)
        self._check(self.L.nnpe_serve_result(self.h, rid, _p(x, C.c_double), _p(v, C.c_double)))
        return x, v

    def stats(self):
        out = np.zeros(8, np.int64)
        self._check(self.L.nnpe_serve_stats(self.h, _p(out, C.c_longlong)))
        keys = ("resident", "atoms", "queued", "finished", "chunks", "relayouts", "arena_bytes", "max_atoms")
        return dict(zip(keys, out.tolist()))

    def close(self):
        if getattr(self, "h", None):
            self.L.nnpe_serve_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
