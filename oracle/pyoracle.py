"""ctypes wrapper of the CPU oracle (liboracle.so, built from oracle/oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, as the test reference.  prec=0 reproduces the reference's float32 arithmetic,
prec=1 is the FP64 oracle.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_lib = None


def build(force=False):
    if force or not os.path.exists(LIB_PATH) or \
            os.path.getmtime(LIB_PATH) < os.path.getmtime(os.path.join(HERE, "oracle.cpp")):
        subprocess.check_call(["make", "-s", "-C", HERE, "liboracle.so"])


def lib():
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.orc_create.restype = vp
        L.orc_create.argtypes = [C.c_int]
        L.orc_destroy.argtypes = [vp]
        for name in dir(L):
            pass
        L.orc_list_size.restype = C.c_longlong
        L.orc_list_export.restype = C.c_longlong
        L.orc_compute_ke.restype = C.c_double
        L.orc_mlp_eval.restype = C.c_double
        L.orc_rmdf_step.restype = C.c_double
        for fn in ("orc_init", "orc_neigh_init", "orc_num_gx", "orc_rebuild", "orc_update_ghost_pos",
                   "orc_isvalid", "orc_list_size", "orc_list_export", "orc_get", "orc_set",
                   "orc_get_cell_ids", "orc_get_xmol_codes", "orc_clear_f", "orc_pair_set",
                   "orc_pair_compute", "orc_bond_compute", "orc_angle_compute",
                   "orc_dihedral_compute", "orc_dihedral_poly_compute", "orc_set_dt", "orc_update_x", "orc_update_v",
                   "orc_compute_ke", "orc_gr_init", "orc_gr_tally", "orc_gr_get", "orc_sq",
                   "orc_nnp_tbl_init", "orc_nnp_tbl_tables", "orc_nnp_tbl_descriptor",
                   "orc_nnp_tbl_set_dedg", "orc_nnp_tbl_force", "orc_mlp_set", "orc_mlp_eval",
                   "orc_sf_set", "orc_sf_descriptor", "orc_sf_force", "orc_sqd"):
            getattr(L, fn)
        _lib = L
    return _lib


def _p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def _f32(a):
    return np.ascontiguousarray(a, np.float32)


def _i32(a):
    return np.ascontiguousarray(a, np.int32)


def _f64(a):
    return np.ascontiguousarray(a, np.float64)


class Oracle:
    """One system in the CPU oracle.  Mirrors the call sequence of the reference MD loop."""

    def __init__(self, sysm, prec=0):
        self.L = lib()
        self.sys = sysm
        self.prec = prec
        self.h = C.c_void_p(self.L.orc_create(prec))
        offs, flat = sysm.special_flat()
        self._keep = [_f32(sysm.x), _f32(sysm.v), _f32(sysm.q), _f32(sysm.mass), _i32(sysm.typ),
                      _i32(sysm.mol_id), _i32(sysm.intra_id), _i32(sysm.mol_type),
                      _i32([m.n_atom for m in sysm.moltypes]), _i32(offs), _i32(flat)]
        k = self._keep
        self.L.orc_init(self.h, sysm.natom, sysm.ntype, sysm.nmol, len(sysm.moltypes),
                        C.c_float(sysm.bl), _p(k[0], C.c_float), _p(k[1], C.c_float),
                        _p(k[2], C.c_float), _p(k[3], C.c_float), _p(k[4], C.c_int),
                        _p(k[5], C.c_int), _p(k[6], C.c_int), _p(k[7], C.c_int),
                        _p(k[8], C.c_int), _p(k[9], C.c_int), _p(k[10], C.c_int))
        self.natom = sysm.natom
        self.gr_nbin = 0

    def __del__(self):
        try:
            self.L.orc_destroy(self.h)
        except Exception:
            pass

    # neighbor list ------------------------------------------------------------------------
    def neigh_init(self, pair_cutoff, skin, gr_cutoff):
        self.L.orc_neigh_init(self.h, C.c_float(pair_cutoff), C.c_float(skin), C.c_float(gr_cutoff))

    def rebuild(self, full=False):
        return self.L.orc_rebuild(self.h, int(full))

    def update_ghost_pos(self):
        self.L.orc_update_ghost_pos(self.h)

    def isvalid(self):
        return bool(self.L.orc_isvalid(self.h))

    def list_entries(self):
        n = self.L.orc_list_size(self.h)
        out = [np.zeros(n, np.int32) for _ in range(4)]
        self.L.orc_list_export(self.h, *[_p(o, C.c_int) for o in out])
        return np.stack(out, 1)

    def cell_ids(self):
        out = np.zeros(self.natom, np.int32)
        self.L.orc_get_cell_ids(self.h, _p(out, C.c_int))
        return out

    def xmol_codes(self):
        out = np.zeros(self.natom, np.int32)
        self.L.orc_get_xmol_codes(self.h, _p(out, C.c_int))
        return out

    # state --------------------------------------------------------------------------------
    def get(self, what):
        out = np.zeros((self.natom, 3), np.float64)
        self.L.orc_get(self.h, {"x": 0, "v": 1, "f": 2, "f_prev": 3}[what], _p(out, C.c_double))
        return out

    def set(self, what, arr):
        a = _f64(arr)
        self.L.orc_set(self.h, {"x": 0, "v": 1, "f": 2, "f_prev": 3}[what], _p(a, C.c_double))

    def clear_f(self):
        self.L.orc_clear_f(self.h)

    # forces -------------------------------------------------------------------------------
    def pair_set(self, cutoff, force_tbl, pot_tbl):
        self._tbl = (_f32(force_tbl), _f32(pot_tbl))
        self.L.orc_pair_set(self.h, C.c_float(cutoff), force_tbl.shape[1] - 1,
                            _p(self._tbl[0], C.c_float), _p(self._tbl[1], C.c_float))

    def pair_compute(self, ev=True):
        e, w = C.c_double(0), C.c_double(0)
        self.L.orc_pair_compute(self.h, int(ev), C.byref(e), C.byref(w))
        return e.value, w.value

    def bonded_compute(self, ev=True):
        s = self.sys
        e, w = C.c_double(0), C.c_double(0)
        bonds, angles, dih = s.expand("bond"), s.expand("angle"), s.expand("dihedral")
        if len(bonds):
            self.L.orc_bond_compute(self.h, len(bonds), _p(bonds, C.c_int), _p(_f32(s.bond_coef), C.c_float),
                                    int(ev), C.byref(e), C.byref(w))
        if len(angles):
            self.L.orc_angle_compute(self.h, len(angles), _p(angles, C.c_int), _p(_f32(s.angle_coef), C.c_float),
                                     int(ev), C.byref(e))
        rb = getattr(s, "dihedral_rb", None)
        if len(dih) and rb is not None:
            self.L.orc_dihedral_poly_compute(self.h, len(dih), _p(dih, C.c_int), _p(_f32(rb), C.c_float),
                                             int(ev), C.byref(e))
        elif len(dih):
            self.L.orc_dihedral_compute(self.h, len(dih), _p(dih, C.c_int), _p(_f32(s.dihedral_coef), C.c_float),
                                        int(ev), C.byref(e))
        return e.value, w.value

    # integrator ---------------------------------------------------------------------------
    def set_dt(self, dt):
        self.L.orc_set_dt(self.h, C.c_double(dt))

    def update_x(self):
        self.L.orc_update_x(self.h)

    def update_v(self):
        self.L.orc_update_v(self.h)

    def compute_ke(self):
        return self.L.orc_compute_ke(self.h)

    # g(r) / S(q) --------------------------------------------------------------------------
    def gr_init(self, nbin_inner, r_max, r_max_inner):
        self.gr_nbin = self.L.orc_gr_init(self.h, nbin_inner, C.c_float(r_max), C.c_float(r_max_inner))
        return self.gr_nbin

    def gr_tally(self, full):
        self.L.orc_gr_tally(self.h, int(full))

    def gr_get(self):
        npair = self.sys.npair
        counts = np.zeros((npair, self.gr_nbin), np.int64)
        h = np.zeros((npair, self.gr_nbin), np.float64)
        self.L.orc_gr_get(self.h, _p(counts, C.c_longlong), _p(h, C.c_double))
        return counts, h

    def sq(self, nq, q_min, q_max):
        out = np.zeros((self.sys.npair, nq), np.float64)
        self.L.orc_sq(self.h, nq, C.c_double(q_min), C.c_double(q_max), _p(out, C.c_double))
        return out

    # RMDF ---------------------------------------------------------------------------------
    def rmdf_init(self, data, q_max, strength, gamma, target=0):
        """data: [nsq][nq][3 + ntype] rows (q, S_exp, weight, form factors)."""
        self._rm = _f32(data)
        self.rm_shape = self._rm.shape
        self.L.orc_rmdf_init(self.h, self._rm.shape[0], self._rm.shape[1], C.c_double(q_max), _p(self._rm, C.c_float),
                             C.c_float(strength), C.c_float(gamma), target)

    def rmdf_step(self, q_max, potential=True):
        nsq, nq = self.rm_shape[0], self.rm_shape[1]
        sq = np.zeros((nsq, nq)); ds = np.zeros((nsq, nq)); fq = np.zeros((self.sys.npair, nq))
        e = self.L.orc_rmdf_step(self.h, C.c_double(q_max), int(potential), _p(sq, C.c_double), _p(ds, C.c_double),
                                 _p(fq, C.c_double))
        return e, sq, ds, fq

    def rmdf_tables(self, shape):
        f = np.zeros(shape); u = np.zeros(shape)
        self.L.orc_rmdf_tables(self.h, _p(f, C.c_double), _p(u, C.c_double))
        return f, u

    # NNP (reference tabulated form) -------------------------------------------------------
    def nnp_tbl_init(self, cutoff, nbin, coef_p, coef_w, coef_d, coef_r):
        self._nnp = [_f32(coef_p), _f32(coef_w), _f32(coef_d), _f32(coef_r)]
        self.L.orc_nnp_tbl_init(self.h, C.c_float(cutoff), nbin, *[_p(a, C.c_float) for a in self._nnp])
        self.nnp_nbin = nbin

    def nnp_tbl_tables(self):
        npair, nb1 = self.sys.npair, self.nnp_nbin + 1
        td = np.zeros((npair, 80, nb1)); tf = np.zeros((npair, 80, nb1))
        tp = np.zeros((npair, 4, nb1)); tfp = np.zeros((npair, 4, nb1))
        self.L.orc_nnp_tbl_tables(self.h, *[_p(a, C.c_double) for a in (td, tf, tp, tfp)])
        return td, tf, tp, tfp

    def nnp_tbl_descriptor(self):
        out = np.zeros((self.natom, 20), np.float64)
        self.L.orc_nnp_tbl_descriptor(self.h, _p(out, C.c_double))
        return out

    def nnp_tbl_force(self, dedg, ev=True):
        d = _f64(dedg)
        self.L.orc_nnp_tbl_set_dedg(self.h, _p(d, C.c_double))
        e, w = C.c_double(0), C.c_double(0)
        self.L.orc_nnp_tbl_force(self.h, int(ev), C.byref(e), C.byref(w))
        return e.value, w.value

    # MLP ---------------------------------------------------------------------------------
    def mlp_set(self, widths, params, out_scale=1.0, shift=None, scale=None):
        """shift, scale: [ntype][D] input standardisation x = (G - shift) * scale (None: identity)."""
        self._w = _i32(widths)
        self._par = _f32(params)
        self._sh = None if shift is None else _f32(shift)
        self._sc = None if scale is None else _f32(scale)
        self.L.orc_mlp_set(self.h, len(widths) - 1, _p(self._w, C.c_int), _p(self._par, C.c_float),
                           C.c_float(out_scale), None if self._sh is None else _p(self._sh, C.c_float),
                           None if self._sc is None else _p(self._sc, C.c_float))

    def mlp_eval(self, G):
        G = _f64(G)
        D = G.shape[1]
        dG = np.zeros_like(G)
        e = np.zeros(self.natom, np.float64)
        et = self.L.orc_mlp_eval(self.h, D, _p(G, C.c_double), _p(dG, C.c_double), _p(e, C.c_double))
        return et, e, dG

    # symmetry-function NNP ---------------------------------------------------------------
    def sf_set(self, rc, mu, eta_r, eta_a):
        self._sf = (_f32(mu), _f32(eta_a))
        self.sf_D = self.L.orc_sf_set(self.h, C.c_float(rc), len(mu), _p(self._sf[0], C.c_float),
                                      C.c_float(eta_r), len(eta_a), _p(self._sf[1], C.c_float))
        return self.sf_D

    def sf_descriptor(self):
        out = np.zeros((self.natom, self.sf_D), np.float64)
        self.L.orc_sf_descriptor(self.h, _p(out, C.c_double))
        return out

    def sf_force(self, dG, ev=True):
        d = _f64(dG)
        w = C.c_double(0)
        self.L.orc_sf_force(self.h, _p(d, C.c_double), int(ev), C.byref(w))
        return w.value


def sq_transform(h, dr, rho, nq, q_min, q_max):
    """The S(q) sum of SQ::compute on an arbitrary h(r) sampled at bin centres."""
    h = _f64(h)
    out = np.zeros(nq, np.float64)
    lib().orc_sq_transform(len(h), C.c_double(dr), C.c_double(rho), _p(h, C.c_double), nq, C.c_double(q_min),
                           C.c_double(q_max), _p(out, C.c_double))
    return out


def sqd(x, typ, ntype, bl, nq, q_max):
    """SQD (direct k-space S(q)) of positions x [natom][3] and elements typ: returns qq (FP64 mean
    |G| per bin), qq_ref (the reference's float mean), sq [nq][npair] (FP64) and count [nq]."""
    x = _f32(x).reshape(-1, 3)
    t = _i32(typ)
    npair = ntype * (ntype + 1) // 2
    qq = np.zeros(nq)
    qq_ref = np.zeros(nq)
    sq = np.zeros((nq, npair))
    cnt = np.zeros(nq, np.int64)
    lib().orc_sqd(len(t), ntype, C.c_float(bl), _p(x, C.c_float), _p(t, C.c_int), nq, C.c_float(q_max),
                  _p(qq, C.c_double), _p(qq_ref, C.c_double), _p(sq, C.c_double), _p(cnt, C.c_longlong))
    return qq, qq_ref, sq, cnt


def canonical(entries):
    """Directed list entries (i, j, code, special) -> sorted unique int64 keys.  Real-real entries
    (code 13) are expanded to both directions so a half list and a full list with the same
    interacting pairs give the same key set."""
    e = np.asarray(entries, np.int64)
    i, j, c, s = e[:, 0], e[:, 1], e[:, 2], e[:, 3]
    real = c == 13
    ii = np.concatenate([i, j[real]])
    jj = np.concatenate([j, i[real]])
    cc = np.concatenate([c, c[real]])
    ss = np.concatenate([s, s[real]])
    key = ((ii * (1 << 24) + jj) * 32 + cc) * 4 + ss
    return np.unique(key)
