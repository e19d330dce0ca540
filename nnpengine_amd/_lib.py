"""ctypes binding of libnnpe.so (include/nnpe.h).  The product path has no fallback: if the
library is missing or no device is present, calls raise."""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnnpe.so")

STATUS = {0: "ok", 1: "bad argument", 2: "HIP runtime error", 3: "call out of order",
          4: "nei_list overflow", 5: "ghost overflow", 6: "atom outside the cell grid",
          7: "no HIP device", 8: "unsupported configuration"}

NNPE_X, NNPE_V, NNPE_F, NNPE_FPREV = 0, 1, 2, 3
TERMS = ("pair", "bond", "angle", "dihedral", "nnp_pair", "nn", "sf", "ke", "rmdf")
NTERM = 9

# (name, restype, argtypes) of every entry point declared in include/nnpe.h
_P, _I, _F, _D = C.c_void_p, C.c_int, C.c_float, C.c_double
_PF, _PI, _PD, _PLL = C.POINTER(C.c_float), C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(C.c_longlong)
SIGNATURES = [
    ("nnpe_status_string", C.c_char_p, [_I]),
    ("nnpe_last_error", C.c_char_p, [_P]),
    ("nnpe_create", _I, [C.POINTER(_P), _I]),
    ("nnpe_destroy", _I, [_P]),
    ("nnpe_synchronize", _I, [_P]),
    ("nnpe_system_set", _I, [_P, _I, _I, _F, _PF, _PF, _PF, _PF, _PI, _PI, _PI, _I, _PI, _I, _PI, _PI, _PI]),
    ("nnpe_state_get", _I, [_P, _I, _PD]),
    ("nnpe_state_set", _I, [_P, _I, _PD]),
    ("nnpe_neigh_init", _I, [_P, _F, _F, _F, _F]),
    ("nnpe_neigh_build", _I, [_P]),
    ("nnpe_neigh_isvalid", _I, [_P, _PI]),
    ("nnpe_neigh_size", _I, [_P, _I, _PLL, _PI]),
    ("nnpe_neigh_export", _I, [_P, _I, _PI, _PI, _PI, _PI, C.c_longlong, _PLL]),
    ("nnpe_pair_table", _I, [_P, _F, _I, _PF, _PF]),
    ("nnpe_pair_compute", _I, [_P, _I]),
    ("nnpe_bonded_set", _I, [_P, _I, _I, _PF, _I, _PI]),
    ("nnpe_bonded_compute", _I, [_P, _I]),
    ("nnpe_bonded_compute_noalloc", _I, [_P, _I]),
    ("nnpe_energy_get", _I, [_P, _PD]),
    ("nnpe_energy_clear", _I, [_P]),
    ("nnpe_verlet_set", _I, [_P, _D, _I]),
    ("nnpe_verlet_update_x", _I, [_P]),
    ("nnpe_verlet_update_v", _I, [_P]),
    ("nnpe_verlet_compute_ke", _I, [_P, _PD]),
    ("nnpe_verlet_scale_v", _I, [_P, _D]),
    ("nnpe_clear_force", _I, [_P]),
    ("nnpe_gr_init", _I, [_P, _I, _F, _F, _PI]),
    ("nnpe_gr_tally", _I, [_P, _I]),
    ("nnpe_gr_get", _I, [_P, _PLL, _PD]),
    ("nnpe_sq_init", _I, [_P, _I, _D, _D]),
    ("nnpe_sq_compute", _I, [_P, _PD]),
    ("nnpe_sqd_init", _I, [_P, _I, _F, _PLL]),
    ("nnpe_sqd_compute", _I, [_P]),
    ("nnpe_sqd_get", _I, [_P, _PD, _PD, _PLL]),
    ("nnpe_rmdf_init", _I, [_P, _I, _PF, _F, _F, _I]),
    ("nnpe_rmdf_set_delta", _I, [_P, _PF]),
    ("nnpe_rmdf_step", _I, [_P, _I]),
    ("nnpe_rmdf_get", _I, [_P, _PF, _PF, _PF]),
    ("nnpe_rmdf_tables", _I, [_P, _PF, _PF]),
    ("nnpe_nnp_tbl_set", _I, [_P, _F, _I, _PF, _PF, _PF, _PF]),
    ("nnpe_nnp_tbl_descriptor", _I, [_P]),
    ("nnpe_nnp_tbl_force", _I, [_P, _I]),
    ("nnpe_sf_set", _I, [_P, _F, _I, _PF, _F, _I, _PF]),
    ("nnpe_sf_descriptor", _I, [_P]),
    ("nnpe_sf_force", _I, [_P, _I]),
    ("nnpe_nnp_get", _I, [_P, _I, _PF, _PI]),
    ("nnpe_nnp_set_dedg", _I, [_P, _PF, _I]),
    ("nnpe_mlp_set", _I, [_P, _I, _PI, _PF, _F, _I, _PF, _PF]),
    ("nnpe_mlp_eval", _I, [_P, _I]),
    ("nnpe_md_run", _I, [_P, _P, _I, _PD, _I, _PI]),
    ("nnpe_md_stats", _I, [_P, _PLL]),
    ("nnpe_md_reset_step", _I, [_P]),
    ("nnpe_md_set_step", _I, [_P, C.c_longlong]),
    ("nnpe_profile_stage", _I, [_P, _I, _I, _PD]),
    ("nnpe_serve_create", _I, [_P, C.POINTER(_P)]),
    ("nnpe_serve_destroy", _I, [_P]),
    ("nnpe_serve_submit", _I, [_P, _I, _F, _PF, _PF, _PF, _PF, _PI, _PI, _PI, _I, _PI, _I, _PI, _I, _PI, _I, _PI, _I,
                               _PLL]),
    ("nnpe_serve_advance", _I, [_P, _PLL, _I, _PI]),
    ("nnpe_serve_result", _I, [_P, C.c_longlong, _PD, _PD]),
    ("nnpe_serve_stats", _I, [_P, _PLL]),
    ("nnpe_serve_last_error", C.c_char_p, [_P]),
]

STAGES = {"pair": 0, "bonded": 1, "sf_desc": 2, "mlp": 3, "sf_adjoint": 4, "sf_force": 5, "rebuild": 6,
          "gr_full": 7, "gr_list": 8, "verlet": 9, "sq": 10, "tbl_desc": 11, "tbl_force": 12,
          "copy": 13, "rebuild_gr": 14, "mfma": 15}


class MdParams(C.Structure):
    """struct nnpe_md_params (include/nnpe.h)."""
    _fields_ = [("dt", C.c_double), ("fp64", C.c_int), ("rebuild_every", C.c_int), ("pair", C.c_int),
                ("bonded", C.c_int), ("nnp", C.c_int), ("gr_every", C.c_int), ("gr_long_every", C.c_int),
                ("sq", C.c_int), ("log_every", C.c_int), ("use_graph", C.c_int), ("rmdf", C.c_int)]



class ServeConfig(C.Structure):
    """struct nnpe_serve_config (include/nnpe.h)."""
    _fields_ = [("device", C.c_int), ("max_atoms", C.c_int), ("max_boxes", C.c_int), ("ntype", C.c_int),
                ("pair_cutoff", C.c_float), ("skin", C.c_float), ("nnp_cutoff", C.c_float),
                ("rebuild_every", C.c_int), ("dt", C.c_double), ("fp64", C.c_int),
                ("pair_nbin", C.c_int), ("pair_force", _PF), ("pair_potential", _PF),
                ("nbond_types", C.c_int), ("nangle_types", C.c_int), ("ndihedral_types", C.c_int),
                ("bond_coef", _PF), ("angle_coef", _PF), ("dihedral_coef", _PF),
                ("nmoltype", C.c_int), ("mt_natom", _PI), ("mt_special_off", _PI), ("mt_special", _PI),
                ("sf_mu", _PF), ("sf_eta_r", C.c_float), ("sf_eta", _PF),
                ("nlayer", C.c_int), ("widths", _PI), ("params", _PF), ("out_scale", C.c_float),
                ("precision", C.c_int), ("in_shift", _PF), ("in_scale", _PF)]


_lib = None


class NnpeError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"nnpe error {code} ({STATUS.get(code, '?')}): {msg}")
        self.code = code


def load():
    """Load libnnpe.so (built by nnpengine_amd.build / __graft_entry__.build)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run `python -m nnpengine_amd.build` first "
                               "(there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def declared_symbols():
    """Entry points declared in include/nnpe.h (parsed from the header)."""
    import re
    with open(os.path.join(HERE, "..", "include", "nnpe.h")) as fh:
        txt = fh.read()
    return sorted(set(re.findall(r"\b(nnpe_[a-z0-9_]+)\s*\(", txt)))
