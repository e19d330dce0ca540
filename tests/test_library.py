"""The C-ABI library loads and exports every entry point include/nnpe.h declares (no device
calls), and the Python binding declares a signature for each of them."""
import ctypes
import os

from nnpengine_amd import _lib


def test_library_exports_every_declared_symbol():
    assert os.path.exists(_lib.LIB_PATH), "run __graft_entry__.build() first"
    L = ctypes.CDLL(_lib.LIB_PATH)
    declared = _lib.declared_symbols()
    assert len(declared) >= 40
    missing = [s for s in declared if not hasattr(L, s)]
    assert not missing, missing
    bound = {name for name, _, _ in _lib.SIGNATURES}
    assert set(declared) == bound, sorted(set(declared) ^ bound)


def test_binding_struct_matches_header():
    import re
    with open(os.path.join(os.path.dirname(_lib.LIB_PATH), "..", "include", "nnpe.h")) as fh:
        txt = fh.read()
    body = re.search(r"typedef struct nnpe_md_params \{(.*?)\} nnpe_md_params;", txt, re.S).group(1)
    fields = re.findall(r"\b(?:double|int)\s+(\w+);", body)
    assert fields == [f for f, _ in _lib.MdParams._fields_]


def test_serve_config_struct_matches_header():
    import re
    with open(os.path.join(os.path.dirname(_lib.LIB_PATH), "..", "include", "nnpe.h")) as fh:
        txt = fh.read()
    body = re.search(r"typedef struct nnpe_serve_config \{(.*?)\} nnpe_serve_config;", txt, re.S).group(1)
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    names = []
    for decl in body.split(";"):
        decl = decl.strip()
        if not decl:
            continue
        names += [n.strip().lstrip("*").strip() for n in re.sub(r"^(const\s+)?\w+\s*\*?", "", decl).split(",")]
    assert names == [f for f, _ in _lib.ServeConfig._fields_]
