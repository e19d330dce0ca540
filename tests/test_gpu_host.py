"""The C++ drop-in driver (nnpengine_amd/nnpe_run, src/run.cpp's classical flow over the C
ABI) on the reference's input-file format: the step-0 energy and virial it logs match the FP64
oracle, the run completes its steps with rebuilds, and the g(r) dump has the reference's
running-average layout.  With experimental S(q) data the RMDF refinement runs inside the loop
(the step-0 Q-space energy it logs matches the oracle's, the refined table is dumped, the sqx
dump has the reference's columns), and the end-of-run SQD dump matches the oracle's SQD of the
dumped final positions."""
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


def test_classical_driver_rmdf_sqd(tmp_path, water_c0):
    from tests.test_gpu_rmdf import _data
    sysm = water_c0
    rc, skin, nbin, qmax, nq = 9.0, 0.5, 450, 12.0, 240
    ft, ut = S.lj_table(sysm, rc, nbin, S.WATER_LJ)
    data = _data(sysm, 5, 1)[0]
    assert data.shape == (nq, 3 + sysm.ntype)
    expf = str(tmp_path / "sexp.txt")
    np.savetxt(expf, data, fmt="%.9e")
    out = {k: str(tmp_path / f"{k}.txt") for k in ("xyz", "sqd", "sqx", "force")}
    sqd_q, sqd_n = 6.0, 30
    path = S.write_classical_input(str(tmp_path), sysm, rc, skin, rc, ft, ut, nbin, nq, qmax,
                                   steps=(20, 10, 10, 10, 10), outputs=out,
                                   rmdf=(2.0, 1.0, [expf], str(tmp_path / "none.txt")), sqd=(sqd_q, sqd_n))
    env = dict(os.environ, NNPE_DT="0.5", NNPE_SEED="7")
    print(f"[host] running {EXE} classical {path} (RMDF + SQD)", flush=True)
    res = subprocess.run([EXE, "classical", path], capture_output=True, text=True, env=env, timeout=120)
    assert res.returncode == 0, res.stderr[-3000:]
    rows = [ln.split() for ln in res.stdout.splitlines() if ln.strip()]
    assert [int(r[0]) for r in rows] == [0, 10, 20]
    # step-0 Q-space energy against the oracle in the reference's FP32 arithmetic
    o = pyoracle.Oracle(sysm, 0)
    o.neigh_init(rc, skin, rc)
    o.rebuild(False)
    o.pair_set(rc, ft, ut)
    o.gr_init(nbin, rc, rc)
    o.gr_tally(True)
    o.rmdf_init(data[None], qmax, 2.0, 1.0, target=0)
    erg, _, _, _ = o.rmdf_step(qmax, potential=True)
    assert float(rows[0][4]) == pytest.approx(erg, rel=1e-5)
    # refined table written, and differs from the base table
    f_ref = np.loadtxt(out["force"])
    assert f_ref.shape == (nbin, 1 + 2 * sysm.npair)
    assert np.abs(f_ref[:, 1] - ft[0, :nbin]).max() > 1e-6 * np.abs(ft[0]).max()
    with open(out["sqx"]) as fh:
        assert fh.readline().startswith("# 1")
    sqx = np.loadtxt(out["sqx"], comments="#")
    assert sqx.shape == (nq, 1 + 3)
    # SQD of the final positions
    with open(out["xyz"]) as fh:
        n = int(fh.readline())
        fh.readline()
        xyz = np.array([[float(v) for v in fh.readline().split()[1:4]] for _ in range(n)], np.float32)
    qq, _, osq, cnt = pyoracle.sqd(xyz, sysm.typ, sysm.ntype, sysm.bl, sqd_n, sqd_q)
    got = np.loadtxt(out["sqd"], comments="#")
    nz = cnt > 0
    assert got.shape == (int(nz.sum()), 1 + sysm.npair)
    np.testing.assert_allclose(got[:, 0], qq[nz], rtol=1e-12)
    assert np.abs(got[:, 1:] - osq[nz]).max() < 2e-5
