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


def _run(exe_mode, path, seed=7, timeout=180):
    env = dict(os.environ, NNPE_DT="0.5", NNPE_SEED=str(seed))
    res = subprocess.run([EXE, exe_mode, path], capture_output=True, text=True, env=env, timeout=timeout)
    assert res.returncode == 0, res.stderr[-3000:]
    return [ln.split() for ln in res.stdout.splitlines() if ln.strip()]


def _read_dump(path):
    with open(path) as fh:
        head = fh.readline()
    return int(head.split()[1]), np.loadtxt(path, comments="#")


def test_gr_dump_running_average(tmp_path, water_c0):
    """GR::dump with append (src/grsq.cpp:221-256): the file keeps a running average over runs,
    new = current / n + (1 - 1/n) old with n = min(previous n + 1, 100) in its "# n" header."""
    sysm = water_c0
    rc, skin = 9.0, 0.5
    ft, ut = S.lj_table(sysm, rc, 900, S.WATER_LJ)
    runs = {}
    for name, steps, seed in (("a", 20, 7), ("b", 10, 8)):
        d = tmp_path / name
        d.mkdir()
        p = S.write_classical_input(str(d), sysm, rc, skin, rc, ft, ut, 450, 200, 20.0, steps=(steps, 10, 10, 10, 10),
                                    outputs={"gr": (str(d / "gr.txt"), 0)})
        _run("classical", p, seed)
        runs[name] = (p, _read_dump(str(d / "gr.txt")))
    (na, ga), (nb, gb) = runs["a"][1], runs["b"][1]
    assert na == nb == 1
    assert np.abs(ga[:, 1:] - gb[:, 1:]).max() > 0
    acc = str(tmp_path / "acc.txt")
    import shutil
    shutil.copy(str(tmp_path / "a" / "gr.txt"), acc)
    # run b again, appending to a copy of run a's dump
    d = tmp_path / "b2"
    d.mkdir()
    p = S.write_classical_input(str(d), sysm, rc, skin, rc, ft, ut, 450, 200, 20.0, steps=(10, 10, 10, 10, 10),
                                outputs={"gr": (acc, 1)})
    _run("classical", p, 8)
    n2, g2 = _read_dump(acc)
    assert n2 == 2
    np.testing.assert_array_equal(g2[:, 0], ga[:, 0])
    np.testing.assert_allclose(g2[:, 1:], 0.5 * gb[:, 1:] + 0.5 * ga[:, 1:], rtol=1e-6, atol=1e-6)
    # the count stops at 100
    txt = open(acc).read().split("\n", 1)[1]
    with open(acc, "w") as fh:
        fh.write("# 100\n" + txt)
    _run("classical", p, 8)
    n3, g3 = _read_dump(acc)
    assert n3 == 100
    np.testing.assert_allclose(g3[:, 1:], gb[:, 1:] / 100 + 0.99 * g2[:, 1:], rtol=1e-6, atol=1e-6)


def test_nnp_run_driver(tmp_path, water_c0):
    """nnp_run (src/run.cpp:968-1403) on the reference's input format with the NNPOTENTIAL_TBL
    weight file (src/pair_nnp.cpp:433-598) and a network file in place of the TF SavedModel: the
    step-0 energy it logs (NNP pair tables + bonded + network) matches the FP64 oracle, and the
    run completes its steps with the g(r) dump written."""
    from nnpengine_amd import model as M
    sysm = water_c0
    rc, skin, nbin = 6.0, 0.5, 600
    cp, cw, cd, cr = M.tbl_coefficients(sysm.npair, 5, rc)
    # a water-like pair part that keeps 20 steps of MD sane: O-O Lennard-Jones (SPC/E), small
    # charge terms between molecules, no intramolecular (special-class) pair terms; the
    # descriptor tables stay seeded
    cp = np.zeros_like(cp)
    cp[0, 0, :2] = (0.155, 3.166)
    cp[:, 0, 5] = (0.3, -0.15, 0.075)
    coef = (cp, cw, cd, cr)
    widths = [20, 32, 32, 1]
    par = M.mlp_init(widths, sysm.ntype, 9)
    grf = str(tmp_path / "gr.txt")
    path = S.write_nnp_input(str(tmp_path), sysm, rc, skin, rc, nbin, coef, widths, par, 0.7, "fp32", 300, 120, 12.0,
                             steps=(20, 10, 10, 10, 10), outputs={"gr": grf})
    print(f"[host] running {EXE} nnp_run {path}", flush=True)
    rows = _run("nnp_run", path)
    assert [int(r[0]) for r in rows] == [0, 10, 20]
    erg0 = float(rows[0][3])
    o = pyoracle.Oracle(sysm, 1)
    o.neigh_init(rc, skin, rc)
    o.rebuild(False)
    o.nnp_tbl_init(rc, nbin, *coef)
    G = o.nnp_tbl_descriptor()
    o.mlp_set(widths, par, 0.7)
    et, _, dG = o.mlp_eval(G)
    o.clear_f()
    ep, _ = o.nnp_tbl_force(dG)
    eb, _ = o.bonded_compute()
    ref = et + ep + eb
    print(f"[host] nnp_run step-0 energy {erg0:.7g} vs FP64 oracle {ref:.7g}", flush=True)
    assert abs(erg0 - ref) < 1e-5 * abs(ref) + 1e-3
    n, g = _read_dump(grf)
    assert n == 1 and g.shape == (300, 1 + sysm.npair)
    assert np.all(np.isfinite(g))


def test_initv_andersen(tmp_path, water_c0):
    """ENVIRON::initV (src/environ.cpp:190-205): Maxwell-Boltzmann velocities at the start
    temperature (the logged step-0 temperature within 5 standard deviations of its sampling
    spread, at two temperatures), and the Andersen re-draw of a tenth of the atoms every
    scale_interval steps (src/run.cpp:349-373) holds a constant set point and pulls the
    temperature up a ramp to twice the start."""
    sysm = water_c0
    rc, skin = 9.0, 0.5
    ft, ut = S.lj_table(sysm, rc, 900, S.WATER_LJ)
    sigma_rel = np.sqrt(2.0 / (3 * sysm.natom))
    for T in (300.0, 450.0):
        d = tmp_path / f"t{int(T)}"
        d.mkdir()
        p = S.write_classical_input(str(d), sysm, rc, skin, rc, ft, ut, 450, 200, 20.0, temp=(T, T),
                                    steps=(10, 10, 10, 10, 10))
        rows = _run("classical", p, 11)
        t0 = float(rows[0][1])
        print(f"[host] initV at {T} K: step-0 temperature {t0:.1f} K (spread {T * sigma_rel:.1f} K)", flush=True)
        assert abs(t0 - T) < 5 * T * sigma_rel
    finals = {}
    for name, T1, seed in (("hold", 300.0, 12), ("ramp", 600.0, 13)):
        d = tmp_path / name
        d.mkdir()
        p = S.write_classical_input(str(d), sysm, rc, skin, rc, ft, ut, 450, 200, 20.0, temp=(300.0, T1),
                                    steps=(400, 100, 10, 100, 100))
        rows = _run("classical", p, seed)
        temps = [float(r[1]) for r in rows]
        print(f"[host] Andersen 300 -> {T1:.0f} K: {['%.0f' % t for t in temps]}", flush=True)
        assert abs(temps[0] - 300) < 5 * 300 * sigma_rel
        finals[name] = temps
    # the lattice start heats the liquid (to ~500 K by step 100); the held run then relaxes back
    # towards its set point step by step, and the ramp ends well above it
    hold = finals["hold"]
    assert all(a > b for a, b in zip(hold[1:], hold[2:])) and hold[-1] < hold[1] - 100
    assert finals["ramp"][-1] > hold[-1] + 100
