"""The butane refinement workload at its full size on the GPU: 4000 united-atom chains (16 000
sites, bonds, angles and the Ryckaert-Bellemans torsion) in an 87.3 A box.

* list (11 A) bit-exact against the FP32 oracle;
* pair + bond + angle + dihedral forces, energy and virial against the FP64 oracle (1e-5);
* g(r) counts bit-exact (full tally to 12 A), then one RMDF step (S(q) -> residual -> table
  rewrite, src/grsq.cpp:1079-1377) against the FP32 oracle on the same counts;
* the MD step of the workload (BF16 NNP, g(r) + S(q) + RMDF every 10 steps, table rewrites on
  logged steps): captured graphs equal eager launches bit for bit over 40 steps.
"""
import numpy as np
import pytest

from nnpengine_amd import workloads as W
from oracle import pyoracle
from tests.conftest import rel_rms

pytestmark = pytest.mark.gpu

C3 = W.BUTANE_C3


@pytest.fixture(scope="module")
def c3():
    sysm = W.butane_c3_system()
    ft, ut = W.butane_tables(sysm)
    return sysm, ft, ut


def _oracle(sysm, prec):
    o = pyoracle.Oracle(sysm, prec)
    o.neigh_init(C3["rc"], C3["skin"], C3["gr_rmax"])
    return o


def test_c3_lists_and_forces(c3):
    sysm, ft, ut = c3
    print(f"[c3] natom={sysm.natom} L={sysm.bl}", flush=True)
    e = W.classical_engine(sysm, ft, ut)
    e.build()
    o32 = _oracle(sysm, 0)
    ng = o32.rebuild(False)
    assert np.array_equal(pyoracle.canonical(e.list_entries()), pyoracle.canonical(o32.list_entries()))
    assert e.list_size()[1] == ng
    e.energy_clear()
    e.pair_compute(True)
    e.bonded_compute(True)
    fg = e.get("f")
    en = e.energy()
    o64 = _oracle(sysm, 1)
    o64.rebuild(False)
    o64.pair_set(C3["rc"], ft, ut)
    o64.clear_f()
    eb = o64.bonded_compute()
    ep = o64.pair_compute()
    f64 = o64.get("f")
    err = rel_rms(fg, f64)
    eg = sum(en[t][0] for t in ("pair", "bond", "angle", "dihedral"))
    wg = en["pair"][1] + en["bond"][1]
    print(f"[c3] force rel rms vs fp64 {err:.2e}; energy {eg:.6f} vs {ep[0] + eb[0]:.6f}; "
          f"virial {wg:.4f} vs {ep[1] + eb[1]:.4f}", flush=True)
    assert err < 1e-5
    assert abs(eg - (ep[0] + eb[0])) < 1e-5 * abs(ep[0] + eb[0])
    assert abs(wg - (ep[1] + eb[1])) < 1e-5 * abs(ep[1] + eb[1])
    assert abs(en["dihedral"][0]) > 0
    e.close()


def test_c3_gr_and_refinement_step(c3):
    sysm, ft, ut = c3
    e = W.classical_engine(sysm, ft, ut)
    e.build()
    o = _oracle(sysm, 0)
    o.rebuild(False)
    o.pair_set(C3["rc"], ft, ut)
    assert o.gr_init(C3["nbin"], C3["gr_rmax"], C3["rc"]) == e.gr_nbin
    e.gr_tally(full=True)
    o.gr_tally(True)
    cg, _ = e.gr_get()
    co, _ = o.gr_get()
    print(f"[c3] g(r) total counts {int(cg.sum())}", flush=True)
    assert np.array_equal(cg, co)
    # a target S(q) of the right shape: the present total S(q) damped towards 1 (the refinement
    # step is checked, not the physics of the target)
    s_now = W.total_sq(e, sysm)
    q = (np.arange(C3["nq"]) + 0.5) * C3["q_max"] / C3["nq"]
    s_exp = 1.0 + (s_now - 1.0) * (1.0 - 0.3 * np.exp(-0.2 * q))
    data = W.rmdf_data(s_exp, sysm.ntype, C3["q_max"])
    e.rmdf_init(data, 1.0, 0.5, target=0)
    e.rmdf_step(potential=True)
    o.rmdf_init(data, C3["q_max"], 1.0, 0.5, target=0)
    erg, sq_o, ds_o, fq_o = o.rmdf_step(C3["q_max"], potential=True)
    sq_g, ds_g, fq_g = e.rmdf_get()

    def scale_err(a, b):
        return np.max(np.abs(a - b)) / max(1e-30, np.max(np.abs(b)))

    print(f"[c3] rmdf: S(q) {scale_err(sq_g, sq_o):.2e} residual {scale_err(ds_g, ds_o):.2e} force_q {scale_err(fq_g, fq_o):.2e}",
          flush=True)
    assert scale_err(sq_g, sq_o) < 1e-5
    assert scale_err(ds_g, ds_o) < 1e-5
    assert scale_err(fq_g, fq_o) < 1e-5
    assert e.energy()["rmdf"][0] == pytest.approx(erg, rel=1e-5)
    nb = C3["nbin"]
    f_g, u_g = e.rmdf_tables(nb)
    f_o, u_o = o.rmdf_tables((sysm.npair, nb + 1))
    print(f"[c3] refined tables rel rms: force {rel_rms(f_g[:, :nb], f_o[:, :nb]):.2e} potential {rel_rms(u_g[:, :nb], u_o[:, :nb]):.2e}",
          flush=True)
    assert rel_rms(f_g[:, :nb], f_o[:, :nb]) < 1e-5
    assert rel_rms(u_g[:, :nb], u_o[:, :nb]) < 1e-5
    # the rewrite moved the tables
    assert np.abs(f_g[:, 150:nb] - ft[:, 150:nb]).max() > 0
    e.close()


def test_c3_md_step_graph_equals_eager(c3):
    sysm, ft, ut = c3
    runs = []
    for use_graph in (True, False):
        e = W.classical_engine(sysm, ft, ut, nnp=True)
        W.nnp_attach(e, sysm, "bf16x1")
        e.gr_tally(full=True)
        s_now = W.total_sq(e, sysm)
        e.rmdf_init(W.rmdf_data(0.98 * s_now + 0.02, sysm.ntype, C3["q_max"]), 1.0, 0.5, target=0)
        log = e.md_run(40, dt=C3["dt"], fp64=True, rebuild_every=10, pair=True, bonded=True, nnp=2, gr_every=10,
                       gr_long_every=10, sq=True, rmdf=True, log_every=20, use_graph=use_graph)
        runs.append((e.get("x"), e.get("v"), log))
        e.close()
    print(f"[c3] md 40 steps: log rows {len(runs[0][2])}", flush=True)
    assert np.all(np.isfinite(runs[0][2]))
    assert np.array_equal(runs[0][0], runs[1][0]) and np.array_equal(runs[0][1], runs[1][1])
    assert np.array_equal(runs[0][2], runs[1][2])
