"""GPU parity of the NNP path: symmetry-function descriptors and forces, the MFMA MLP in FP32 /
BF16 / FP16, the reference's tabulated descriptor form, and the graph-captured MD loop.
Tolerances (relative RMS vs the FP64 oracle): FP32 forms 1e-5; BF16 NNP forces 1e-2 (both the
split hi + lo form "bf16" and the one-term form "bf16x1"); FP16 5e-3."""
import numpy as np
import pytest

from nnpengine_amd import model as M
from nnpengine_amd import system as S
from oracle import pyoracle
from tests.conftest import rel_rms

pytestmark = pytest.mark.gpu

# the 768-atom box (L = 19.73 A) needs skin + rc < L/2 = 9.865 A
RC_PAIR, SKIN, RC_NNP = 9.0, 0.5, 6.0
MD_SKIN = 0.8


def _setup(sysm, widths, seed=11, out_scale=0.5):
    """Engine + FP64 oracle on the same system with the same (standardised) network."""
    from nnpengine_amd.engine import Engine
    sfp = M.sf_params(RC_NNP)
    o = pyoracle.Oracle(sysm, 1)
    o.neigh_init(RC_PAIR, SKIN, RC_PAIR)
    o.rebuild(True)
    D = o.sf_set(sfp["rc"], sfp["mu"], sfp["eta_r"], sfp["eta_a"])
    G64 = o.sf_descriptor()
    norm = M.standardisation(G64, sysm.typ, sysm.ntype)
    par = M.mlp_init(widths, sysm.ntype, seed)
    o.mlp_set(widths, par, out_scale, *norm)
    e = Engine(sysm)
    e.neigh_init(RC_PAIR, SKIN, RC_PAIR, RC_NNP)
    e.build()
    e.sf_set(sfp["rc"], sfp["mu"], sfp["eta_r"], sfp["eta_a"])
    assert D == M.sf_dim(sysm.ntype)
    return e, o, par, G64, norm


PRECS = [("fp32", 1e-5, 1e-5), ("bf16", 2e-2, 3e-2), ("fp16", 5e-3, 5e-3), ("bf16x1", 2e-2, 3e-2)]


# hidden widths all 128: the uniform-width kernel of the 16-bit forms; 64: the general kernel
@pytest.mark.parametrize("widths", [[80, 64, 64, 1], [80, 128, 128, 1]], ids=["w64", "w128"])
@pytest.mark.parametrize("prec,tol_e,tol_g", PRECS)
def test_mlp_matches_fp64(water_c0, prec, tol_e, tol_g, widths):
    e, o, par, G64, norm = _setup(water_c0, widths)
    e.sf_descriptor()
    G = e.nnp_get("G").astype(np.float64)
    print(f"[mlp] prec={prec} descriptor rel-rms vs fp64 = {rel_rms(G, G64):.2e}", flush=True)
    assert rel_rms(G, G64) < 1e-5
    e.mlp_set(widths, par, 0.5, prec, *norm)
    e.energy_clear()
    e.mlp_eval(True)
    ea = e.nnp_get("e_atom").astype(np.float64)
    dG = e.nnp_get("dG").astype(np.float64)
    et, ea64, dG64 = o.mlp_eval(G)
    print(f"[mlp] prec={prec} E rel-rms {rel_rms(ea, ea64):.2e} dG rel-rms {rel_rms(dG, dG64):.2e}")
    assert rel_rms(ea, ea64) < tol_e
    assert rel_rms(dG, dG64) < tol_g
    en = e.energy()
    assert abs(en["nn"][0] - et) < max(tol_e, 1e-6) * (np.abs(ea64).sum() + 1e-9) * 10


@pytest.mark.parametrize("widths", [[80, 128, 96, 32, 1], [80, 128, 128, 128, 1]], ids=["mixed", "w128x3"])
@pytest.mark.parametrize("prec,tol_e,tol_g", PRECS)
def test_mlp_ragged_tiles_mixed_widths(prec, tol_e, tol_g, widths):
    """300 O and 600 H atoms: the last 32-row tile of each element is partial (12 and 24 rows);
    hidden widths 128, 96, 32 leave one, then three, of the four column blocks idle in the
    forward and backward passes, and 3 x 128 is the uniform-width kernel (16-bit forms) on
    ragged tiles."""
    sysm = S.water_box(300, 21.0, seed=3)
    e, o, par, G64, norm = _setup(sysm, widths, seed=5)
    e.sf_descriptor()
    G = e.nnp_get("G").astype(np.float64)
    e.mlp_set(widths, par, 0.5, prec, *norm)
    e.energy_clear()
    e.mlp_eval(True)
    ea = e.nnp_get("e_atom").astype(np.float64)
    dG = e.nnp_get("dG").astype(np.float64)
    _, ea64, dG64 = o.mlp_eval(G)
    print(f"[mlp] ragged prec={prec} E rel-rms {rel_rms(ea, ea64):.2e} dG rel-rms {rel_rms(dG, dG64):.2e}", flush=True)
    assert np.all(np.isfinite(ea)) and np.all(np.isfinite(dG))
    assert rel_rms(ea, ea64) < tol_e
    assert rel_rms(dG, dG64) < tol_g


@pytest.mark.parametrize("case", ["c0", "ragged"])
def test_sf_forces_fp32(case, water_c0):
    """FP32 NNP forces, energy and virial vs FP64; "ragged" is 900 atoms, so the last 16-lane
    group of the descriptor and force passes is partial."""
    sysm = water_c0 if case == "c0" else S.water_box(300, 21.0, seed=3)
    widths = [80, 64, 64, 1]
    e, o, par, G64, norm = _setup(sysm, widths)
    e.mlp_set(widths, par, 0.5, "fp32", *norm)
    e.sf_descriptor()
    e.clear_force()
    e.energy_clear()
    e.mlp_eval(True)
    e.sf_force(True)
    fg = e.get("f")
    en = e.energy()
    et, ea64, dG64 = o.mlp_eval(G64)
    o.clear_f()
    w64 = o.sf_force(dG64)
    f64 = o.get("f")
    print(f"[sf] {case} force rel-rms {rel_rms(fg, f64):.2e} virial {en['sf'][1]:.6f} vs {w64:.6f}")
    assert rel_rms(fg, f64) < 1e-5
    assert abs(en["sf"][1] - w64) < 1e-4 * abs(w64) + 1e-6
    # the total is a sum of per-atom energies of both signs: bound it by their magnitude
    assert abs(en["nn"][0] - et) < 1e-5 * np.abs(ea64).sum()
    # Newton's third law: the NNP forces sum to zero
    assert np.abs(fg.sum(0)).max() < 1e-3 * np.abs(fg).max()


def test_sf_forces_bf16_error(water_c1):
    """BF16 MFMA network: NNP force error vs FP64 within 1e-2 relative RMS (24k atoms, 3x128)."""
    widths = [80, 128, 128, 128, 1]
    e, o, par, G64, norm = _setup(water_c1, widths)
    _, _, dG64 = o.mlp_eval(G64)
    o.clear_f()
    o.sf_force(dG64)
    f64 = o.get("f")
    # force kernel alone on the oracle's dE/dG
    e.mlp_set(widths, par, 0.5, "fp32", *norm)
    e.sf_descriptor()
    G = e.nnp_get("G").astype(np.float64)
    print(f"[sf 24k] descriptor rel-rms vs fp64 {rel_rms(G, G64):.2e}", flush=True)
    e.nnp_set_dedg(dG64.astype(np.float32))
    e.clear_force()
    e.sf_force(False)
    print(f"[sf 24k] force kernel on fp64 dE/dG: rel-rms {rel_rms(e.get('f'), f64):.2e}", flush=True)
    errs = {}
    for prec in ("fp32", "bf16", "fp16", "bf16x1"):
        e.mlp_set(widths, par, 0.5, prec, *norm)
        e.sf_descriptor()
        e.clear_force()
        e.mlp_eval(False)
        dG = e.nnp_get("dG").astype(np.float64)
        e.sf_force(False)
        errs[prec] = rel_rms(e.get("f"), f64)
        print(f"[sf 24k] {prec}: dE/dG rel-rms {rel_rms(dG, dG64):.2e} force rel-rms {errs[prec]:.2e}", flush=True)
    print("[sf 24k] NNP force rel-rms vs fp64:", errs)
    assert errs["fp32"] < 1e-5
    assert errs["bf16"] < 1e-2
    assert errs["fp16"] < 5e-3
    assert errs["bf16x1"] < 1e-2


def test_precision_forms_100_snapshots():
    """NNP force error of each precision form vs FP64 on 100 seeded 768-atom snapshots (the
    water lattice start with independent orientations per seed), 2 x 64 tanh network, one
    standardisation (seed 0).  Bars on every snapshot: FP32 1e-5, BF16 (both forms) 1e-2,
    FP16 5e-3."""
    from nnpengine_amd.engine import Engine
    widths = [80, 64, 64, 1]
    sfp = M.sf_params(RC_NNP)
    par = M.mlp_init(widths, 2, 5)
    norm = None
    errs = {"fp32": [], "bf16": [], "fp16": [], "bf16x1": []}
    for seed in range(100):
        sysm = S.water_box(256, 19.73, seed=1000 + seed)
        o = pyoracle.Oracle(sysm, 1)
        o.neigh_init(RC_PAIR, SKIN, RC_PAIR)
        o.rebuild(True)
        o.sf_set(sfp["rc"], sfp["mu"], sfp["eta_r"], sfp["eta_a"])
        G64 = o.sf_descriptor()
        if norm is None:
            norm = M.standardisation(G64, sysm.typ, sysm.ntype)
        o.mlp_set(widths, par, 0.5, *norm)
        _, _, dG64 = o.mlp_eval(G64)
        o.clear_f()
        o.sf_force(dG64)
        f64 = o.get("f")
        e = Engine(sysm)
        e.neigh_init(RC_PAIR, SKIN, RC_PAIR, RC_NNP)
        e.build()
        e.sf_set(sfp["rc"], sfp["mu"], sfp["eta_r"], sfp["eta_a"])
        for prec in errs:
            e.mlp_set(widths, par, 0.5, prec, *norm)
            e.sf_descriptor()
            e.clear_force()
            e.mlp_eval(False)
            e.sf_force(False)
            errs[prec].append(rel_rms(e.get("f"), f64))
        e.close()
    for prec, v in errs.items():
        v = np.array(v)
        print(f"[snapshots] {prec}: NNP force rel-rms vs fp64 over {len(v)} snapshots: mean {v.mean():.2e} "
              f"max {v.max():.2e}", flush=True)
    assert max(errs["fp32"]) < 1e-5
    assert max(errs["bf16"]) < 1e-2
    assert max(errs["fp16"]) < 5e-3
    assert max(errs["bf16x1"]) < 1e-2


def _nnp_forces_fp64(sysm, widths, par, norm, sfp):
    o = pyoracle.Oracle(sysm, 1)
    o.neigh_init(RC_PAIR, 2.0, RC_PAIR)
    o.rebuild(True)
    o.sf_set(sfp["rc"], sfp["mu"], sfp["eta_r"], sfp["eta_a"])
    G64 = o.sf_descriptor()
    if norm is None:
        norm = M.standardisation(G64, sysm.typ, sysm.ntype)
    o.mlp_set(widths, par, 0.5, *norm)
    _, _, dG64 = o.mlp_eval(G64)
    o.clear_f()
    o.sf_force(dG64)
    return o.get("f"), norm


def test_precision_forms_24k_snapshots():
    """The same bars on the benchmark's own size and network: 24000-atom water (L = 62.1 A), the
    3 x 128 tanh network, NNP forces of every precision form vs the FP64 oracle on 8 snapshots --
    4 seeded lattice starts and 4 liquid configurations from 1 ps of MD (FP64 Verlet, classical
    + FP32 NNP forces) of the first.  One standardisation (first snapshot's descriptors)."""
    from nnpengine_amd.engine import Engine
    widths = [80, 128, 128, 128, 1]
    sfp = M.sf_params(RC_NNP)
    par = M.mlp_init(widths, 2, 2024)
    norm = None
    snaps = [S.water_box(8000, 62.1, seed=7 + 100 * k) for k in range(4)]
    # liquid snapshots: MD of the first start (pair table, bonded terms, the FP32 network)
    f0, norm = _nnp_forces_fp64(snaps[0], widths, par, norm, sfp)
    ft, ut = S.lj_table(snaps[0], RC_PAIR, 900, S.WATER_LJ)
    md = Engine(snaps[0])
    md.neigh_init(RC_PAIR, 2.0, RC_PAIR, RC_NNP)
    md.pair_table(RC_PAIR, ft, ut)
    md.bonded_set_from_system()
    md.sf_set(sfp["rc"], sfp["mu"], sfp["eta_r"], sfp["eta_a"])
    md.mlp_set(widths, par, 0.5, "fp32", *norm)
    for _ in range(4):
        md.md_run(500, dt=0.5, fp64=True, rebuild_every=10, pair=True, bonded=True, nnp=2, log_every=0)
        liq = snaps[0].copy()
        liq.x = np.ascontiguousarray(md.get("x"), np.float32)
        snaps.append(liq)
    md.close()
    errs = {"fp32": [], "bf16": [], "fp16": [], "bf16x1": []}
    for k, sysm in enumerate(snaps):
        print(f"[snapshots24k] snapshot {k}", flush=True)
        f64 = f0 if k == 0 else _nnp_forces_fp64(sysm, widths, par, norm, sfp)[0]
        e = Engine(sysm)
        e.neigh_init(RC_PAIR, 2.0, RC_PAIR, RC_NNP)
        e.build()
        e.sf_set(sfp["rc"], sfp["mu"], sfp["eta_r"], sfp["eta_a"])
        for prec in errs:
            e.mlp_set(widths, par, 0.5, prec, *norm)
            e.sf_descriptor()
            e.clear_force()
            e.mlp_eval(False)
            e.sf_force(False)
            errs[prec].append(rel_rms(e.get("f"), f64))
        e.close()
    for prec, v in errs.items():
        v = np.array(v)
        print(f"[snapshots24k] {prec}: NNP force rel-rms vs fp64 over {len(v)} snapshots: "
              + " ".join(f"{x:.2e}" for x in v), flush=True)
    assert max(errs["fp32"]) < 1e-5
    assert max(errs["bf16"]) < 1e-2
    assert max(errs["fp16"]) < 5e-3
    assert max(errs["bf16x1"]) < 1e-2


def test_tbl_descriptor_and_force(water_c0):
    sysm = water_c0
    from nnpengine_amd.engine import Engine
    cp, cw, cd, cr = M.tbl_coefficients(sysm.npair, 5, 6.0)
    res = {}
    for prec in (0, 1):
        o = pyoracle.Oracle(sysm, prec)
        o.neigh_init(6.0, 1.0, 6.0)
        o.rebuild(False)
        o.nnp_tbl_init(6.0, 600, cp, cw, cd, cr)
        res[prec] = o
    e = Engine(sysm)
    e.neigh_init(6.0, 1.0, 6.0)
    e.build()
    e.nnp_tbl_set(6.0, 600, cp, cw, cd, cr)
    e.nnp_tbl_descriptor()
    G = e.nnp_get("G").astype(np.float64)
    G32 = res[0].nnp_tbl_descriptor()
    G64 = res[1].nnp_tbl_descriptor()
    print(f"[tbl] desc rel-rms vs fp32 {rel_rms(G, G32):.2e} vs fp64 {rel_rms(G, G64):.2e}")
    assert rel_rms(G, G32) < 2e-6
    assert rel_rms(G, G64) < 1e-5
    rng = np.random.default_rng(3)
    dedg = (0.01 * rng.standard_normal(G.shape)).astype(np.float32)
    e.nnp_set_dedg(dedg)
    e.clear_force()
    e.energy_clear()
    e.nnp_tbl_force(True)
    fg = e.get("f")
    en = e.energy()
    res[1].clear_f()
    eo, wo = res[1].nnp_tbl_force(dedg.astype(np.float64))
    f64 = res[1].get("f")
    print(f"[tbl] force rel-rms vs fp64 {rel_rms(fg, f64):.2e}")
    assert rel_rms(fg, f64) < 1e-5
    assert abs(en["nnp_pair"][0] - eo) < 1e-5 * abs(eo) + 1e-6


def test_md_graph_equals_eager(water_c0):
    """Replayed hipGraph steps give bit-identical trajectories to eagerly launched steps."""
    from nnpengine_amd.engine import Engine
    sysm = water_c0
    ft, ut = S.lj_table(sysm, RC_PAIR, 900, S.WATER_LJ)
    sfp = M.sf_params(RC_NNP)
    widths = [80, 64, 64, 1]
    par = M.mlp_init(widths, sysm.ntype, 2) * 0.1
    xs = []
    for graph in (False, True):
        e = Engine(sysm)
        e.neigh_init(RC_PAIR, MD_SKIN, RC_PAIR, RC_NNP)
        e.pair_table(RC_PAIR, ft, ut)
        e.bonded_set_from_system()
        e.sf_set(sfp["rc"], sfp["mu"], sfp["eta_r"], sfp["eta_a"])
        e.mlp_set(widths, par, 0.5, "bf16")
        e.gr_init(450, RC_PAIR, RC_PAIR)
        log = e.md_run(40, dt=0.5, rebuild_every=10, nnp=2, gr_every=10, log_every=10, use_graph=graph)
        assert log.shape[0] == 4
        xs.append((e.get("x"), log))
    assert np.array_equal(xs[0][0], xs[1][0])
    assert np.array_equal(xs[0][1], xs[1][1])


def test_md_gr_on_second_stream_equals_standalone(water_c1):
    """NNP MD steps whose g(r) tally (folded into the rebuild) and S(q) run on the second stream
    beside the NNP passes: the counts and h(r) left by the step equal a standalone tally on the
    same positions, bit for bit (24k atoms, 12 A, BF16 one-term network)."""
    from nnpengine_amd.engine import Engine
    sysm = water_c1
    ft, ut = S.lj_table(sysm, RC_PAIR, 900, S.WATER_LJ)
    sfp = M.sf_params(RC_NNP)
    widths = [80, 128, 128, 128, 1]
    par = M.mlp_init(widths, sysm.ntype, 5) * 0.1
    e = Engine(sysm)
    e.neigh_init(RC_PAIR, 2.0, 12.0, RC_NNP)
    e.pair_table(RC_PAIR, ft, ut)
    e.bonded_set_from_system()
    e.sf_set(sfp["rc"], sfp["mu"], sfp["eta_r"], sfp["eta_a"])
    e.mlp_set(widths, par, 0.5, "bf16x1")
    e.gr_init(450, 12.0, RC_PAIR)
    e.sq_init(1000, 0.3, 20.0)
    # steps 0..20: step 20 rebuilds and tallies; the positions stay as tallied
    e.md_run(21, dt=0.5, fp64=True, rebuild_every=10, nnp=2, gr_every=10, gr_long_every=10, sq=True)
    side, h_side = e.gr_get()
    e.gr_tally(full=True)
    alone, h_alone = e.gr_get()
    print(f"[gr side] total counts {int(side.sum())} vs {int(alone.sum())}", flush=True)
    assert side.sum() > 0
    np.testing.assert_array_equal(side, alone)
    np.testing.assert_array_equal(h_side, h_alone)
    assert np.all(np.isfinite(e.sq_compute()))


def test_md_nve_energy_conservation(water_c0):
    """FP64 integration, dt = 0.5 fs, 400 steps: total energy drift small against its fluctuation."""
    from nnpengine_amd.engine import Engine
    sysm = water_c0
    ft, ut = S.lj_table(sysm, RC_PAIR, 900, S.WATER_LJ)
    e = Engine(sysm)
    e.neigh_init(RC_PAIR, MD_SKIN, RC_PAIR)
    e.pair_table(RC_PAIR, ft, ut)
    e.bonded_set_from_system()
    log = e.md_run(400, dt=0.5, fp64=True, rebuild_every=4, log_every=20)
    et = log[:, 0] + log[:, 2] + log[:, 4] + log[:, 6] + log[:, 14]
    print("[nve] E_tot first/last", et[1], et[-1], "ke", log[1, 14], log[-1, 14])
    st = e.md_stats()
    assert st["skin_violations"] == 0
    assert abs(et[-1] - et[1]) < 0.02 * abs(log[1, 14])


@pytest.mark.parametrize("every", [10, 5])
def test_md_step_equals_call_sequence(every, water_c0):
    """nnpe_md_run's NNP step -- graph-replayed, classical terms on a second stream, the FP64
    half-kick fused into the SF force pass, and in the captured interval of steps 10-19 the next
    step's position update fused there too (two position buffers trading places) -- against the
    same step spelt out through the C ABI (update_x, build every 10th step, pair, bonded, SF
    descriptors, network, SF force, update_v): bit-identical positions and velocities after 25
    steps with a rebuild every 10 steps, BF16 network.  With a rebuild every 5 steps (an odd
    interval) the position update stays a kernel of its own in every step."""
    widths = [80, 64, 64, 1]
    runs = []
    for mode in ("md", "calls"):
        e, o, par, G64, norm = _setup(water_c0, widths)
        e.pair_table(RC_PAIR, *S.lj_table(water_c0, RC_PAIR, 900, S.WATER_LJ))
        e.bonded_set_from_system()
        e.mlp_set(widths, par, 0.5, "bf16", *norm)
        if mode == "md":
            e.md_run(25, dt=0.5, fp64=True, rebuild_every=every, pair=True, bonded=True, nnp=2, use_graph=True)
        else:
            e.verlet_set(0.5, fp64=True)
            for step in range(25):
                if step > 0:
                    e.update_x()
                if step % every == 0:
                    e.build()
                e.pair_compute(False)
                e.bonded_compute(False)
                e.sf_descriptor()
                e.mlp_eval(False)
                e.sf_force(False)
                e.update_v()
        runs.append((e.get("x"), e.get("v")))
        print(f"[md vs calls] {mode}: |v| rms {np.sqrt((runs[-1][1] ** 2).mean()):.6f}", flush=True)
    assert np.array_equal(runs[0][0], runs[1][0])
    assert np.array_equal(runs[0][1], runs[1][1])


def test_md_checkpoint_rerun_bit_identical(water_c0):
    """A chunk repeated from a checkpoint (nnpe_state_set of x, v, f, f_prev and nnpe_md_set_step,
    the path that repeats a chunk after a list overflow) reproduces the first attempt bit for bit,
    NNP on, FP64 integration."""
    widths = [80, 64, 64, 1]
    e, o, par, G64, norm = _setup(water_c0, widths)
    e.pair_table(RC_PAIR, *S.lj_table(water_c0, RC_PAIR, 900, S.WATER_LJ))
    e.bonded_set_from_system()
    e.mlp_set(widths, par, 0.5, "bf16", *norm)
    kw = dict(dt=0.5, fp64=True, rebuild_every=10, nnp=2, log_every=10)
    e.md_run(100, **kw)
    step0 = e.md_stats()["steps"]
    saved = {w: e.get(w) for w in ("x", "v", "f", "f_prev")}
    log_a = e.md_run(100, **kw)
    xa, va = e.get("x"), e.get("v")
    for w, a in saved.items():
        e.set(w, a)
    e.md_set_step(step0)
    log_b = e.md_run(100, **kw)
    assert np.array_equal(xa, e.get("x")) and np.array_equal(va, e.get("v"))
    assert np.array_equal(log_a, log_b)
    # the checkpointed driver over the same steps gives the same trajectory
    for w, a in saved.items():
        e.set(w, a)
    e.md_set_step(step0)
    log_c, redo = e.md_run_checkpointed(100, chunk=50, **kw)
    assert redo == 0 and np.array_equal(log_a, log_c) and np.array_equal(xa, e.get("x"))
