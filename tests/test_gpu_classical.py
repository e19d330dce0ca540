"""GPU parity of the classical path (neighbour list, tabulated pair, bonded, Verlet, g(r),
S(q)) against the oracle on identical seeded inputs.  Bars: bit-exact list sets and g(r)
counts; forces/energies within 1e-5 relative RMS of the FP64 oracle (and ~1e-6 of the FP32
oracle, which differs only in summation order)."""
import dataclasses
import os

import numpy as np
import pytest

from nnpengine_amd import system as S
from oracle import pyoracle
from tests.conftest import rel_rms
from tests.golden.make_golden import make_system
from tests.golden.make_golden import tables as golden_tables

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

pytestmark = pytest.mark.gpu

RC, SKIN = 9.0, 0.5


def _engine(sysm, rc=RC, skin=SKIN, gr=RC, short=0.0):
    from nnpengine_amd.engine import Engine
    e = Engine(sysm)
    e.neigh_init(rc, skin, gr, short)
    return e


def _oracle(sysm, prec, rc=RC, skin=SKIN, gr=RC):
    o = pyoracle.Oracle(sysm, prec)
    o.neigh_init(rc, skin, gr)
    return o


@pytest.mark.parametrize("case", ["c0", "ragged", "c1"])
def test_list_bit_exact(case, water_c0, water_c1):
    if case == "c0":
        sysm, rc, skin, gr = water_c0, RC, SKIN, RC
    elif case == "ragged":
        sysm, rc, skin, gr = S.water_box(300, 21.0, seed=3), 8.0, 1.0, 8.0
    else:
        sysm, rc, skin, gr = water_c1, RC, 2.0, 12.0
    print(f"[list] case={case} natom={sysm.natom}", flush=True)
    e = _engine(sysm, rc, skin, gr)
    e.build()
    o = _oracle(sysm, 0, rc, skin, gr)
    ng_o = o.rebuild(False)
    kg = pyoracle.canonical(e.list_entries())
    ko = pyoracle.canonical(o.list_entries())
    assert len(kg) == len(ko)
    assert np.array_equal(kg, ko)
    _, ng = e.list_size()
    assert ng == ng_o
    # refolded positions identical
    assert np.array_equal(e.get("x"), o.get("x"))


def _chain_system(base, lens):
    """The base water box with its first sum(lens) atoms regrouped into linear chain molecules of
    the given lengths (bonds, angles and dihedrals along the chain): molecules larger and
    smaller than 64 atoms in one system, for the two ways the list build looks up special
    classes."""
    n_ch = int(sum(lens))
    assert n_ch % 3 == 0
    moltypes = [base.moltypes[0]]
    kind = {}
    for L in sorted(set(lens)):
        kind[L] = len(moltypes)
        moltypes.append(S.MolType(L, bonds=[(0, k, k + 1) for k in range(L - 1)],
                                  angles=[(0, k, k + 1, k + 2) for k in range(L - 2)],
                                  dihedrals=[(0, k, k + 1, k + 2, k + 3) for k in range(L - 3)]))
    mol_id, intra, mtyp = [], [], []
    for m, L in enumerate(lens):
        mol_id += [m] * L
        intra += list(range(L))
        mtyp.append(kind[L])
    nw = (base.natom - n_ch) // 3
    mol_id += list(np.repeat(np.arange(len(lens), len(lens) + nw), 3))
    intra += [0, 1, 2] * nw
    mtyp += [0] * nw
    s = base.copy()
    s.mol_id = np.asarray(mol_id, np.int32)
    s.intra_id = np.asarray(intra, np.int32)
    s.mol_type = np.asarray(mtyp, np.int32)
    s.moltypes = moltypes
    return s


def test_list_bit_exact_large_molecules(water_c0):
    sysm = _chain_system(water_c0, [81, 63, 81, 63])
    print(f"[list] chains 81/63 natom={sysm.natom}", flush=True)
    e = _engine(sysm, RC, SKIN, RC, 6.0)
    e.build()
    o = _oracle(sysm, 0)
    o.rebuild(False)
    full = e.list_entries()
    kg = pyoracle.canonical(full)
    ko = pyoracle.canonical(o.list_entries())
    assert np.array_equal(kg, ko)
    # special classes 1..3 of chain atoms on both sides of the 64-atom boundary
    for lo, hi in ((0, 81), (81, 144)):
        m = (full[:, 0] >= lo) & (full[:, 0] < hi)
        assert set(np.unique(full[m, 3])) >= {0, 1, 2, 3}
    # short rows: exactly the full-list entries within (6 + skin), by the same FP32 operations
    # (image shift, differences, squared distance summed x, y, z)
    x = e.get("x").astype(np.float32)
    bl = np.float32(sysm.bl)
    i, j, code = full[:, 0], full[:, 1], full[:, 2]
    shift = np.array([-bl, np.float32(0), bl], np.float32)
    img = x[j] + np.stack([shift[code // 9], shift[(code % 9) // 3], shift[code % 3]], 1)
    img[code == 13] = x[j[code == 13]]
    d = x[i] - img
    rsq = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
    sc = np.float32(6.0) + np.float32(SKIN)
    want = pyoracle.canonical(full[rsq < sc * sc])
    got = pyoracle.canonical(e.list_entries(1))
    assert len(got) > 0 and np.array_equal(got, want)


def _capacity0(natom, bl, rc):
    """The list capacity a context starts with: 1.5x the ideal-gas count within rc at the mean
    density, +64, rounded up to 64 (nnpe_neigh_init)."""
    rho = natom / float(bl) ** 3
    k = int(1.5 * (4.0 / 3.0 * np.pi * rc ** 3 * rho)) + 64
    return (k + 63) // 64 * 64


def test_list_capacity_growth_dense_slab():
    """A liquid squeezed into a slab of 1/16 of the box height (16x the mean density): rows far
    beyond the starting capacities of both lists, and cells of more than 64 atoms (the sort's
    long-cell path).  The reference stops with "nei_list overflow" (src/neighlist.cpp:327)
    beyond its fixed 2048 columns; this library grows the capacity and builds again, and the
    result must still be the oracle's list, bit for bit."""
    sysm = S.water_box(1000, 31.0, seed=5)
    x = sysm.x.astype(np.float32).copy()
    x[:, 2] *= np.float32(1.0 / 16.0)
    sysm.x = x
    sc = 4.0
    print(f"[list] dense slab natom={sysm.natom}", flush=True)
    e = _engine(sysm, RC, SKIN, RC, sc)
    e.build()
    o = _oracle(sysm, 0)
    ng_o = o.rebuild(False)
    full = e.list_entries()
    assert np.array_equal(pyoracle.canonical(full), pyoracle.canonical(o.list_entries()))
    assert e.list_size()[1] == ng_o
    # the growth path ran: the longest row exceeds the starting capacity
    rows = np.bincount(full[:, 0], minlength=sysm.natom)
    assert rows.max() > _capacity0(sysm.natom, sysm.bl, RC + SKIN)
    # and cells of more than 64 atoms were sorted (cell edge (rc + skin) / 2 from the grid origin)
    xr = e.get("x").astype(np.float64)
    g_dx = (RC + SKIN) / 2.0
    cell = np.floor((xr + (RC + SKIN) + sysm.bl / 2.0) / g_dx).astype(np.int64)
    _, occ = np.unique(cell[:, 0] * 1_000_000 + cell[:, 1] * 1000 + cell[:, 2], return_counts=True)
    assert occ.max() > 64
    # short rows: the full-list entries within sc + skin (same FP32 operations), past their
    # own starting capacity
    xf = xr.astype(np.float32)
    bl = np.float32(sysm.bl)
    i, j, code = full[:, 0], full[:, 1], full[:, 2]
    shift = np.array([-bl, np.float32(0), bl], np.float32)
    img = xf[j] + np.stack([shift[code // 9], shift[(code % 9) // 3], shift[code % 3]], 1)
    img[code == 13] = xf[j[code == 13]]
    d = xf[i] - img
    rsq = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
    c2 = np.float32(sc) + np.float32(SKIN)
    short = e.list_entries(1)
    assert np.array_equal(pyoracle.canonical(short), pyoracle.canonical(full[rsq < c2 * c2]))
    # the short lists are per neighbour element: a row of one element is what must fit
    per = np.bincount(short[:, 0] * sysm.ntype + sysm.typ[short[:, 1]], minlength=sysm.natom * sysm.ntype)
    assert per.max() > _capacity0(sysm.natom, sysm.bl, sc + SKIN)


def test_call_order_errors(water_c0):
    """Calls out of order fail with NNPE_ERR_STATE and a message, as the reference's
    END_PROGRAM would stop the program (src/util.cpp:117-124); the context stays usable."""
    from nnpengine_amd.engine import Engine, NnpeError
    e = Engine(water_c0)
    with pytest.raises(NnpeError) as ei:
        e.build()
    assert ei.value.code == 3 and "nnpe_neigh_init" in str(ei.value)
    e.neigh_init(RC, SKIN, RC, 0.0)
    e.build()
    o = _oracle(water_c0, 0)
    o.rebuild(False)
    assert np.array_equal(pyoracle.canonical(e.list_entries()), pyoracle.canonical(o.list_entries()))


def _forces_case(case, water_c0):
    """(system, pair cutoff, tables, reference fixture or None) of a force-parity case."""
    if case == "c0":
        sysm = water_c0
    elif case == "ragged":
        sysm = S.water_box(300, 21.0, seed=3)
    else:
        sysm = make_system("butane")
        if case == "butane_rb":
            sysm = S.with_rb_dihedral(sysm)
    if case.startswith("butane"):
        ft, ut = golden_tables("butane", sysm)
        fix = np.load(os.path.join(GOLD, "butane_snap.npz")) if case == "butane" else None
    else:
        ft, ut = S.lj_table(sysm, RC, 900, S.WATER_LJ)
        fix = np.load(os.path.join(GOLD, "water_c0_snap.npz")) if case == "c0" else None
    return sysm, ft, ut, fix


def _oracle_forces(sysm, ft, ut, prec, pair=True):
    o = _oracle(sysm, prec)
    o.rebuild(False)
    o.pair_set(RC, ft, ut)
    o.clear_f()
    eb = o.bonded_compute()
    ep = o.pair_compute() if pair else (0.0, 0.0)
    return o.get("f"), ep, eb


ENERGY_TERMS = ("pair", "bond", "angle", "dihedral")


@pytest.mark.parametrize("case", ["c0", "ragged", "butane", "butane_rb"])
def test_pair_bonded_forces(case, water_c0):
    """Tabulated pair + bonded forces, energies and virials against the FP64 oracle (1e-5) and,
    where the reference built from source wrote one, its own FP32 result.  "ragged" is 900
    atoms, so the last workgroup of the pair pass is partial; "butane" is 64 united-atom chains
    with bonds, angles and the reference's cubic dihedral; "butane_rb" the same chains with the
    Ryckaert-Bellemans quintic dihedral."""
    sysm, ft, ut, fix = _forces_case(case, water_c0)
    print(f"[pair] case={case} natom={sysm.natom}", flush=True)
    e = _engine(sysm)
    e.build()
    e.pair_table(RC, ft, ut)
    e.bonded_set_from_system()
    e.energy_clear()
    e.pair_compute(True)
    e.bonded_compute(True)
    fg = e.get("f")
    en = e.energy()
    f32, ep32, eb32 = _oracle_forces(sysm, ft, ut, 0)
    f64, ep64, eb64 = _oracle_forces(sysm, ft, ut, 1)
    print(f"[pair] {case} force rel rms vs fp32 oracle", rel_rms(fg, f32), "vs fp64", rel_rms(fg, f64))
    assert rel_rms(fg, f32) < 2e-6
    assert rel_rms(fg, f64) < 1e-5
    if fix is not None:
        print(f"[pair] {case} vs the reference's own FP32 forces", rel_rms(fg, fix["f"]))
        assert rel_rms(fg, fix["f"]) < 5e-6
    e_gpu = sum(en[t][0] for t in ENERGY_TERMS)
    w_gpu = sum(en[t][1] for t in ENERGY_TERMS)
    for k in (0, 1):
        r64 = ep64[k] + eb64[k]
        got = e_gpu if k == 0 else w_gpu
        print(f"[pair] {case} {'energy' if k == 0 else 'virial'} {got:.8f} fp64 {r64:.8f} rel {abs(got - r64) / abs(r64):.2e}")
        assert abs(got - r64) < 1e-5 * abs(r64)


@pytest.mark.parametrize("case", ["c0", "butane", "butane_rb"])
def test_bonded_forces_alone(case, water_c0):
    """The bonded pass alone (bond, angle, dihedral) against the FP64 oracle: the pair forces
    would otherwise hide its error.  Every site is displaced by up to 0.05 A (seeded), so the
    molecules leave their equilibrium geometry (the water box starts at it, with zero bonded
    force)."""
    sysm, ft, ut, _ = _forces_case(case, water_c0)
    sysm = sysm.copy()
    rng = np.random.default_rng(17)
    sysm.x = (sysm.x + rng.uniform(-0.05, 0.05, sysm.x.shape)).astype(np.float32)
    e = _engine(sysm)
    e.build()
    e.bonded_set_from_system()
    e.clear_force()
    e.energy_clear()
    e.bonded_compute(True)
    fg = e.get("f")
    en = e.energy()
    f64, _, eb64 = _oracle_forces(sysm, ft, ut, 1, pair=False)
    err = rel_rms(fg, f64)
    eg = en["bond"][0] + en["angle"][0] + en["dihedral"][0]
    print(f"[bonded] {case} force rel rms vs fp64 {err:.2e}; energy {eg:.8f} vs {eb64[0]:.8f}; "
          f"virial {en['bond'][1]:.6f} vs {eb64[1]:.6f}", flush=True)
    assert err < 1e-5
    assert abs(eg - eb64[0]) < 1e-5 * abs(eb64[0])
    assert abs(en["bond"][1] - eb64[1]) < 1e-5 * max(abs(eb64[1]), 1e-3 * abs(eb64[0]))
    if case.startswith("butane"):
        assert abs(en["dihedral"][0]) > 0
    # every atom's bonded force has one writer: a second pass adds exactly the same forces
    e.bonded_compute(False)
    assert np.array_equal(e.get("f"), 2 * fg)


def test_md_steps_match_fp32_oracle(water_c0):
    """20 velocity-Verlet steps with skin-triggered rebuilds, same call sequence as the
    reference MD loop (src/run.cpp:230-346)."""
    sysm = water_c0
    ft, ut = S.lj_table(sysm, RC, 900, S.WATER_LJ)
    e = _engine(sysm)
    e.pair_table(RC, ft, ut)
    e.bonded_set_from_system()
    e.verlet_set(0.5, fp64=False)
    o = _oracle(sysm, 0)
    o.pair_set(RC, ft, ut)
    o.set_dt(0.5)
    nb_g = nb_o = 0
    for step in range(21):
        if step > 0:
            e.update_x()
            o.update_x()
        vg, vo = e.isvalid(), o.isvalid()
        assert vg == vo
        if not vg:
            e.build(); o.rebuild(False)
            nb_g += 1; nb_o += 1
        else:
            o.update_ghost_pos()
        e.pair_compute(False); e.bonded_compute(False)
        o.clear_f(); o.bonded_compute(False); o.pair_compute(False)
        e.update_v(); o.update_v()
    assert nb_g == nb_o >= 2
    assert rel_rms(e.get("x"), o.get("x")) < 1e-6
    assert rel_rms(e.get("v"), o.get("v")) < 1e-4


@pytest.mark.parametrize("fp64", [False, True])
def test_md_run_counts_skin_violations(fp64, water_c0):
    """nnpe_md_run with a fixed rebuild interval counts the rebuilds that came too late (an atom
    moved more than skin/2 since the last build; the check rides on the position update of the
    rebuild step): the same count as NEIGHLIST::isvalid (src/neighlist.cpp:178-187) asked before
    each rebuild of the same trajectory spelt out through the C ABI, with bit-identical positions.
    With a 0.7 A skin the first 10-step interval of this box is short enough and the later ones
    are not."""
    ft, ut = S.lj_table(water_c0, RC, 900, S.WATER_LJ)
    nstep, every = 80, 10
    out = []
    for mode in ("md", "calls"):
        e = _engine(water_c0, skin=0.7)
        e.pair_table(RC, ft, ut)
        e.bonded_set_from_system()
        late = 0
        if mode == "md":
            e.md_run(nstep, dt=0.5, fp64=fp64, rebuild_every=every, pair=True, bonded=True, use_graph=True)
            late = e.md_stats()["skin_violations"]
        else:
            e.verlet_set(0.5, fp64=fp64)
            for step in range(nstep):
                if step > 0:
                    e.update_x()
                if step % every == 0:
                    if step > 0 and not e.isvalid():
                        late += 1
                    e.build()
                e.pair_compute(False)
                e.bonded_compute(False)
                e.update_v()
        out.append((late, e.get("x")))
        e.close()
    print(f"[skin] late rebuilds: md {out[0][0]}, calls {out[1][0]} of {nstep // every - 1}", flush=True)
    assert out[0][0] == out[1][0]
    assert 0 < out[1][0] < nstep // every
    assert np.array_equal(out[0][1], out[1][1])


def _trajectory(e, o, nstep):
    """The reference MD loop's call sequence (src/run.cpp:230-346) on the engine and the oracle
    side by side; returns the number of rebuilds (equal on both sides)."""
    nb = 0
    for step in range(nstep + 1):
        if step > 0:
            e.update_x()
            o.update_x()
        vg, vo = e.isvalid(), o.isvalid()
        assert vg == vo, f"step {step}: rebuild decision differs"
        if not vg:
            e.build(); o.rebuild(False)
            nb += 1
        else:
            o.update_ghost_pos()
        e.pair_compute(False); e.bonded_compute(False)
        o.clear_f(); o.bonded_compute(False); o.pair_compute(False)
        e.update_v(); o.update_v()
    return nb


def test_md_steps_fp64_form_match_fp64_oracle(water_c0):
    """The FP64 integration form (positions and velocities in FP64) and the FP32 form against
    the FP64 oracle's trajectory over 40 steps with rebuilds.  Both sides differ by the FP32
    force evaluation on the engine (the chaotic growth of its ~1e-6 relative error gives ~1e-8
    in x over 20 fs); the FP32 form adds FP32 rounding of x and v on every step and lands
    several times further away."""
    sysm = water_c0
    ft, ut = S.lj_table(sysm, RC, 900, S.WATER_LJ)
    errs = {}
    for fp64 in (True, False):
        e = _engine(sysm)
        e.pair_table(RC, ft, ut)
        e.bonded_set_from_system()
        e.verlet_set(0.5, fp64=fp64)
        o = _oracle(sysm, 1)
        o.pair_set(RC, ft, ut)
        o.set_dt(0.5)
        nb = _trajectory(e, o, 40)
        assert nb >= 3
        errs[fp64] = (rel_rms(e.get("x"), o.get("x")), rel_rms(e.get("v"), o.get("v")))
        e.close()
    print(f"[md] 40 steps vs FP64 oracle: FP64 form x {errs[True][0]:.2e} v {errs[True][1]:.2e}; "
          f"FP32 form x {errs[False][0]:.2e} v {errs[False][1]:.2e}", flush=True)
    assert errs[True][0] < 2e-7 and errs[True][1] < 5e-5
    assert errs[True][0] < 0.5 * errs[False][0]


def test_fp64_verlet_is_the_fp64_restatement(water_c0):
    """The FP64 form's update itself (src/verlet.cpp:29-97 restated in double) against the FP64
    oracle's update given the same forces: every step the engine's FP32 forces, widened, are
    the oracle's forces.  Masses 16 and 1 (their reciprocals exact in FP32) give both sides the
    same 1/m, so the two integrations are the same double arithmetic: positions and velocities
    agree to 1e-14 over 40 steps with list rebuilds and re-sorts on the engine."""
    import dataclasses
    sysm = dataclasses.replace(water_c0, mass=np.where(water_c0.typ == 0, 16.0, 1.0).astype(np.float32))
    ft, ut = S.lj_table(sysm, RC, 900, S.WATER_LJ)
    e = _engine(sysm)
    e.pair_table(RC, ft, ut)
    e.bonded_set_from_system()
    e.verlet_set(0.5, fp64=True)
    o = pyoracle.Oracle(sysm, 1)
    o.set_dt(0.5)
    o.set("x", e.get("x"))
    o.set("v", e.get("v"))
    o.set("f_prev", e.get("f_prev"))
    worst, nb = [0.0, 0.0], 0
    for step in range(41):
        if step > 0:
            e.update_x()
            o.update_x()
        if not e.isvalid():
            e.build()
            nb += 1
        e.pair_compute(False)
        e.bonded_compute(False)
        o.set("f", e.get("f"))
        e.update_v()
        o.update_v()
        for k, w in enumerate(("x", "v")):
            a, b = e.get(w), o.get(w)
            worst[k] = max(worst[k], float(np.abs(a - b).max() / np.abs(b).max()))
    print(f"[md] FP64 update vs FP64 restatement, same forces: max rel x {worst[0]:.1e} v {worst[1]:.1e}; "
          f"{nb} builds", flush=True)
    assert nb >= 3
    assert worst[0] <= 1e-14 and worst[1] <= 1e-14


def test_gr_counts_bit_exact(water_c1):
    sysm = water_c1
    e = _engine(sysm, RC, 2.0, 12.0)
    e.build()
    nb = e.gr_init(450, 12.0, 9.0)
    o = _oracle(sysm, 0, RC, 2.0, 12.0)
    o.rebuild(False)
    assert o.gr_init(450, 12.0, 9.0) == nb == 600
    e.gr_tally(full=True)
    o.gr_tally(True)
    cg, hg = e.gr_get()
    co, ho = o.gr_get()
    assert np.array_equal(cg, co)
    # real-ghost pairs count 1 from each side; the two sides see FP32-rounded image positions
    # and may land in different bins, so only real-real pairs are guaranteed to add 2
    assert cg.sum() > 0
    # inner tally keeps outer bins of the previous full tally (reference behaviour)
    e.gr_tally(full=False)
    o.gr_tally(False)
    cg2, _ = e.gr_get()
    co2, ho2 = o.gr_get()
    assert np.array_equal(cg2, co2)
    assert np.array_equal(cg2[:, 450:], cg[:, 450:])
    e.sq_init(1000, 0.3, 20.0)
    sg = e.sq_compute()
    so = o.sq(1000, 0.3, 20.0)
    assert np.max(np.abs(sg - so)) < 1e-9 * max(1.0, np.max(np.abs(so)))


def test_gr_fused_with_rebuild_equals_standalone(water_c1):
    """The full g(r) tally folded into a list rebuild (the MD loop's rebuild + g(r) steps) gives
    the same counts, bit for bit, as the standalone full tally on the same positions (24k atoms,
    12 A, pair stencil 11 A: pairs in the outer cells come from the shell pass)."""
    from nnpengine_amd.engine import Engine
    sysm = water_c1
    ft, ut = S.lj_table(sysm, RC, 900, S.WATER_LJ)
    e = Engine(sysm)
    e.neigh_init(RC, 2.0, 12.0, 0.0)
    e.pair_table(RC, ft, ut)
    e.bonded_set_from_system()
    e.gr_init(450, 12.0, RC)
    # steps 0..20: step 20 rebuilds and tallies (fused); the positions stay as tallied
    e.md_run(21, dt=0.5, fp64=True, rebuild_every=10, gr_every=10, gr_long_every=10)
    fused, _ = e.gr_get()
    e.gr_tally(full=True)
    alone, _ = e.gr_get()
    print(f"[gr fused] total counts {int(fused.sum())} vs {int(alone.sum())}", flush=True)
    assert fused.sum() > 0
    np.testing.assert_array_equal(fused, alone)


def _many_element_box(base, nt):
    """The water box with its atoms relabelled into nt elements (oxygens 0/1 by molecule,
    hydrogens 2 .. nt-1 by molecule and position) and a distinct LJ pair for every element pair.
    Four elements is the most a list entry names (the pair pass and the inner tally take the
    neighbour's element from the entry); with five they read it from the element array."""
    sysm = base.copy()
    typ = np.where(base.typ == 0, base.mol_id % 2, 2 + (base.mol_id + base.intra_id) % (nt - 2)).astype(np.int32)
    sysm = dataclasses.replace(sysm, ntype=nt, typ=typ)
    lj = {}
    for a in range(nt):
        for b in range(a, nt):
            lj[(a, b)] = (0.05 + 0.02 * (a + 2 * b), 2.2 + 0.15 * (a + b))
    return sysm, lj


@pytest.mark.parametrize("nt", [4, 5])
def test_many_elements_pair_and_gr(nt, water_c0):
    """Four and five elements: pair forces against the FP32 and FP64 oracles, and the full and
    inner g(r) tallies (10 and 15 element pairs) bit-exact against the oracle."""
    sysm, lj = _many_element_box(water_c0, nt)
    print(f"[elements] ntype={nt} counts={np.bincount(sysm.typ).tolist()}", flush=True)
    ft, ut = S.lj_table(sysm, RC, 900, lj)
    e = _engine(sysm)
    e.build()
    e.pair_table(RC, ft, ut)
    e.pair_compute(False)
    fg = e.get("f")
    fo = {}
    for prec in (0, 1):
        o = _oracle(sysm, prec)
        o.rebuild(False)
        o.pair_set(RC, ft, ut)
        o.clear_f()
        o.pair_compute()
        fo[prec] = o.get("f")
    print(f"[elements] {nt}: pair force rel rms vs fp32 oracle", rel_rms(fg, fo[0]), "vs fp64", rel_rms(fg, fo[1]), flush=True)
    assert rel_rms(fg, fo[0]) < 2e-6
    assert rel_rms(fg, fo[1]) < 1e-5
    nb = e.gr_init(450, RC, RC)
    o = _oracle(sysm, 0)
    o.rebuild(False)
    assert o.gr_init(450, RC, RC) == nb
    e.gr_tally(full=True)
    o.gr_tally(True)
    cg, _ = e.gr_get()
    co, _ = o.gr_get()
    assert cg.shape[0] == nt * (nt + 1) // 2 and cg.sum() > 0
    assert np.array_equal(cg, co)
    e.gr_tally(full=False)
    o.gr_tally(False)
    cg2, _ = e.gr_get()
    co2, _ = o.gr_get()
    assert np.array_equal(cg2, co2)
