"""Set-up of the benchmark workloads on an Engine (host side; every step runs in the library).

Reverse-MD refinement on a molecular liquid with all bonded terms (the butane workload):
4000 united-atom CH3-CH2-CH2-CH3 chains, 16 000 sites, in a cubic box of 87.3 A (0.58 g/cm^3);
CH3/CH2 LJ sigma 3.75/3.95 A, epsilon 0.195/0.091 kcal/mol (Lorentz-Berthelot), tabulated to
9 A in the 0.02 A bins of the g(r) grid (RMDF refines the table on that grid, as the reference
ties the table spacing to the g(r) spacing: src/run.cpp:147-153); bonds 1.54 A, angles 114 deg,
the Ryckaert-Bellemans torsion; lists to 11 A (2 A skin) rebuilt every 10 steps; the
symmetry-function NNP (6 A, 16 radial + 16 angular per element pair) with a 3 x 128 tanh
network; g(r) of CH3-CH3, CH3-CH2, CH2-CH2 on 0-12 A in 0.02 A bins and S(q) on 1000 points of
(0, 20] 1/A (the reference's S(q) grid for RMDF starts at q = 0: q_i = (i + 1/2) q_max / nq,
src/grsq.cpp:258-280).  The target S(q) comes from a run with every LJ epsilon scaled by 1.1;
one refinement iteration is an MD run with the g(r) -> S(q) -> RMDF step every 10 steps and the
table rewritten on the logged steps (src/run.cpp:285-298).
"""
from __future__ import annotations

import numpy as np

from . import model as M
from . import system as S

BUTANE_C3 = dict(nmol=4000, bl=87.3, seed=31, rc=9.0, skin=2.0, nbin=450, gr_rmax=12.0, nq=1000, q_max=20.0,
                 rc_nnp=6.0, widths=(80, 128, 128, 128, 1), dt=0.5, rebuild_every=10, gr_every=10)


def butane_c3_system(nmol=None, bl=None, seed=None):
    c = BUTANE_C3
    sysm = S.butane_box(nmol or c["nmol"], bl or c["bl"], seed=c["seed"] if seed is None else seed)
    return S.with_rb_dihedral(sysm)


def butane_tables(sysm, eps_scale=1.0):
    c = BUTANE_C3
    return S.lj_table(sysm, c["rc"], c["nbin"], S.butane_lj(eps_scale))


def classical_engine(sysm, ft, ut, nnp=False):
    """Engine with lists, the pair table, bonded terms, g(r) and S(q) of the butane workload."""
    from .engine import Engine
    c = BUTANE_C3
    e = Engine(sysm)
    e.neigh_init(c["rc"], c["skin"], c["gr_rmax"], c["rc_nnp"] if nnp else 0.0)
    e.pair_table(c["rc"], ft, ut)
    e.bonded_set_from_system()
    e.gr_init(c["nbin"], c["gr_rmax"], c["rc"])
    e.sq_init(c["nq"], 0.0, c["q_max"])
    return e


def nnp_attach(e, sysm, precision="bf16x1", seed=2024):
    """The symmetry-function NNP with seeded weights and the descriptor standardisation of the
    current positions (descriptors from the engine itself)."""
    c = BUTANE_C3
    sfp = M.sf_params(c["rc_nnp"])
    e.sf_set(sfp["rc"], sfp["mu"], sfp["eta_r"], sfp["eta_a"])
    e.build()
    e.sf_descriptor()
    G0 = e.nnp_get("G")
    # G is by atom id; element of each atom
    norm = M.standardisation(G0.astype(np.float64), sysm.typ, sysm.ntype)
    par = M.mlp_init(list(c["widths"]), sysm.ntype, seed)
    e.mlp_set(list(c["widths"]), par, 0.5, precision, *norm)
    return par, norm


def rmdf_data(s_exp, ntype, q_max, weight=1.0):
    """RMDF data set [1][nq][3 + ntype]: (q, S_exp, weight, form factor per element), q on the S(q)
    grid, unit form factors (united-atom sites)."""
    nq = len(s_exp)
    q = (np.arange(nq) + 0.5) * q_max / nq
    rows = np.concatenate([q[:, None], np.asarray(s_exp, np.float64)[:, None], np.full((nq, 1), weight),
                           np.ones((nq, ntype))], 1)
    return rows[None].astype(np.float32)


def total_sq(e, sysm):
    """The total S(q) the refinement compares with the data for unit form factors (united-atom
    sites): the Faber-Ziman sum of the partial S_ab(q) of the last g(r) tally with the weights
    c_a c_b (2 - delta_ab), c the element fractions -- the combination RMDF makes of them
    (src/grsq.cpp:899-1000)."""
    sq = e.sq_compute()
    c = np.bincount(sysm.typ, minlength=sysm.ntype) / sysm.natom
    out = np.zeros(sq.shape[1])
    p = 0
    for a in range(sysm.ntype):
        for b in range(a, sysm.ntype):
            out += sq[p] * c[a] * c[b] * (1.0 if a == b else 2.0)
            p += 1
    return out


def butane_target(sysm, steps, samples=4):
    """Target total S(q) of the workload: MD of the chains with every LJ epsilon scaled by 1.1
    (classical terms, FP64 Verlet), averaged over `samples` full g(r) tallies spread over the
    second half of `steps`."""
    c = BUTANE_C3
    ft, ut = butane_tables(sysm, 1.1)
    e = classical_engine(sysm, ft, ut)
    kw = dict(dt=c["dt"], fp64=True, rebuild_every=c["rebuild_every"], pair=True, bonded=True, nnp=0)
    first = steps // 2 - steps // 2 % c["rebuild_every"]
    e.md_run(first, **kw)
    acc = np.zeros(c["nq"])
    chunk = max(c["rebuild_every"], (steps - first) // samples // c["rebuild_every"] * c["rebuild_every"])
    for _ in range(samples):
        e.md_run(chunk, **kw)
        e.gr_tally(full=True)
        acc += total_sq(e, sysm)
    x = e.get("x")
    e.close()
    return acc / samples, x
