"""Pins the oracle functions that no reference build covers (CPU-only tests).

* Angular symmetry functions: the oracle sums them through per-atom moments (A0, A1, A2, S);
  here they are summed directly over neighbour triplets (j, k) from their definition.
* g(r) counts: an exact float32 brute force that forms every periodic image the way the
  reference forms its ghosts (x + (+-L), src/environ.cpp:431-433), the distance and the bin the
  way GR::tally does (src/grsq.cpp:94-205), and the 2/1 counting of real and ghost partners;
  the full tally, and the inner (list) tally with its rule that bins past nbin_inner keep the
  previous full tally's counts.
* h(r) normalisation (GR ctor, src/grsq.cpp:56-70): against the closed-form volume of each
  shell, (4 pi / 3)((k+1)^3 - k^3) dr^3, and, statistically, on an ideal gas.
* The Ryckaert-Bellemans dihedral form (no reference counterpart): finite differences, and its
  reduction to the reference's cubic when C4 = C5 = 0.
"""
import dataclasses
import itertools

import numpy as np
import pytest

from nnpengine_amd import model as M
from nnpengine_amd import system as S
from oracle import pyoracle
from tests.conftest import rel_rms


def _images(x, L):
    """All 27 image shifts of every position: [27, N, 3] float64 and the shift multipliers."""
    m = np.array(list(itertools.product((-1, 0, 1), repeat=3)), np.float64)
    return x[None, :, :] + m[:, None, :] * L, m


# ---- angular / radial symmetry functions -------------------------------------------------------

def _sf_brute(x, typ, ntype, L, rc, mu, eta_r, eta_a):
    """Symmetry functions from their definition, FP64, every periodic image within rc.
    Radial: G[a, n] = sum_{j in a} exp(-eta_r (r_ij - mu_n)^2) fc(r_ij).
    Angular, element pair a <= b, exponent eta, (lambda, zeta) in (1,1), (-1,1), (1,2), (-1,2):
      G = 2^(1-zeta) sum_{j in a} sum_{k in b, k != j} (1 + lambda cos t_jik)^zeta w_j w_k,
      w = exp(-eta r^2) fc(r), fc(r) = (cos(pi r / rc) + 1) / 2."""
    N = len(x)
    nrad, neta = len(mu), len(eta_a)
    npair = ntype * (ntype + 1) // 2
    D = nrad * ntype + 4 * neta * npair
    G = np.zeros((N, D))
    img, _ = _images(x, L)
    pairs = [(a, b) for a in range(ntype) for b in range(a, ntype)]
    for i in range(N):
        d = (img - x[i]).reshape(-1, 3)
        jt = np.tile(typ, 27)
        r = np.sqrt((d * d).sum(1))
        sel = (r < rc) & (r > 0)
        d, r, jt = d[sel], r[sel], jt[sel]
        fc = 0.5 * (np.cos(np.pi * r / rc) + 1.0)
        for a in range(ntype):
            ra = r[jt == a]
            fa = fc[jt == a]
            for n in range(nrad):
                G[i, a * nrad + n] = (np.exp(-eta_r * (ra - mu[n]) ** 2) * fa).sum()
        u = d / r[:, None]
        cos = u @ u.T
        for pp, (a, b) in enumerate(pairs):
            ja, kb = np.nonzero(jt == a)[0], np.nonzero(jt == b)[0]
            for m, eta in enumerate(eta_a):
                w = np.exp(-eta * r * r) * fc
                c = cos[np.ix_(ja, kb)]
                ww = np.outer(w[ja], w[kb])
                not_self = ja[:, None] != kb[None, :]
                for t, (lam, zeta) in enumerate(((1, 1), (-1, 1), (1, 2), (-1, 2))):
                    G[i, nrad * ntype + (pp * neta + m) * 4 + t] = \
                        2.0 ** (1 - zeta) * ((1 + lam * c) ** zeta * ww * not_self).sum()
    return G


def test_sf_descriptors_equal_triplet_sums():
    """81 atoms (27 waters, L = 13.4 A, rc = 6 A): the moment-identity descriptors equal the
    direct sums over neighbour pairs of every atom, including the j = k exclusion and the
    off-diagonal products of the cos^2 term."""
    sysm = S.water_box(27, 13.4, seed=11)
    sfp = M.sf_params(6.0)
    o = pyoracle.Oracle(sysm, 1)
    o.neigh_init(6.0, 0.5, 6.0)
    o.rebuild(True)
    D = o.sf_set(sfp["rc"], sfp["mu"], sfp["eta_r"], sfp["eta_a"])
    G = o.sf_descriptor()
    x = o.get("x")
    # the oracle's rows are by atom id; its positions are refolded into the box
    Gb = _sf_brute(x, sysm.typ, sysm.ntype, float(np.float32(sysm.bl)), float(np.float32(sfp["rc"])),
                   np.asarray(sfp["mu"], np.float64), float(np.float32(sfp["eta_r"])),
                   np.asarray(sfp["eta_a"], np.float64))
    assert G.shape == Gb.shape == (sysm.natom, D)
    assert np.abs(Gb).max() > 1.0
    assert np.abs(G - Gb).max() < 1e-12 * np.abs(Gb).max()


# ---- g(r) counts, exact ------------------------------------------------------------------------

def _gr_brute_f32(x, typ, ntype, L, nbi, r_max, r_max_inner, inner):
    """GR::tally in float32 by brute force over the 27 images.  Returns (counts, nbin)."""
    f = np.float32
    x = x.astype(f)
    L = f(L)
    r_dr = f(nbi) / f(r_max_inner)
    dr = f(r_max_inner) / f(nbi)
    nbin = int(f(r_max) * r_dr)
    r_maxsq = f(f(nbin) * dr) * f(f(nbin) * dr)
    if inner:
        nlim, rsq_lim = nbi, f(r_max_inner) * f(r_max_inner)
    else:
        nlim, rsq_lim = nbin, r_maxsq
    pt = np.zeros((ntype, ntype), np.int64)
    k = 0
    for a in range(ntype):
        for b in range(a, ntype):
            pt[a, b] = pt[b, a] = k
            k += 1
    counts = np.zeros((k, nbin), np.int64)
    shift = np.array([-L, f(0), L], f)
    N = len(x)
    for code in range(27):
        ma, mb, mc = code // 9, (code // 3) % 3, code % 3
        # ghost coordinates: x + (+-L) per shifted axis, x itself on the others
        y = x.copy()
        for ax, mm in enumerate((ma, mb, mc)):
            if mm != 1:
                y[:, ax] = x[:, ax] + shift[mm]
        for i in range(N):
            d = x[i][None, :] - y  # float32
            rsq = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
            if code == 13:
                rsq[i] = np.inf
            ok = rsq < rsq_lim
            j = np.nonzero(ok)[0]
            b = (np.sqrt(rsq[j]) * r_dr).astype(np.int64)
            keep = b < nlim
            j, b = j[keep], b[keep]
            # a real partner is found from both sides (+1 each: the reference's +2 once); a ghost
            # partner from this side only
            np.add.at(counts, (pt[typ[i], typ[j]], b), 1)
    return counts, nbin


def test_gr_counts_exact_full_and_inner():
    sysm = S.water_box(81, 13.4, seed=4)
    nbi, r_max, r_inner = 250, 6.0, 5.0
    o = pyoracle.Oracle(sysm, 0)
    o.neigh_init(5.5, 0.5, r_max)
    o.rebuild(False)
    nb = o.gr_init(nbi, r_max, r_inner)
    o.gr_tally(True)
    full0, _ = o.gr_get()
    x0 = o.get("x")
    want0, nbin = _gr_brute_f32(x0, sysm.typ, sysm.ntype, sysm.bl, nbi, r_max, r_inner, False)
    assert nb == nbin
    assert full0.sum() > 10000
    assert np.array_equal(full0, want0)
    # move every atom by < skin/2 (the list stays valid), then an inner tally from the list
    rng = np.random.default_rng(3)
    x1 = (x0 + rng.uniform(-0.1, 0.1, x0.shape)).astype(np.float32).astype(np.float64)
    o.set("x", x1)
    o.update_ghost_pos()
    assert o.isvalid()
    o.gr_tally(False)
    inner1, _ = o.gr_get()
    want1, _ = _gr_brute_f32(x1, sysm.typ, sysm.ntype, sysm.bl, nbi, r_max, r_inner, True)
    assert np.array_equal(inner1[:, :nbi], want1[:, :nbi])
    # bins past nbin_inner keep the previous full tally
    assert np.array_equal(inner1[:, nbi:], full0[:, nbi:])
    assert full0[:, nbi:].sum() > 0
    # and a full tally of the moved positions refreshes them
    o.gr_tally(True)
    full1, _ = o.gr_get()
    want_f1, _ = _gr_brute_f32(x1, sysm.typ, sysm.ntype, sysm.bl, nbi, r_max, r_inner, False)
    assert np.array_equal(full1, want_f1)


# ---- h(r) normalisation --------------------------------------------------------------------------

def _gas(n, L, seed, ntype=2):
    rng = np.random.default_rng(seed)
    mt = S.MolType(1, bonds=np.zeros((0, 3)), angles=np.zeros((0, 4)), dihedrals=np.zeros((0, 5)))
    x = rng.uniform(-L / 2, L / 2, (n, 3)).astype(np.float32)
    typ = (np.arange(n) % ntype).astype(np.int32)
    return S.System(bl=float(np.float32(L)), ntype=ntype, x=x, v=np.zeros((n, 3), np.float32),
                    q=np.zeros(n, np.float32), mass=np.ones(n, np.float32), typ=typ,
                    mol_id=np.arange(n, dtype=np.int32), intra_id=np.zeros(n, np.int32),
                    mol_type=np.zeros(n, np.int32), moltypes=[mt], bond_coef=np.zeros((0, 4), np.float32),
                    angle_coef=np.zeros((0, 4), np.float32), dihedral_coef=np.zeros((0, 4), np.float32))


def test_h_normalisation_shell_volumes_ideal_gas():
    """h_k + 1 = count_k / E_k with E_k the ideal-gas expectation of the 2-per-pair count in
    shell k: N_a (N_a - 1) (a = b) or 2 N_a N_b (a != b) times (4 pi / 3)((k+1)^3 - k^3) dr^3 / L^3.
    On an ideal gas the mean of h over the bins is then 0 within its counting noise."""
    L, n = 30.0, 3000
    sysm = _gas(n, L, seed=9)
    nbi, r_max = 200, 8.0
    o = pyoracle.Oracle(sysm, 0)
    o.neigh_init(7.0, 0.5, r_max)
    o.rebuild(False)
    nb = o.gr_init(nbi, r_max, 0.0)
    assert nb == nbi
    o.gr_tally(True)
    counts, h = o.gr_get()
    dr = r_max / nbi
    k = np.arange(nbi, dtype=np.float64)
    shell = 4.0 * np.pi / 3.0 * ((k + 1) ** 3 - k ** 3) * dr ** 3
    Lf = float(np.float32(L))
    na = np.bincount(sysm.typ, minlength=2).astype(np.float64)
    for p, (a, b) in enumerate(((0, 0), (0, 1), (1, 1))):
        npairs2 = na[a] * (na[a] - 1) if a == b else 2 * na[a] * na[b]
        E = npairs2 * shell / Lf ** 3
        assert np.allclose(h[p] + 1.0, counts[p] / E, rtol=2e-6, atol=0)
        # statistics: bins past 2 A hold hundreds of counts; sum (count - E) / sqrt(sum E)
        sel = k * dr > 2.0
        z = (counts[p][sel] - E[sel]).sum() / np.sqrt(2 * E[sel].sum())
        assert abs(z) < 5.0
        assert abs(np.average(h[p][sel], weights=E[sel])) < 0.02


# ---- Ryckaert-Bellemans dihedral -----------------------------------------------------------------

def _butane_one(rb=True, phi_deg=63.0):
    s = S.butane_box(1, 40.0, seed=0)
    # a gauche-like conformer: rotate the last site about the 1-2 bond
    x = s.x.astype(np.float64)
    b = x[2] - x[1]
    b /= np.linalg.norm(b)
    v = x[3] - x[2]
    t = np.radians(180.0 - phi_deg)
    v_rot = v * np.cos(t) + np.cross(b, v) * np.sin(t) + b * (b @ v) * (1 - np.cos(t))
    x[3] = x[2] + v_rot
    s.x = x.astype(np.float32)
    return S.with_rb_dihedral(s) if rb else s


def _bonded_energy_force(sysm, xx):
    o = pyoracle.Oracle(sysm, 1)
    o.neigh_init(6.0, 0.5, 6.0)
    o.rebuild(False)
    o.set("x", xx)
    o.clear_f()
    e, _ = o.bonded_compute()
    return e, o.get("f")


def test_rb_dihedral_finite_difference():
    sysm = _butane_one(True)
    x = sysm.x.astype(np.float64)
    e, f = _bonded_energy_force(sysm, x)
    h = 1e-6
    fd = np.zeros_like(x)
    for a in range(4):
        for d in range(3):
            xp, xm = x.copy(), x.copy()
            xp[a, d] += h
            xm[a, d] -= h
            fd[a, d] = -(_bonded_energy_force(sysm, xp)[0] - _bonded_energy_force(sysm, xm)[0]) / (2 * h)
    assert np.abs(fd - f).max() < 1e-5 * np.abs(f).max()
    # the dihedral energy against its closed form at the conformer's cos(phi)
    xi, xj, xk, xl = x
    m = np.cross(xj - xi, xk - xj)
    n = np.cross(xk - xj, xl - xk)
    c = m @ n / np.sqrt((m @ m) * (n @ n))
    mt = sysm.moltypes[0]
    no_dih = dataclasses.replace(sysm, moltypes=[S.MolType(4, bonds=mt.bonds, angles=mt.angles,
                                                           dihedrals=np.zeros((0, 5)))], dihedral_rb=None)
    e_no_dih, _ = _bonded_energy_force(no_dih, x)
    C = S.rb_butane_coef()[0].astype(np.float64)
    assert abs((e - e_no_dih) - sum(C[q] * c ** q for q in range(6))) < 1e-9 * max(1.0, abs(e))


def test_rb_reduces_to_reference_cubic():
    """C4 = C5 = 0: the quintic path equals the reference's cubic (src/dihedral.cpp:27-35) for the
    same torsion written as U = c0 + c1 cos + c2 cos 2p + c3 cos 3p."""
    c0, c1, c2, c3 = 0.0, 0.7055, -0.1355, 1.5725
    C = np.array([[c0 - c2, c1 - 3 * c3, 2 * c2, 4 * c3, 0.0, 0.0]], np.float32)
    cub = _butane_one(False)
    rb = cub.copy()
    rb.dihedral_rb = C
    x = cub.x.astype(np.float64)
    e1, f1 = _bonded_energy_force(cub, x)
    e2, f2 = _bonded_energy_force(rb, x)
    assert abs(e1 - e2) < 1e-6 * max(1.0, abs(e1))
    assert rel_rms(f2, f1) < 1e-6
