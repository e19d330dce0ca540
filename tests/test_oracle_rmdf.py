"""Pinning of the oracle's RMDF restatement (src/grsq.cpp:899-1377).  The reference's RMDF lives in
grsq.cpp, which includes the TensorFlow C header and cannot be compiled here, so the oracle is
pinned by what the mathematics fixes:
  * the total S(q) is the Faber-Ziman combination of the partials;
  * DA/RMD (gamma = 1) makes the relaxed residual the plain residual, and the Q-space energy is
    N/4 sum_q W (S - S_exp)^2;
  * FMIRL (gamma < 1) relaxes geometrically on a fixed g(r);
  * a zero residual leaves the tables at their base;
  * one q point gives table updates c j0(qr) (potential) and c (j0(qr) - cos(qr)) / r^2 (force),
    and the force update is -(dU/dr) / r of the potential update (finite differences);
  * on the NNP tables the update goes to special class 0 and half of it to class 3.
"""
import numpy as np
import pytest

from nnpengine_amd import model as M
from nnpengine_amd import system as S
from oracle import pyoracle

RC, NBIN, QMAX, NQ = 9.0, 450, 12.0, 120


def _setup(sysm, prec=1, target=0):
    o = pyoracle.Oracle(sysm, prec)
    if target == 0:
        o.neigh_init(RC, 0.5, RC)
        o.rebuild(False)
        ft, ut = S.lj_table(sysm, RC, NBIN, S.WATER_LJ)
        o.pair_set(RC, ft, ut)
        o.gr_init(NBIN, RC, RC)
        base = (np.asarray(ft, np.float64).reshape(sysm.npair, -1), np.asarray(ut, np.float64).reshape(sysm.npair, -1))
    else:
        rc = 6.0
        o.neigh_init(rc, 0.5, rc)
        o.rebuild(False)
        nbin = 300
        o.nnp_tbl_init(rc, nbin, *M.tbl_coefficients(sysm.npair, 3, rc))
        o.gr_init(nbin, rc, rc)
        _, _, tp, tfp = o.nnp_tbl_tables()
        base = (tfp, tp)
    o.gr_tally(False)
    return o, base


def _fz_weights(sysm, ff):
    """Faber-Ziman weights (2 - delta_ab) c_a c_b f_a f_b / (sum_a c_a f_a)^2 per pair and q."""
    c = np.bincount(sysm.typ, minlength=sysm.ntype) / sysm.natom
    norm = (c[None, :] * ff).sum(1) ** 2
    w = []
    for a in range(sysm.ntype):
        for b in range(a, sysm.ntype):
            w.append((1 if a == b else 2) * c[a] * c[b] * ff[:, a] * ff[:, b] / norm)
    return np.array(w)  # [npair][nq]


def _data(sysm, sq_exp, weight, ff):
    q = (np.arange(NQ) + 0.5) * QMAX / NQ
    rows = np.concatenate([q[:, None], sq_exp[:, None], weight[:, None], ff], axis=1)
    return rows[None].astype(np.float32)


def _ff(sysm, seed=0):
    rng = np.random.default_rng(seed)
    q = (np.arange(NQ) + 0.5) * QMAX / NQ
    # element form factors that differ per element and vary with q
    return np.stack([(1.0 + 0.5 * t) * np.exp(-(0.02 + 0.01 * t) * q * q) + 0.1 * rng.random(NQ)
                     for t in range(sysm.ntype)], 1).astype(np.float32)


def test_total_sq_is_faber_ziman(water_c0):
    sysm = water_c0
    o, _ = _setup(sysm)
    partial = o.sq(NQ, 0.0, QMAX)
    ff = _ff(sysm)
    o.rmdf_init(_data(sysm, np.zeros(NQ), np.ones(NQ), ff), QMAX, 1.0, 1.0)
    _, sq, _, _ = o.rmdf_step(QMAX)
    expect = (_fz_weights(sysm, ff.astype(np.float64)) * partial).sum(0)
    assert np.max(np.abs(sq[0] - expect)) < 1e-5 * max(1.0, np.max(np.abs(expect)))


def test_da_residual_energy_and_force_q(water_c0):
    sysm = water_c0
    o, _ = _setup(sysm)
    partial = o.sq(NQ, 0.0, QMAX)
    ff = _ff(sysm, 1)
    rng = np.random.default_rng(2)
    w = rng.random(NQ).astype(np.float32)
    strength = 3.0
    ffd = ff.astype(np.float64)
    sfz = (_fz_weights(sysm, ffd) * partial).sum(0)
    sexp = (sfz + 0.05 * np.sin(np.arange(NQ))).astype(np.float32)
    o.rmdf_init(_data(sysm, sexp, w, ff), QMAX, strength, 1.0)
    erg, sq, ds, fq = o.rmdf_step(QMAX)
    d = sq[0] - sexp
    assert np.allclose(ds[0], d, rtol=0, atol=1e-12)
    w_eff = w.astype(np.float64) * strength
    assert erg == pytest.approx(sysm.natom / 4 * np.sum(w_eff * d * d), rel=1e-10)
    # force_q of pair ab: residual x weight x f_a f_b / (sum c f)^2
    c = np.bincount(sysm.typ, minlength=sysm.ntype) / sysm.natom
    norm = (c[None, :] * ffd).sum(1) ** 2
    p = 0
    for a in range(sysm.ntype):
        for b in range(a, sysm.ntype):
            assert np.allclose(fq[p], d * w_eff * ffd[:, a] * ffd[:, b] / norm, rtol=1e-9, atol=1e-12)
            p += 1


def test_fmirl_relaxation(water_c0):
    sysm = water_c0
    o, _ = _setup(sysm)
    ff = _ff(sysm, 4)
    sexp = np.full(NQ, 0.9, np.float32)
    g = float(np.float32(0.3))  # gamma is passed as FP32
    o.rmdf_init(_data(sysm, sexp, np.ones(NQ, np.float32), ff), QMAX, 1.0, g)
    _, sq1, ds1, fq1 = o.rmdf_step(QMAX)
    _, sq2, ds2, fq2 = o.rmdf_step(QMAX)
    d = sq1[0] - sexp
    assert np.allclose(sq1, sq2, rtol=0, atol=0)
    assert np.allclose(ds1[0], g * d, rtol=1e-12, atol=1e-14)
    assert np.allclose(ds2[0], g * d * (2 - g), rtol=1e-12, atol=1e-14)
    assert np.allclose(fq2, fq1 * (2 - g), rtol=1e-12, atol=1e-14)


def test_zero_residual_keeps_base_tables(water_c0):
    sysm = water_c0
    o, base = _setup(sysm)
    ff = _ff(sysm, 5)
    o.rmdf_init(_data(sysm, np.zeros(NQ, np.float32), np.ones(NQ, np.float32), ff), QMAX, 1.0, 1.0)
    _, sq, _, _ = o.rmdf_step(QMAX)
    o2, _ = _setup(sysm)
    o2.rmdf_init(_data(sysm, sq[0].astype(np.float32), np.ones(NQ, np.float32), ff), QMAX, 1.0, 1.0)
    erg, _, ds, fq = o2.rmdf_step(QMAX)
    # S_exp is S(q) rounded to FP32: residuals at that rounding only
    assert np.max(np.abs(ds)) < 1e-6 and np.max(np.abs(fq)) < 1e-5
    f, u = o2.rmdf_tables(base[0].shape)
    scale = np.max(np.abs(base[0][:, :NBIN]))
    assert np.max(np.abs(f[:, :NBIN] - base[0][:, :NBIN])) < 1e-6 * scale


@pytest.mark.parametrize("k", [7, 60])
def test_single_q_closed_form(water_c0, k):
    sysm = water_c0
    o, base = _setup(sysm)
    ff = _ff(sysm, 6)
    w = np.zeros(NQ, np.float32)
    w[k] = 1.0
    o.rmdf_init(_data(sysm, np.zeros(NQ, np.float32), w, ff), QMAX, 1.0, 1.0)
    _, _, _, fq = o.rmdf_step(QMAX, potential=True)
    assert np.count_nonzero(fq[:, np.arange(NQ) != k]) == 0
    f, u = o.rmdf_tables(base[0].shape)
    r = (np.arange(NBIN) + 0.5) * (RC / NBIN)
    qk = (k + 0.5) * QMAX / NQ
    x = qk * r
    j0 = np.sin(x) / x
    du = u[:, :NBIN] - base[1][:, :NBIN]
    df = f[:, :NBIN] - base[0][:, :NBIN]
    for p in range(sysm.npair):
        c = fq[p, k]
        assert np.allclose(du[p], c * j0, rtol=1e-6, atol=1e-9 * abs(c))
        assert np.allclose(df[p], c * (j0 - np.cos(x)) / r ** 2, rtol=1e-6, atol=1e-9 * abs(c))
        # force table = -(dU/dr) / r: central differences of the potential update
        dr = RC / NBIN
        fd = -(du[p, 2:] - du[p, :-2]) / (2 * dr) / r[1:-1]
        err = np.abs(fd - df[p, 1:-1]) / np.max(np.abs(df[p, 1:-1]))
        assert np.max(err[20:]) < 2e-3


def test_nnp_target_special_classes(water_c0):
    sysm = water_c0
    o, (tfp0, tp0) = _setup(sysm, target=1)
    ff = _ff(sysm, 8)
    rng = np.random.default_rng(9)
    o.rmdf_init(_data(sysm, (0.5 + rng.random(NQ)).astype(np.float32), np.ones(NQ, np.float32), ff), QMAX, 1.0, 1.0,
                target=1)
    _, _, _, fq = o.rmdf_step(QMAX, potential=True)
    f, u = o.rmdf_tables(tfp0.shape)
    nr = 300
    dF = f - tfp0
    dU = u - tp0
    assert np.max(np.abs(dF[:, 0, :nr])) > 0
    assert np.array_equal(dF[:, 1:3], np.zeros_like(dF[:, 1:3]))
    assert np.array_equal(dU[:, 1:3], np.zeros_like(dU[:, 1:3]))
    # (base + update) - base loses the update's low bits where the base is large (r -> 0)
    tolF = 1e-15 * (np.abs(tfp0[:, 0, :nr]) + np.abs(tfp0[:, 3, :nr])) + 1e-12 * np.max(np.abs(dF))
    tolU = 1e-15 * (np.abs(tp0[:, 0, :nr]) + np.abs(tp0[:, 3, :nr])) + 1e-12 * np.max(np.abs(dU))
    assert np.all(np.abs(dF[:, 3, :nr] - 0.5 * dF[:, 0, :nr]) <= tolF + 1e-9 * np.abs(dF[:, 0, :nr]))
    assert np.all(np.abs(dU[:, 3, :nr] - 0.5 * dU[:, 0, :nr]) <= tolU + 1e-9 * np.abs(dU[:, 0, :nr]))
    # the class-0 update is force_q . kernel on the g(r) grid
    r = (np.arange(nr) + 0.5) * (6.0 / nr)
    q = (np.arange(NQ) + 0.5) * QMAX / NQ
    x = np.outer(r, q)
    j0 = np.sin(x) / x
    assert np.all(np.abs(dU[:, 0, :nr] - fq @ j0.T) <= tolU + 1e-9 * np.abs(dU[:, 0, :nr]))
