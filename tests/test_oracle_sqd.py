"""Pinning of the oracle's SQD (direct k-space S(q), src/grsq.cpp:362-606) restatement.

Checks that do not retype the oracle's loops:
  * vector set and bins: a vectorised numpy enumeration of the integer lattice (the half space
    l > 0 | (l = 0, k > 0) | (l = k = 0, h > 0), n^2 <= cuti2) with the bin decided by
    searchsorted against the float32 edges i * dq -- per-bin counts must agree exactly, and the
    total equals half the lattice points of the ball;
  * closed form 1: a simple-cubic lattice of n^3 identical atoms filling the box has
    |rho(G)|^2 = N^2 when h, k, l are all multiples of n and 0 otherwise, so
    S(bin) = N m_bin / count_bin - 1 (m = lattice vectors in the bin);
  * closed form 2: two atoms of different elements: S_00 = S_11 = 0 and
    S_01 = 2 mean_G cos(G . (x_0 - x_1)).
"""
import numpy as np

from oracle import pyoracle


def half_space(bl, q_max, nq):
    """Integer vectors, |G| (float32, the reference's arithmetic) and bins, by z slices."""
    two_pi = np.float32(np.float32(np.pi) * np.float32(2))
    g = np.float32(two_pi / np.float32(bl))
    qg = np.float32(np.float32(q_max) / g)
    K = int(qg)
    cuti2 = int(np.float32(qg * qg))
    dq = np.float32(np.float32(q_max) / np.float32(nq))
    edges = np.arange(1, nq, dtype=np.float32) * dq
    r = np.arange(-K, K + 1)
    hh, kk = np.meshgrid(r, r, indexing="ij")
    hh, kk = hh.ravel(), kk.ravel()
    out = []
    for l in range(0, K + 1):
        n2 = hh * hh + kk * kk + l * l
        keep = n2 <= cuti2
        if l == 0:
            keep &= (kk > 0) | ((kk == 0) & (hh > 0))
        h, k, n = hh[keep], kk[keep], n2[keep]
        gs = np.sqrt(n.astype(np.float32)) * g
        b = np.searchsorted(edges, gs, side="right")
        out.append(np.stack([h, k, np.full_like(h, l), b], 1))
    v = np.concatenate(out)
    gs = np.sqrt((v[:, 0] ** 2 + v[:, 1] ** 2 + v[:, 2] ** 2).astype(np.float32)) * g
    return v[:, :3], gs, v[:, 3], K, cuti2


def test_sqd_vector_set_and_bins():
    bl, q_max, nq = 19.73, 6.0, 30
    x = np.zeros((3, 3), np.float32)
    t = np.array([0, 1, 1])
    qq, qq_ref, sq, cnt = pyoracle.sqd(x, t, 2, bl, nq, q_max)
    hkl, gs, b, K, cuti2 = half_space(bl, q_max, nq)
    np.testing.assert_array_equal(cnt, np.bincount(b, minlength=nq))
    # half the nonzero lattice points of the ball within the cube
    r = np.arange(-K, K + 1)
    n2 = r[:, None, None] ** 2 + r[None, :, None] ** 2 + r[None, None, :] ** 2
    assert cnt.sum() * 2 == int(((n2 <= cuti2) & (n2 > 0)).sum())
    mean = np.bincount(b, weights=gs.astype(np.float64), minlength=nq) / np.maximum(cnt, 1)
    np.testing.assert_allclose(qq, mean, rtol=1e-12, atol=0)
    nz = cnt > 0
    np.testing.assert_allclose(qq_ref[nz], qq[nz], rtol=1e-5)
    # every bin's mean lies inside its float edges
    dq = q_max / nq
    assert np.all(qq[nz][1:] >= np.arange(nq)[nz][1:] * dq - 1e-6)


def test_sqd_lattice_closed_form():
    n, a = 4, 2.0
    bl = n * a
    ijk = np.array(np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")).reshape(3, -1).T
    x = ((ijk + 0.25) * a - bl / 2).astype(np.float32)   # exact in float32
    N = len(x)
    q_max, nq = 8.0, 16
    qq, _, sq, cnt = pyoracle.sqd(x, np.zeros(N, np.int32), 1, bl, nq, q_max)
    hkl, gs, b, _, _ = half_space(bl, q_max, nq)
    bragg = np.all(hkl % n == 0, axis=1)
    m = np.bincount(b, weights=bragg.astype(np.float64), minlength=nq)
    nz = cnt > 0
    expect = N * m[nz] / cnt[nz] - 1.0
    assert m.sum() > 0
    np.testing.assert_allclose(sq[nz, 0], expect, atol=1e-9 * N, rtol=0)


def test_sqd_two_atoms_closed_form():
    bl, q_max, nq = 10.0, 5.0, 20
    x = np.array([[0.3, -1.2, 2.0], [1.7, 0.4, -3.1]], np.float32)
    qq, _, sq, cnt = pyoracle.sqd(x, np.array([0, 1]), 2, bl, nq, q_max)
    hkl, gs, b, _, _ = half_space(bl, q_max, nq)
    d = x[0].astype(np.float64) - x[1].astype(np.float64)
    c = np.cos(2 * np.pi / np.float64(np.float32(bl)) * (hkl @ d))
    nz = cnt > 0
    mean_c = np.bincount(b, weights=c, minlength=nq)[nz] / cnt[nz]
    np.testing.assert_allclose(sq[nz, 1], 2 * mean_c, atol=1e-10)
    np.testing.assert_allclose(sq[nz, 0], 0.0, atol=1e-12)
    np.testing.assert_allclose(sq[nz, 2], 0.0, atol=1e-12)
    # a planted mistake in the oracle (the cross term's sign) is caught by this closed form
    assert np.abs(2 * mean_c).max() > 1e-3
