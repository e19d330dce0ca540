"""NNP model parameters: symmetry-function settings, seeded per-element MLP weights, the
per-element descriptor standardisation of the network input, and the reference's NNPOTENTIAL_TBL
coefficient arrays (the layout its weight file holds, src/pair_nnp.cpp:505-517)."""
from __future__ import annotations

import numpy as np

NRAD = 16
NETA = 4


def sf_params(rc=6.0):
    """16 Gaussian radial centres from 0.8 A to rc - 0.2 A with width equal to their spacing, and
    4 angular exponents spanning short to long range."""
    mu = np.linspace(0.8, rc - 0.2, NRAD).astype(np.float32)
    sp = float(mu[1] - mu[0])
    eta_r = np.float32(0.5 / sp ** 2)
    eta_a = np.array([0.02, 0.06, 0.15, 0.4], np.float32)
    return dict(rc=float(rc), mu=mu, eta_r=float(eta_r), eta_a=eta_a)


def sf_dim(ntype):
    return NRAD * ntype + NETA * 4 * ntype * (ntype + 1) // 2


def mlp_init(widths, ntype, seed):
    """Weights N(0, 1/fan_in), zero biases; flat float32 array, per element then per layer
    W [in][out] row-major followed by b [out]."""
    rng = np.random.default_rng(seed)
    parts = []
    for _ in range(ntype):
        for l in range(len(widths) - 1):
            fi, fo = widths[l], widths[l + 1]
            parts.append(rng.standard_normal((fi, fo)) / np.sqrt(fi))
            parts.append(np.zeros(fo))
    return np.concatenate([p.reshape(-1) for p in parts]).astype(np.float32)


def standardisation(G, typ, ntype, floor=0.5, abs_floor=0.1):
    """Per-element input standardisation x = (G - shift) * scale of the network: shift = mean,
    scale = 1 / max(std, floor * rms, abs_floor * RMS), rms the component's root mean square and
    RMS that of all the element's components.  The first floor keeps components whose spread is
    tiny against their magnitude (e.g. on a lattice snapshot) from amplifying descriptor
    rounding; the second keeps components that are nearly zero on the snapshot (short-range
    radial terms of a lattice, ~1e-10 there but O(0.1) in the liquid) from scales of 1e10.
    Returns shift, scale as [ntype, D] float32."""
    D = G.shape[1]
    shift = np.zeros((ntype, D))
    scale = np.ones((ntype, D))
    for t in range(ntype):
        g = G[typ == t].astype(np.float64)
        if len(g):
            m = g.mean(0)
            sd = g.std(0)
            rms = np.sqrt(m * m + sd * sd)
            big = np.sqrt((rms * rms).mean())
            shift[t] = m
            scale[t] = 1.0 / np.maximum(np.maximum(np.maximum(sd, floor * rms), abs_floor * big), 1e-12)
    return shift.astype(np.float32), scale.astype(np.float32)


def tbl_coefficients(npair, seed, cutoff):
    """Seeded NNPOTENTIAL_TBL coefficients in the weight-file layout: coef_p [npair][4][6]
    (eps, sigma, c8, c10, c12, qij for each special class) and coef_w/d/r [npair][80]."""
    rng = np.random.default_rng(seed)
    cp = np.zeros((npair, 4, 6), np.float32)
    for p in range(npair):
        for s in range(4):
            c = rng.standard_normal(3)
            c /= np.linalg.norm(c)
            cp[p, s] = [0.1 + 0.1 * rng.random(), 2.8 + 0.6 * rng.random(), abs(c[0]) * 0.5, abs(c[1]) * 0.2,
                        abs(c[2]) * 0.1, 0.2 * rng.standard_normal()]
    cw = (0.2 + 0.8 * rng.random((npair, 80))).astype(np.float32)
    cd = (-(0.5 + 3.0 * rng.random((npair, 80)))).astype(np.float32)
    cr = (0.8 + (cutoff - 1.0) * rng.random((npair, 80))).astype(np.float32)
    return cp, cw, cd, cr
