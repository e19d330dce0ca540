"""Serving form (nnpe_serve_*): independent water boxes of different sizes share one resident
arena and advance in the same launches.  Each result must be the same alone and in the mix (the
per-atom arithmetic of a box does not depend on the other boxes, so bit for bit), the same as the
single-box engine's MD loop on that box, and the multi-image neighbour list of a box smaller than
twice the list reach (L/2 < rc + skin < L, which the reference's single-sided ghosts cannot
handle) must equal a brute-force search over the 26 neighbouring images."""
import itertools

import numpy as np
import pytest

from nnpengine_amd import model as M
from nnpengine_amd import system as S
from oracle import pyoracle

pytestmark = pytest.mark.gpu

RC, SKIN, RC_NNP, STEPS = 9.0, 2.0, 6.0, 40
WIDTHS = [80, 64, 64, 1]


def _box(nmol, seed):
    return S.water_box(nmol, 62.1 * (nmol / 8000.0) ** (1.0 / 3.0), seed=seed)


def _network():
    ref = _box(216, 5)
    sfp = M.sf_params(RC_NNP)
    o = pyoracle.Oracle(ref, 1)
    o.neigh_init(RC, 0.5, RC)
    o.rebuild(True)
    o.sf_set(sfp["rc"], sfp["mu"], sfp["eta_r"], sfp["eta_a"])
    shift, scale = M.standardisation(o.sf_descriptor(), ref.typ, ref.ntype)
    return ref, M.mlp_init(WIDTHS, ref.ntype, 17), shift, scale


def _serve(ref, par, shift, scale, max_boxes=8):
    from nnpengine_amd.engine import Serve
    return Serve(ref, WIDTHS, par, shift, scale, precision="bf16", max_atoms=20000, max_boxes=max_boxes, pair_cutoff=RC,
                 skin=SKIN, nnp_cutoff=RC_NNP, rebuild_every=10, dt=0.5, fp64=True)


def _drain(srv, want):
    out = {}
    for _ in range(1000):
        for rid in srv.advance():
            out[rid] = srv.result(rid)
        if len(out) == want:
            return out
    raise AssertionError("requests did not finish")


def test_serve_alone_equals_mix():
    ref, par, shift, scale = _network()
    boxes = [_box(64, 1), _box(200, 2), _box(343, 3)]
    print("[serve] boxes", [(b.natom, round(b.bl, 2)) for b in boxes], flush=True)
    alone = []
    for b in boxes:
        srv = _serve(ref, par, shift, scale)
        rid = srv.submit(b, STEPS)
        alone.append(_drain(srv, 1)[rid])
        srv.close()
    # mix: the first box starts alone, the others arrive one chunk later, so the arena is laid
    # out again when they are admitted and again when the first box leaves
    srv = _serve(ref, par, shift, scale)
    ids = [srv.submit(boxes[0], STEPS)]
    got = {}
    for rid in srv.advance():
        got[rid] = srv.result(rid)
    ids += [srv.submit(b, STEPS) for b in boxes[1:]]
    for _ in range(100):
        for rid in srv.advance():
            got[rid] = srv.result(rid)
        if len(got) == 3:
            break
    st = srv.stats()
    print("[serve] mix stats", st, flush=True)
    assert st["relayouts"] >= 3
    for k, rid in enumerate(ids):
        x, v = got[rid]
        print(f"[serve] box {k}: max |dx| {np.abs(x - alone[k][0]).max():.3e} max |dv| {np.abs(v - alone[k][1]).max():.3e}")
        assert np.array_equal(x, alone[k][0]) and np.array_equal(v, alone[k][1])
    srv.close()


def test_serve_matches_single_box_engine():
    from nnpengine_amd.engine import Engine
    ref, par, shift, scale = _network()
    b = _box(200, 2)
    srv = _serve(ref, par, shift, scale, max_boxes=1)
    rid = srv.submit(b, STEPS)
    xs, vs = _drain(srv, 1)[rid]
    srv.close()
    e = Engine(b)
    ft, ut = S.lj_table(b, RC, 900, S.WATER_LJ)
    sfp = M.sf_params(RC_NNP)
    e.neigh_init(RC, SKIN, RC, RC_NNP)
    e.pair_table(RC, ft, ut)
    e.bonded_set_from_system()
    e.sf_set(sfp["rc"], sfp["mu"], sfp["eta_r"], sfp["eta_a"])
    e.mlp_set(WIDTHS, par, 0.5, "bf16", shift, scale)
    e.md_run(STEPS, dt=0.5, fp64=True, rebuild_every=10, pair=True, bonded=True, nnp=2, use_graph=True)
    xe = e.get("x")
    ve = e.get("v")
    dx = np.abs(xs - xe).max()
    dv = np.abs(vs - ve).max()
    print(f"[serve] vs engine: max |dx| {dx:.3e} max |dv| {dv:.3e}", flush=True)
    assert dx == 0.0 and dv == 0.0


def test_small_box_multi_image_list():
    from nnpengine_amd.engine import Engine
    b = _box(64, 4)
    comm = RC + SKIN
    assert b.bl / 2 < comm < b.bl
    e = Engine(b)
    e.neigh_init(RC, SKIN, RC, RC_NNP)
    e.build()
    ent = e.list_entries()
    ai, aj, code = ent[:, 0], ent[:, 1], ent[:, 2]
    x = e.get("x").astype(np.float32)  # refolded positions by atom id
    L = np.float32(b.bl)
    shift = {0: np.float32(0.0), 1: L, -1: -L}
    cut = np.float32(RC + SKIN) * np.float32(RC + SKIN)
    # the list's entries within the cutoff, as (i, j, image code)
    xi, xj = x[ai], x[aj]
    img = np.stack([code // 9 - 1, (code // 3) % 3 - 1, code % 3 - 1], 1)
    pj = xj.copy()
    for d in range(3):
        for s_ in (-1, 1):
            m = img[:, d] == s_
            pj[m, d] = (xj[m, d] + shift[s_]).astype(np.float32)
    dd = (xi - pj).astype(np.float32)
    r2 = ((dd[:, 0] * dd[:, 0] + dd[:, 1] * dd[:, 1]).astype(np.float32) + dd[:, 2] * dd[:, 2]).astype(np.float32)
    got = {(int(a), int(c), int(k)) for a, c, k, r in zip(ai, aj, code, r2) if r < cut}
    # brute force over all pairs and the 27 images (the image at zero shift excludes i == j)
    want = set()
    n = b.natom
    for a, bb, c in itertools.product((-1, 0, 1), repeat=3):
        k = (a + 1) * 9 + (bb + 1) * 3 + (c + 1)
        pj = x.copy()
        for d, s_ in enumerate((a, bb, c)):
            pj[:, d] = (x[:, d] + shift[s_]).astype(np.float32)
        dd = (x[:, None, :] - pj[None, :, :]).astype(np.float32)
        r2 = ((dd[..., 0] * dd[..., 0] + dd[..., 1] * dd[..., 1]).astype(np.float32) + dd[..., 2] * dd[..., 2]).astype(np.float32)
        ii, jj = np.nonzero(r2 < cut)
        for i_, j_ in zip(ii, jj):
            if k == 13 and i_ == j_:
                continue
            want.add((int(i_), int(j_), k))
    print(f"[serve] small box L={b.bl:.2f}: {len(want)} directed pairs within {RC + SKIN} A", flush=True)
    assert got == want
