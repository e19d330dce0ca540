"""nnpe_bonded_compute_noalloc (the bonded pass with its partial sums in the library's fixed device
array, no allocation) against nnpe_bonded_compute on the same state: forces and the bond,
angle and dihedral energy/virial sums bit for bit, on two calls in a row -- a 900-atom water box,
then a smaller 256-site butane box with dihedrals (a second context on the same device, which
shares the array and finds the first call's partials in it)."""
import numpy as np
import pytest

from nnpengine_amd import system as S
from tests.golden.make_golden import make_system

pytestmark = pytest.mark.gpu

TERMS = ("bond", "angle", "dihedral")


def _perturbed(sysm, seed):
    s = sysm.copy()
    rng = np.random.default_rng(seed)
    s.x = (s.x + rng.uniform(-0.05, 0.05, s.x.shape)).astype(np.float32)
    return s


def _both_paths(sysm):
    from nnpengine_amd.engine import Engine
    e = Engine(sysm)
    e.neigh_init(9.0, 0.5, 9.0)
    e.build()
    e.bonded_set_from_system()
    out = {}
    for name, call in (("alloc", e.bonded_compute), ("noalloc", e.bonded_compute_noalloc)):
        e.clear_force()
        e.energy_clear()
        call(True)
        en = e.energy()
        out[name] = (e.get("f"), np.array([en[t] for t in TERMS]))
    e.close()
    return out


def test_noalloc_bonded_equals_bonded_compute_twice():
    water = _perturbed(S.water_box(300, 21.0, seed=3), 17)
    butane = _perturbed(make_system("butane"), 18)
    assert butane.natom < water.natom
    for sysm in (water, butane):
        print(f"[noalloc] natom={sysm.natom}", flush=True)
        r = _both_paths(sysm)
        f0, e0 = r["alloc"]
        f1, e1 = r["noalloc"]
        assert np.abs(f0).max() > 0
        assert np.array_equal(f0, f1)
        assert np.array_equal(e0, e1)
        assert e0[0, 0] > 0
