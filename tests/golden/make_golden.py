"""Writes the golden fixtures under tests/golden/ from the reference's own code.

The reference C++ (TF-free sources: ENVIRON, NEIGHLIST, LJTableCoulDSF, BOND, ANGLE, DIHEDRAL,
VERLET, ...) is compiled from the reference tree by oracle/Makefile into oracle/_ref/ref_harness,
which runs the reference's classes with its 27 OpenMP threads on seeded inputs.  This script
needs the reference tree (run it where the reference sources are checked out):

    python tests/golden/make_golden.py

Fixtures (inputs are regenerated from the seeds recorded in each file):
  water_c0_snap.npz     768-atom water, rc 9, skin 0.5: forces, energy, virial, refolded x,
                        canonical list key set (sha256 + count), ghost count
  water_small_list.npz  243-atom water, rc 6, skin 0.5: the full canonical key set
  water_c0_md.npz       40 velocity-Verlet steps (dt 0.5 fs, skin-triggered rebuilds): x, v
  butane_snap.npz       64 united-atom butane chains (dihedrals, 1-4 specials): forces, energy
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from nnpengine_amd import system as S  # noqa: E402
from oracle import pyoracle, refrun  # noqa: E402

CASES = {
    "water_c0": dict(builder="water", nmol=256, bl=19.73, seed=1, rc=9.0, skin=0.5),
    "water_small": dict(builder="water", nmol=81, bl=13.4, seed=4, rc=6.0, skin=0.5),
    "butane": dict(builder="butane", nmol=64, bl=22.0, seed=5, rc=9.0, skin=0.5),
}


def make_system(case):
    c = CASES[case]
    if c["builder"] == "water":
        return S.water_box(c["nmol"], c["bl"], seed=c["seed"])
    return S.butane_box(c["nmol"], c["bl"], seed=c["seed"])


def tables(case, sysm):
    c = CASES[case]
    params = S.WATER_LJ if c["builder"] == "water" else S.butane_lj()
    nbin = int(round(c["rc"] / 0.01))
    return S.lj_table(sysm, c["rc"], nbin, params)


def key_hash(keys):
    return hashlib.sha256(np.ascontiguousarray(keys, np.int64).tobytes()).hexdigest()


def main():
    refrun.build()
    assert refrun.available(), "reference harness not built"
    # snapshots
    for case, fname in (("water_c0", "water_c0_snap.npz"), ("butane", "butane_snap.npz")):
        c = CASES[case]
        sysm = make_system(case)
        ft, ut = tables(case, sysm)
        r = refrun.run(sysm, c["rc"], c["skin"], c["rc"], ft, ut, mode="snap")
        keys = pyoracle.canonical(r["list"])
        np.savez_compressed(os.path.join(HERE, fname), f=r["f"], x=r["x"], erg=r["erg"], virial=r["virial"],
                            nghost=r["nghost"], nkeys=len(keys), keys_sha256=key_hash(keys))
        print(case, "entries", int(r["nentries"]), "keys", len(keys), "erg", r["erg"])
    # full key set of a small system
    case = "water_small"
    c = CASES[case]
    sysm = make_system(case)
    ft, ut = tables(case, sysm)
    r = refrun.run(sysm, c["rc"], c["skin"], c["rc"], ft, ut, mode="snap")
    keys = pyoracle.canonical(r["list"])
    np.savez_compressed(os.path.join(HERE, "water_small_list.npz"), keys=keys, f=r["f"], erg=r["erg"])
    print(case, "keys", len(keys))
    # short trajectory
    case = "water_c0"
    c = CASES[case]
    sysm = make_system(case)
    ft, ut = tables(case, sysm)
    r = refrun.run(sysm, c["rc"], c["skin"], c["rc"], ft, ut, mode="md", nsteps=40, dt=0.5)
    np.savez_compressed(os.path.join(HERE, "water_c0_md.npz"), x=r["x"], v=r["v"], nbuild=r["nbuild"],
                        erg=r["erg"])
    print("md nbuild", r["nbuild"])


if __name__ == "__main__":
    main()
