"""Runs the reference harness (oracle/_ref/ref_harness, built from the reference sources by
oracle/Makefile) on a System.  TEST INFRASTRUCTURE ONLY (fixtures and the CPU baseline)."""
from __future__ import annotations

import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
HARNESS = os.path.join(HERE, "_ref", "ref_harness")


def available():
    return os.path.exists(HARNESS)


def build():
    """Compile the harness when the reference tree is present (machines without it use the
    prebuilt files)."""
    if os.path.isdir("/root/reference/src"):
        subprocess.check_call(["make", "-s", "-C", HERE, "ref"])


def write_inputs(d, sysm, pair_cutoff, skin, gr_cutoff, table_force, table_pot, nsteps=0, dt=1.0,
                 angle_params=None, dihedral_params=None):
    """Write the reference-format configuration, molecule, table and harness-params files."""
    from nnpengine_amd import system as S
    mols = []
    for k, mt in enumerate(sysm.moltypes):
        p = os.path.join(d, f"mol{k}.txt")
        S.write_molecule(p, mt)
        mols.append(p)
    cfg = os.path.join(d, "config.txt")
    S.write_config(cfg, sysm, mols)
    tbl = os.path.join(d, "table.txt")
    S.write_table(tbl, sysm, pair_cutoff, table_force, table_pot)
    vel = os.path.join(d, "v.bin")
    np.ascontiguousarray(sysm.v, np.float32).tofile(vel)
    lines = [cfg, f"{pair_cutoff:.9g} {skin:.9g} {gr_cutoff:.9g}", f"{table_force.shape[1] - 1} {tbl}"]
    bc = sysm.bond_coef
    lines.append(str(len(bc)))
    for row in bc:
        lines.append(f"{row[0]:.9g} {row[3] / 2:.9g}")
    # ANGLE::setParam(i, theta0_deg, ka, kb, kc); callers pass the original parameters
    ap = angle_params if angle_params is not None else (sysm.angle_params or [])
    lines.append(str(len(ap)))
    for row in ap:
        lines.append(" ".join(f"{v:.9g}" for v in row))
    dp = dihedral_params if dihedral_params is not None else (sysm.dihedral_params or [])
    lines.append(str(len(dp)))
    for row in dp:
        lines.append(" ".join(f"{v:.9g}" for v in row))
    lines.append(vel)
    lines.append(f"{nsteps} {dt:.9g}")
    prm = os.path.join(d, "params.txt")
    with open(prm, "w") as fh:
        fh.write("\n".join(lines) + "\n")
    return prm


def run(sysm, pair_cutoff, skin, gr_cutoff, table_force, table_pot, mode="snap", nsteps=0, dt=1.0,
        angle_params=None, dihedral_params=None, workdir=None, env=None):
    """Run the harness; returns a dict with x, v, f (by atom id, float64), the directed list
    entries [n,4] and the summary values."""
    d = workdir or tempfile.mkdtemp(prefix="nnpe_ref_")
    prm = write_inputs(d, sysm, pair_cutoff, skin, gr_cutoff, table_force, table_pot, nsteps, dt,
                       angle_params, dihedral_params)
    out = os.path.join(d, "out")
    os.makedirs(out, exist_ok=True)
    res = subprocess.run([HARNESS, mode, prm, out], capture_output=True, text=True, env=env)
    if res.returncode != 0:
        raise RuntimeError(f"ref_harness failed: {res.stderr[-2000:]}")
    n = sysm.natom
    r = {"x": np.fromfile(os.path.join(out, "x.bin"), np.float64).reshape(n, 3),
         "v": np.fromfile(os.path.join(out, "v.bin"), np.float64).reshape(n, 3),
         "f": np.fromfile(os.path.join(out, "f.bin"), np.float64).reshape(n, 3),
         "list": np.fromfile(os.path.join(out, "list.bin"), np.int32).reshape(-1, 4),
         "stdout": res.stdout, "stderr": res.stderr}
    with open(os.path.join(out, "summary.txt")) as fh:
        for ln in fh:
            k, v = ln.split()
            r[k] = float(v)
    return r
