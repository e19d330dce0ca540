"""Seeded molecular-liquid systems and the reference's on-disk input formats.

A :class:`System` holds everything ``ENVIRON::initialize`` and ``MOLECULE::from_file`` read
(reference ``src/environ.cpp:22-141``, ``src/molecule.cpp:25-131``): per-atom type, mass,
charge, position, molecule id and intra-molecule index, plus per-molecule-type topology and the
special-pair class table (1 = 1-2, 2 = 1-3, 3 = 1-4).  Bonded instances are expanded per
molecule exactly as ``BOND/ANGLE/DIHEDRAL::autoGenerate`` do (``src/bond.cpp:45-82``): the
template's intra indices plus the atom id of the molecule's first atom.

The builders here make the water-like and butane-like liquids of the benchmark workloads from a
seed, and write/read the reference's configuration, molecule and pair-table files so that the
same inputs can be fed to the reference binary built from source.
"""
from __future__ import annotations

import dataclasses
import itertools
import math
from typing import List, Optional

import numpy as np

# Units and constants, float32 values as the reference uses them (src/environ.h:57-83).
PI_F = np.float32(3.141592653589793)
ELECTRIC_K = 332.06371
ACCEL_F = 4.184e-4
EKIN_F = 1195.0286802753603
KT_M_SQRT = 0.0009118367518929329
EK_T = 335.47968899917714
EV_GPA = 6.9476954570553735
SPECIAL_LJ = (1.0, 0.0, 0.0, 0.5)          # src/molecule.h:11
SPECIAL_COUL = (1.0, 0.0, 0.0, 5.0 / 6.0)  # src/molecule.h:11


@dataclasses.dataclass
class MolType:
    """One molecule template (``MOLECULE``, src/molecule.h:4-15), all indices 0-based."""
    n_atom: int
    bonds: np.ndarray       # [nb, 3] (type, a, b)
    angles: np.ndarray      # [na, 4] (type, a, centre, c)
    dihedrals: np.ndarray   # [nd, 5] (type, a, b, c, d)
    special: np.ndarray = None  # [n_atom, n_atom] int

    def __post_init__(self):
        self.bonds = np.asarray(self.bonds, np.int32).reshape(-1, 3)
        self.angles = np.asarray(self.angles, np.int32).reshape(-1, 4)
        self.dihedrals = np.asarray(self.dihedrals, np.int32).reshape(-1, 5)
        if self.special is None:
            self.special = special_table(self.n_atom, self.bonds, self.angles, self.dihedrals)


def special_table(n_atom, bonds, angles, dihedrals):
    """Special class of each intra-molecule pair, in the reference's overwrite order:
    1-2 from bonds, then 1-3 from angle ends, then 1-4 from dihedral ends
    (src/molecule.cpp:98-131)."""
    sp = np.zeros((n_atom, n_atom), np.int32)
    for _, a, b in bonds:
        sp[a, b] = sp[b, a] = 1
    for _, a, _c, b in angles:
        sp[a, b] = sp[b, a] = 2
    for _, a, _b, _c, d in dihedrals:
        sp[a, d] = sp[d, a] = 3
    return sp


@dataclasses.dataclass
class System:
    bl: float
    ntype: int
    x: np.ndarray          # [N,3] float32
    v: np.ndarray          # [N,3] float32
    q: np.ndarray          # [N]   float32
    mass: np.ndarray       # [N]   float32
    typ: np.ndarray        # [N]   int32
    mol_id: np.ndarray     # [N]   int32
    intra_id: np.ndarray   # [N]   int32
    mol_type: np.ndarray   # [nmol] int32
    moltypes: List[MolType]
    bond_coef: np.ndarray = None      # [nbtyp, 4]  (b0, 2a, b, 2c)   src/bond.h:11
    angle_coef: np.ndarray = None     # [natyp, 4]  (cos t0, 2a, b, 2c) src/angle.h:11
    dihedral_coef: np.ndarray = None  # [ndtyp, 4]  cubic-in-cos form  src/dihedral.cpp:27-35
    names: tuple = ()
    angle_params: list = None      # raw ANGLE::setParam arguments (theta0_deg, ka, kb, kc)
    dihedral_params: list = None   # raw DIHEDRAL::setParam arguments (c0, c1, c2, c3)
    # Ryckaert-Bellemans dihedral [ndtyp, 6]: U = sum_n C_n cos^n(phi), phi = 0 cis (the reference
    # measures phi that way).  When set it replaces dihedral_coef (nnpe_bonded_set kind 3).
    dihedral_rb: np.ndarray = None

    @property
    def natom(self):
        return int(self.x.shape[0])

    @property
    def nmol(self):
        return int(self.mol_type.shape[0])

    @property
    def npair(self):
        return self.ntype * (self.ntype + 1) // 2

    def pairtype(self):
        """``ENVIRON::pairtype`` (src/environ.cpp:130-139): upper-triangle numbering."""
        pt = np.zeros((self.ntype, self.ntype), np.int32)
        k = 0
        for i in range(self.ntype):
            for j in range(i, self.ntype):
                pt[i, j] = pt[j, i] = k
                k += 1
        return pt

    def mol_first(self):
        first = np.zeros(self.nmol, np.int32)
        first[self.mol_id[self.intra_id == 0]] = np.nonzero(self.intra_id == 0)[0]
        return first

    def expand(self, kind):
        """``autoGenerate`` for kind in {'bond','angle','dihedral'}: [nterm, 1+k] int32 with
        atom ids, molecule after molecule."""
        width = {"bond": 3, "angle": 4, "dihedral": 5}[kind]
        rows = []
        first = self.mol_first()
        for m in range(self.nmol):
            mt = self.moltypes[self.mol_type[m]]
            tmpl = {"bond": mt.bonds, "angle": mt.angles, "dihedral": mt.dihedrals}[kind]
            if len(tmpl):
                r = tmpl.copy()
                r[:, 1:] += first[m]
                rows.append(r)
        if not rows:
            return np.zeros((0, width), np.int32)
        return np.ascontiguousarray(np.concatenate(rows), np.int32)

    def special_flat(self):
        """Concatenated per-molecule-type special tables and their offsets."""
        offs, flat, o = [], [], 0
        for mt in self.moltypes:
            offs.append(o)
            flat.append(mt.special.reshape(-1))
            o += mt.n_atom * mt.n_atom
        return np.asarray(offs, np.int32), np.ascontiguousarray(np.concatenate(flat), np.int32)

    def copy(self):
        return dataclasses.replace(self, x=self.x.copy(), v=self.v.copy())


def random_rotations(rng, n):
    """Uniformly distributed rotation matrices from the QR factorisation of Gaussian matrices
    (sign-fixed so the distribution is Haar)."""
    a = rng.standard_normal((n, 3, 3))
    qm, rm = np.linalg.qr(a)
    d = np.sign(np.einsum("nii->ni", rm))
    qm = qm * d[:, None, :]
    det = np.linalg.det(qm)
    qm[det < 0, :, 0] *= -1
    return qm


def _place_rigid(rng, centres, tmpl, bl, sites, n_side, min_dist=1.7, tries=400):
    """Place one rigid template per lattice site at a uniformly random orientation, re-drawing an
    orientation (up to `tries` times) while any of its atoms lies within `min_dist` of an atom of
    an already placed molecule on a neighbouring site (minimum image).  Returns [n, k, 3]."""
    n, k = centres.shape[0], tmpl.shape[0]
    x = np.zeros((n, k, 3))
    occ = -np.ones((n_side,) * 3, np.int64)
    for m, s in enumerate(sites):
        occ[tuple(s)] = m
    offs = np.array(list(itertools.product((-1, 0, 1), repeat=3)))
    placed = np.zeros(n, bool)
    for m in range(n):
        nb = occ[tuple(((sites[m] + offs) % n_side).T)]
        nb = nb[(nb >= 0) & placed[np.maximum(nb, 0)] & (nb != m)]
        other = x[nb].reshape(-1, 3)
        for t in range(tries):
            rot = random_rotations(rng, 1)[0]
            cand = centres[m] + tmpl @ rot.T
            if len(other) == 0:
                break
            d = cand[:, None, :] - other[None, :, :]
            d -= bl * np.round(d / bl)
            if np.sqrt((d * d).sum(-1)).min() >= min_dist:
                break
        x[m] = cand
        placed[m] = True
    return x


def maxwell_boltzmann(rng, mass, temperature):
    """v ~ N(0, sqrt(kT/m)) per component, in A/fs, as ``ENVIRON::initV`` draws them
    (src/environ.cpp:190-205); the net momentum is then removed."""
    sd = np.sqrt(temperature / mass) * KT_M_SQRT
    v = rng.standard_normal((mass.shape[0], 3)) * sd[:, None]
    p = (v * mass[:, None]).sum(0) / mass.sum()
    return (v - p).astype(np.float32)


WATER_Q = (-0.82, 0.41, 0.41)
WATER_M = (15.999, 1.008, 1.008)


def water_moltype():
    # O = intra 0, H = intra 1, 2; bonds O-H, angle H-O-H.
    return MolType(3, bonds=[(0, 0, 1), (0, 0, 2)], angles=[(0, 1, 0, 2)], dihedrals=np.zeros((0, 5)))


def water_box(nmol: int, bl: float, seed: int, temperature: float = 300.0, r0: float = 1.0,
              theta0_deg: float = 109.47, k_bond: float = 1059.0, k_angle: float = 75.9) -> System:
    """Flexible 3-site water-like liquid: O on a seeded simple-cubic lattice (a random subset of
    sites when nmol is not a cube), uniform random orientations, Maxwell-Boltzmann velocities.

    Bond: the reference's quartic form with a = 0, i.e. U = k (r - r0)^2.
    Angle: the reference's form is harmonic in cos(theta) (src/angle.cpp:113-115, :131-134),
    U = c (cos t - cos t0)^2; c = k_angle / sin^2(t0) matches a harmonic-in-theta k_angle
    (kcal/mol/rad^2) at small deviations.
    """
    rng = np.random.default_rng(seed)
    n_side = int(round(nmol ** (1.0 / 3.0)))
    while n_side ** 3 < nmol:
        n_side += 1
    sites = np.array(list(itertools.product(range(n_side), repeat=3)), np.float64)
    if len(sites) > nmol:
        pick = np.sort(rng.permutation(len(sites))[:nmol])
        sites = sites[pick]
    a = bl / n_side
    o = (sites + 0.5) * a - bl / 2
    th = math.radians(theta0_deg)
    h1 = np.array([math.sin(th / 2), 0.0, math.cos(th / 2)]) * r0
    h2 = np.array([-math.sin(th / 2), 0.0, math.cos(th / 2)]) * r0
    x = _place_rigid(rng, o, np.stack([np.zeros(3), h1, h2]), bl, sites.astype(np.int64), n_side)
    x = x.reshape(-1, 3).astype(np.float32)
    n = 3 * nmol
    typ = np.tile(np.array([0, 1, 1], np.int32), nmol)
    q = np.tile(np.array(WATER_Q, np.float32), nmol)
    mass = np.tile(np.array(WATER_M, np.float32), nmol)
    mol_id = np.repeat(np.arange(nmol, dtype=np.int32), 3)
    intra = np.tile(np.array([0, 1, 2], np.int32), nmol)
    v = maxwell_boltzmann(rng, mass.astype(np.float64), temperature)
    cos0 = np.float32(math.cos(np.float32(theta0_deg) * PI_F / np.float32(180)))
    bond_coef = np.array([[r0, 0.0, 0.0, 2 * k_bond]], np.float32)
    c_ang = k_angle / math.sin(th) ** 2
    angle_coef = np.array([[cos0, 0.0, 0.0, 2 * c_ang]], np.float32)
    return System(bl=float(np.float32(bl)), ntype=2, x=x, v=v, q=q, mass=mass, typ=typ,
                  mol_id=mol_id, intra_id=intra, mol_type=np.zeros(nmol, np.int32),
                  moltypes=[water_moltype()], bond_coef=bond_coef, angle_coef=angle_coef,
                  dihedral_coef=np.zeros((0, 4), np.float32), names=("O", "H"),
                  angle_params=[(theta0_deg, 0.0, 0.0, c_ang)], dihedral_params=[])


def dihedral_coef_from_fourier(c0, c1, c2, c3):
    """``DIHEDRAL::setParam(i,c0,c1,c2,c3)`` (src/dihedral.cpp:27-35): U = c0 + c1 cos(p) +
    c2 cos(2p) + c3 cos(3p) rewritten as a cubic in cos(p); stored (k0, k1, 4*k2, 12*k3) so the
    force polynomial is (coef3 * c + coef2) * c + coef1."""
    return np.array([c0 - c2, c1 - 3 * c3, c2 * 2 * 2, c3 * 4 * 3], np.float32)


def butane_moltype():
    # united-atom CH3-CH2-CH2-CH3: types CH3=0, CH2=1
    return MolType(4, bonds=[(0, 0, 1), (0, 1, 2), (0, 2, 3)],
                   angles=[(0, 0, 1, 2), (0, 1, 2, 3)], dihedrals=[(0, 0, 1, 2, 3)])


def butane_box(nmol: int, bl: float, seed: int, temperature: float = 300.0) -> System:
    """United-atom 4-site butane-like chains in trans conformation at random orientations on a
    seeded lattice (masses CH3 15.035, CH2 14.027; no charges).  The dihedral is the reference's
    cubic-in-cos form (src/dihedral.cpp:27-35) from OPLS-like Fourier coefficients."""
    rng = np.random.default_rng(seed)
    n_side = int(round(nmol ** (1.0 / 3.0)))
    while n_side ** 3 < nmol:
        n_side += 1
    sites = np.array(list(itertools.product(range(n_side), repeat=3)), np.float64)
    if len(sites) > nmol:
        sites = sites[np.sort(rng.permutation(len(sites))[:nmol])]
    a = bl / n_side
    c = (sites + 0.5) * a - bl / 2
    b, th = 1.54, math.radians(114.0)
    # planar zig-zag centred at the origin
    dx, dy = b * math.sin(th / 2), b * math.cos(th / 2)
    tmpl = np.array([[-1.5 * dx, 0, 0], [-0.5 * dx, dy, 0], [0.5 * dx, 0, 0], [1.5 * dx, dy, 0]])
    tmpl -= tmpl.mean(0)
    x = _place_rigid(rng, c, tmpl, bl, sites.astype(np.int64), n_side, min_dist=3.2)
    x = x.reshape(-1, 3).astype(np.float32)
    typ = np.tile(np.array([0, 1, 1, 0], np.int32), nmol)
    mass = np.tile(np.array([15.035, 14.027, 14.027, 15.035], np.float32), nmol)
    q = np.zeros(4 * nmol, np.float32)
    v = maxwell_boltzmann(rng, mass.astype(np.float64), temperature)
    cos0 = np.float32(math.cos(np.float32(114.0) * PI_F / np.float32(180)))
    c_ang = 62.1 / math.sin(th) ** 2
    return System(bl=float(np.float32(bl)), ntype=2, x=x, v=v, q=q, mass=mass, typ=typ,
                  mol_id=np.repeat(np.arange(nmol, dtype=np.int32), 4),
                  intra_id=np.tile(np.arange(4, dtype=np.int32), nmol),
                  mol_type=np.zeros(nmol, np.int32), moltypes=[butane_moltype()],
                  bond_coef=np.array([[1.54, 0, 0, 2 * 260.0]], np.float32),
                  angle_coef=np.array([[cos0, 0, 0, 2 * c_ang]], np.float32),
                  dihedral_coef=dihedral_coef_from_fourier(0.0, 0.7055, -0.1355, 1.5725)[None, :],
                  names=("CH3", "CH2"), angle_params=[(114.0, 0.0, 0.0, c_ang)],
                  dihedral_params=[(0.0, 0.7055, -0.1355, 1.5725)])


def rb_butane_coef():
    """Ryckaert-Bellemans butane torsion, C0..C5 = 9.28, 12.16, -13.12, -3.06, 26.24, -31.5 kJ/mol
    in cos(psi) with psi = phi - 180 deg (trans = 0), converted to kcal/mol and to cos(phi) of
    the reference's convention (cos psi = -cos phi, so C_n picks up (-1)^n)."""
    c = np.array([9.28, 12.16, -13.12, -3.06, 26.24, -31.5]) / 4.184
    return (c * (-1.0) ** np.arange(6)).astype(np.float32)[None, :]


def with_rb_dihedral(sysm: System) -> System:
    """The system with its dihedral replaced by the Ryckaert-Bellemans quintic."""
    s = sysm.copy()
    s.dihedral_rb = rb_butane_coef()
    return s


# ---------------------------------------------------------------------------------------------
# Pair tables (format of LJTableCoulDSF::loadPotential, src/pair.cpp:197-263)


def lj_table(sys: System, cutoff: float, nbin: int, params: dict, r_clamp_frac: float = 0.6):
    """Shifted-force LJ tables F(r)/r and U(r) at r = (i + 1/2) dr for every pair type.

    ``params`` maps (ti, tj) -> (eps, sigma); missing pairs are zero.  Below
    ``r_clamp_frac * sigma`` the values are held at that radius so the float32 table stays
    finite.  Returns (force[npair, nbin+1], potential[npair, nbin+1]) float32 with the last entry
    zero, as loadPotential sets it.
    """
    dr = np.float32(cutoff) / np.float32(nbin)
    r = (np.arange(nbin, dtype=np.float64) + 0.5) * float(dr)
    pt = sys.pairtype()
    f = np.zeros((sys.npair, nbin + 1), np.float64)
    u = np.zeros((sys.npair, nbin + 1), np.float64)
    for (ti, tj), (eps, sig) in params.items():
        p = pt[ti, tj]
        rr = np.maximum(r, r_clamp_frac * sig)

        def uf(x):
            s6 = (sig / x) ** 6
            return 4 * eps * (s6 * s6 - s6), 24 * eps * (2 * s6 * s6 - s6) / x

        uc, fc = uf(float(cutoff))
        ur, fr = uf(rr)
        u[p, :nbin] = ur - uc + (rr - cutoff) * fc
        f[p, :nbin] = (fr - fc) / rr
    return f.astype(np.float32), u.astype(np.float32)


def write_table(path, sys: System, cutoff: float, force, potential):
    nbin = force.shape[1] - 1
    dr = np.float32(cutoff) / np.float32(nbin)
    with open(path, "w") as fh:
        for i in range(nbin):
            cols = [f"{(i + 0.5) * float(dr):.9e}"]
            for p in range(sys.npair):
                cols.append(f"{force[p, i]:.9e}")
                cols.append(f"{potential[p, i]:.9e}")
            fh.write(" ".join(cols) + "\n")


def read_table(path, npair, nbin):
    """Parse a pair-table file into float32 (force, potential) with the trailing zero entry."""
    dat = np.loadtxt(path, dtype=np.float64, ndmin=2)
    f = np.zeros((npair, nbin + 1), np.float32)
    u = np.zeros((npair, nbin + 1), np.float32)
    for p in range(npair):
        f[p, :nbin] = dat[:nbin, 1 + 2 * p]
        u[p, :nbin] = dat[:nbin, 2 + 2 * p]
    return f, u


# ---------------------------------------------------------------------------------------------
# Configuration / molecule files (src/environ.h:109-114, src/molecule.cpp:25-97)


def write_molecule(path, mt: MolType):
    with open(path, "w") as fh:
        fh.write(f"{mt.n_atom} atoms\n{len(mt.bonds)} bonds\n{len(mt.angles)} angles\n"
                 f"{len(mt.dihedrals)} dihedrals\n\n")
        if len(mt.bonds):
            fh.write("Bonds\n\n")
            for k, (t, a, b) in enumerate(mt.bonds):
                fh.write(f"{k + 1} {t + 1} {a + 1} {b + 1}\n")
            fh.write("\n")
        if len(mt.angles):
            fh.write("Angles\n\n")
            for k, (t, a, b, c) in enumerate(mt.angles):
                fh.write(f"{k + 1} {t + 1} {a + 1} {b + 1} {c + 1}\n")
            fh.write("\n")
        if len(mt.dihedrals):
            fh.write("Dihedrals\n\n")
            for k, (t, a, b, c, d) in enumerate(mt.dihedrals):
                fh.write(f"{k + 1} {t + 1} {a + 1} {b + 1} {c + 1} {d + 1}\n")


def write_config(path, sys: System, mol_files):
    """Configuration file: ``nmoltype nmol`` / molecule files / ``ntype natom BL`` /
    ``mol_type typ mass charge x y z`` per atom.  Values are printed with 9 significant digits so
    the float32 parse reproduces them exactly."""
    with open(path, "w") as fh:
        fh.write(f"{len(sys.moltypes)} {sys.nmol}\n")
        for mf in mol_files:
            fh.write(f"{mf}\n")
        fh.write(f"{sys.ntype} {sys.natom} {np.float32(sys.bl):.9g}\n")
        mt_atom = sys.mol_type[sys.mol_id]
        for i in range(sys.natom):
            fh.write(f"{mt_atom[i]} {sys.typ[i]} {sys.mass[i]:.9g} {sys.q[i]:.9g} "
                     f"{sys.x[i, 0]:.9g} {sys.x[i, 1]:.9g} {sys.x[i, 2]:.9g}\n")


def write_classical_input(d, sysm, pair_cutoff, skin, gr_cutoff, force_tbl, pot_tbl, nbin_gr, nq, q_max,
                          temp=(300.0, 300.0), steps=(100, 10, 10, 10, 10), outputs=None, rmdf=None, sqd=(10.0, 10)):
    """The positional input file of ``classical`` (src/run.cpp:25-50) plus the configuration,
    molecule and table files it names, written into directory d.  Returns the input path.
    rmdf = (strength, gamma, [experimental S(q) files], initial-weights file) or None (no data
    sets); sqd = (q_max, nq) of the end-of-run direct S(q) (written when outputs names 'sqd')."""
    import os
    mols = []
    for k, mt in enumerate(sysm.moltypes):
        p = os.path.join(d, f"mol{k}.txt")
        write_molecule(p, mt)
        mols.append(p)
    cfg = os.path.join(d, "config.txt")
    write_config(cfg, sysm, mols)
    tbl = os.path.join(d, "table.txt")
    write_table(tbl, sysm, pair_cutoff, force_tbl, pot_tbl)
    lines = [cfg, f"{pair_cutoff:.9g} {skin:.9g} {gr_cutoff:.9g}", f"{force_tbl.shape[1] - 1} {tbl}",
             str(len(sysm.bond_coef))]
    lines += [f"{r[0]:.9g} {r[3] / 2:.9g}" for r in sysm.bond_coef]
    ap = sysm.angle_params or []
    lines += [str(len(ap))] + [" ".join(f"{v:.9g}" for v in r) for r in ap]
    dp = sysm.dihedral_params or []
    lines += [str(len(dp))] + [" ".join(f"{v:.9g}" for v in r) for r in dp]
    strength, gamma, files, delta = rmdf if rmdf else (0.0, 1.0, [], "*")
    lines += [f"{nbin_gr} {nq} {q_max:.9g}", f"{strength:.9g} {gamma:.9g}", " ".join([str(len(files))] + list(files)),
              delta, f"{temp[0]:.9g} {temp[1]:.9g}", " ".join(str(s) for s in steps)]
    lines += _output_lines(outputs, sqd, force=True)
    path = os.path.join(d, "in.classical")
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")
    return path


def _output_lines(outputs, sqd, force):
    """The dump lines at the end of the reference's inputs: name and append flag for the xyz,
    g(r), S(q), SQD (+ q_max, nq), sqx and (classical) refined-table dumps; an output may be
    given as a name or as (name, append)."""
    o = outputs or {}

    def nm(k):
        v = o.get(k, "*")
        return (v, 0) if isinstance(v, str) else (v[0], int(v[1]))

    out = [f"{nm('xyz')[0]} {nm('xyz')[1]}", f"{nm('gr')[0]} {nm('gr')[1]}", f"{nm('sq')[0]} {nm('sq')[1]}",
           f"{nm('sqd')[0]} {nm('sqd')[1]} {sqd[0]:.9g} {sqd[1]}", f"{nm('sqx')[0]} {nm('sqx')[1]}"]
    if force:
        out.append(f"{nm('force')[0]} {nm('force')[1]}")
    return out


def write_nnp_input(d, sysm, pair_cutoff, skin, gr_cutoff, nbin_tbl, tbl_coef, widths, params, out_scale, precision,
                    nbin_gr, nq, q_max, temp=(300.0, 300.0), steps=(100, 10, 10, 10, 10), outputs=None, rmdf=None,
                    sqd=(10.0, 10)):
    """The positional input file of ``nnp_run`` (src/run.cpp:968-1100) and the files it names:
    configuration and molecules, the parameter file (its first token names the network file;
    the reference names a TF SavedModel there), the network file (text: nlayer, widths,
    out_scale, precision, then per element and layer W [in][out] and b), and the binary weight
    file the reference reads (src/pair_nnp.cpp:505-517): bond, angle and dihedral coefficients
    as ``setParam()`` expects them (the factor-2 terms not yet doubled), then coef_p
    [npair][4][6], coef_w, coef_d, coef_r [npair][80] as float32.  Returns the input path."""
    import os
    mols = []
    for k, mt in enumerate(sysm.moltypes):
        p = os.path.join(d, f"mol{k}.txt")
        write_molecule(p, mt)
        mols.append(p)
    cfg = os.path.join(d, "config.txt")
    write_config(cfg, sysm, mols)
    net = os.path.join(d, "net.txt")
    with open(net, "w") as fh:
        fh.write(f"{len(widths) - 1} " + " ".join(str(w) for w in widths) + f" {out_scale:.9g} {precision}\n")
        fh.write("\n".join(f"{v:.9e}" for v in np.asarray(params, np.float32)) + "\n")
    prm = os.path.join(d, "nnp_param.txt")
    with open(prm, "w") as fh:
        fh.write(net + "\n")
    wgt = os.path.join(d, "weights.bin")

    def undouble(coef, cols):
        c = np.array(coef, np.float32).reshape(-1, 4).copy()
        for k in cols:
            c[:, k] /= 2
        return c

    cp, cw, cd, cr = tbl_coef
    dih = np.zeros((0, 4), np.float32) if sysm.dihedral_coef is None else np.asarray(sysm.dihedral_coef, np.float32)
    # DIHEDRAL::setParam() multiplies coef[2] by 2 and coef[3] by 3 (src/dihedral.cpp:37-44)
    dih_file = dih.reshape(-1, 4).copy()
    dih_file[:, 2] /= 2
    dih_file[:, 3] /= 3
    parts = [undouble(sysm.bond_coef, (1, 3)), undouble(sysm.angle_coef, (1, 3)), dih_file,
             np.asarray(cp, np.float32), np.asarray(cw, np.float32), np.asarray(cd, np.float32), np.asarray(cr, np.float32)]
    with open(wgt, "wb") as fh:
        for a in parts:
            fh.write(np.ascontiguousarray(a, np.float32).tobytes())
    strength, gamma, files, delta = rmdf if rmdf else (0.0, 1.0, [], "*")
    lines = [cfg, f"{pair_cutoff:.9g} {skin:.9g} {gr_cutoff:.9g}", str(len(sysm.bond_coef)), str(len(sysm.angle_coef)),
             str(len(dih)), prm, wgt, str(nbin_tbl), f"{nbin_gr} {nq} {q_max:.9g}", f"{strength:.9g} {gamma:.9g}",
             " ".join([str(len(files))] + list(files)), delta, f"{temp[0]:.9g} {temp[1]:.9g}",
             " ".join(str(s) for s in steps)]
    lines += _output_lines(outputs, sqd, force=False)
    path = os.path.join(d, "in.nnp")
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")
    return path


WATER_LJ = {(0, 0): (0.155, 3.166)}
BUTANE_LJ_BASE = {0: (0.195, 3.75), 1: (0.091, 3.95)}


def butane_lj(eps_scale=1.0):
    out = {}
    for i in range(2):
        for j in range(i, 2):
            ei, si = BUTANE_LJ_BASE[i]
            ej, sj = BUTANE_LJ_BASE[j]
            out[(i, j)] = (eps_scale * math.sqrt(ei * ej), 0.5 * (si + sj))
    return out
