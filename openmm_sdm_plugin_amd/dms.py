"""Desmond DMS (SQLite) reader for the reference's inputs (example/desmonddmsfile75.py).

`DesmondDMSFile(file)` mirrors the reference class: the topology, positions and
velocities (`:147-233`), and `createSystem(...)` (`:360-545`), which here returns a
`DMSForceField` holding every force-field term the reference builds, in OpenMM's
internal units (nm, ps, amu, e, kJ/mol, rad):

  bonds        stretch_harm: k doubled (Desmond writes k without the 1/2, `:563-565`);
               constrained ones become constraints
  angles       angle_harm: k doubled (`:597-599`); constrained ones become a constraint
               between the outer atoms (law of cosines, `:587-592`)
  constraints  the above, constraint_ah* (`:615-625`) and the H-H constraint of
               constraint_hoh (`:627-637`)
  torsions     dihedral_trig: one periodic term k (1 + cos(n phi - phi0)) per nonzero
               fc_n, n = 0..6 (`:666-682`); with OPLS the reference's CustomTorsionForce
               f cos(n phi - phi0) (`:652-656`, `:679-680`), written as the two periodic
               terms (f/2, phi0) and (-f/2, phi0 + pi), whose sum it is exactly
  impropers    improper_harm: k (phi - phi0)^2 (`:684-717`)
  exclusions   exclusion (`:803-808`)
  exceptions   pair_12_6_es: qq = q_ij, sigma = (a/b)^(1/6), epsilon = b^2/(4a)
               (`:818-845`); each must also be an exclusion
  posre        posre_harmflatbottom: hk (dist - tol)^2 beyond tol, hk = fc/2 (`:972-1017`);
               posre_harm: sum_d hk_d (x_d - x0_d)^2, hk = fc/2 (`:917-970`)
  nonbonded    charge, sigma, epsilon per particle; Lorentz-Berthelot, or with
               createSystem(OPLS=True) the geometric sigma of the reference's
               CustomNonbondedForce (`:780-801`: LJ sqrt(s1 s2), sqrt(e1 e2), plain cutoff,
               no long-range correction; charges stay in the NonbondedForce), which the
               engine evaluates in its pair kernels (sdm_system_desc.lj_rule)

  virtual sites virtual_lc2: x = (1-c1) x1 + c1 x2; virtual_lc3: x = (1-c1-c2) x1 + c1 x2 +
               c2 x3; virtual_out3: x = x1 + c1 r12 + c2 r13 + c3 (r12 x r13), c3 passed on
               unconverted as the reference does (`:852-915`); the engine places them and
               moves their forces (sdm_set_virtual_sites)

Multi-file inputs are concatenated with index offsets as in the reference (`:79-117`).
Unsupported terms raise NotImplementedError as the reference does (CMAP, fdat3 virtual
sites, flat-bottom stretch/angle/improper terms, non-12-6 van der Waals, Drude particles);
so do virtual-site tables other than lc2, lc3, out3 and fdat3, which the reference's
reader passes over without placing the sites.

`to_sdm_system(ff, ...)` turns a DMSForceField into the engine's SDMSystem with the
reaction field in place of Ewald/PME (the engine's only long-range treatment).
"""
from __future__ import annotations

import dataclasses
import math
import os
import sqlite3
import types
from typing import List, Optional

import numpy as np

KCAL = 4.184  # kJ per kcal
ANG = 0.1  # nm per angstrom
DEG = math.pi / 180.0

# the virtual-site tables the reference's reader places sites from (:852-911)
VSITE_TABLES = ("virtual_lc2_term", "virtual_lc3_term", "virtual_out3_term")
NONBONDED_METHODS = ("NoCutoff", "CutoffNonPeriodic", "CutoffPeriodic", "Ewald", "PME", "LJPME")


@dataclasses.dataclass
class DMSAtom:
    index: int
    id: int  # particle.id + 1 (the reference's topology ids are 1-based)
    name: str
    anum: int
    resname: str
    resid: int
    chain: str


@dataclasses.dataclass
class DMSTopology:
    atoms_: List[DMSAtom]
    bonds_: np.ndarray  # (b, 2)
    box: Optional[np.ndarray]  # (3,) nm, rectangular; None without global_cell

    def atoms(self):
        return iter(self.atoms_)

    def getNumAtoms(self):
        return len(self.atoms_)

    def bonds(self):
        return iter(map(tuple, self.bonds_))

    def getUnitCellDimensions(self):
        return None if self.box is None else self.box.copy()


@dataclasses.dataclass
class DMSForceField:
    """The OpenMM System that the reference's createSystem builds, as arrays."""

    mass: np.ndarray
    charge: np.ndarray
    sigma: np.ndarray
    epsilon: np.ndarray
    opls: bool
    exclusions: np.ndarray  # (m, 2)
    exception_pairs: np.ndarray  # (k, 2)
    exception_qq: np.ndarray  # e^2
    exception_sigma: np.ndarray  # nm
    exception_epsilon: np.ndarray  # kJ/mol
    bonds: np.ndarray  # (b, 2)
    bond_r0: np.ndarray
    bond_k: np.ndarray  # kJ/mol/nm^2, E = k/2 (r - r0)^2
    angles: np.ndarray  # (a, 3)
    angle_t0: np.ndarray
    angle_k: np.ndarray  # kJ/mol/rad^2, E = k/2 (theta - theta0)^2
    constraints: np.ndarray  # (c, 2)
    constraint_r: np.ndarray
    torsions: np.ndarray  # (t, 4)
    torsion_n: np.ndarray
    torsion_phase: np.ndarray
    torsion_k: np.ndarray  # kJ/mol, E = k (1 + cos(n phi - phase))
    impropers: np.ndarray  # (i, 4)
    improper_t0: np.ndarray
    improper_k: np.ndarray  # kJ/mol/rad^2, E = k (phi - phi0)^2
    posre_fb_atoms: np.ndarray
    posre_fb_x0: np.ndarray  # (p, 3) nm
    posre_fb_hk: np.ndarray  # kJ/mol/nm^2
    posre_fb_tol: np.ndarray  # nm
    posre_h_atoms: np.ndarray
    posre_h_x0: np.ndarray
    posre_h_hk: np.ndarray  # (p, 3)
    box: Optional[np.ndarray]
    nonbonded_method: str
    cutoff: float
    remove_cm_motion: bool
    bonded_force_group: int = 1
    nonbonded_force_group: int = 2
    # virtual sites: kind (0 two-particle average, 1 three-particle average, 2 out of
    # plane), atoms (v, 4): site, p1, p2, p3; weights (v, 3)
    vsite_kind: np.ndarray = dataclasses.field(
        default_factory=lambda: np.zeros(0, dtype=np.int32))
    vsite_atoms: np.ndarray = dataclasses.field(
        default_factory=lambda: np.zeros((0, 4), dtype=np.int32))
    vsite_weights: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros((0, 3)))

    @property
    def n(self):
        return int(self.mass.shape[0])


class DesmondDMSFile:
    """Reads one DMS file or a list of them (one molecule each)."""

    def __init__(self, file, verbose=False):
        self._verbose = verbose
        self._file = list(file) if isinstance(file, (list, tuple)) else [file]
        self._conn, self._tables, self._natoms = [], [], []
        for f in self._file:
            if not os.path.exists(str(f)):
                raise IOError("No such file or directory: %s" % str(f))
            conn = sqlite3.connect(str(f))
            tables = self._read_schemas(conn)
            if not tables:
                raise IOError("DMS file %s was not loaded sucessfully. No tables found" % str(f))
            if "nbtype" not in tables["particle"]:
                raise ValueError("No nonbonded parameters associated with DMS file %s" % str(f))
            self._conn.append(conn)
            self._tables.append(tables)
        self.topology, self.positions, self.velocities = self._create_topology()
        self._offset = np.concatenate([[0], np.cumsum(self._natoms)]).astype(int)
        self._bonded_force_group = 1
        self._nonbonded_force_group = 2

    # ---- the reference's accessors ----
    def getPositions(self):
        return self.positions.copy()

    def getVelocities(self):
        return self.velocities.copy()

    def getTopology(self):
        return self.topology

    def getBondedForceGroup(self):
        return self._bonded_force_group

    def getNonBondedForceGroup(self):
        return self._nonbonded_force_group

    def setPositions(self, positions):
        """Writes positions (nm) back to the files (the reference's setPositions)."""
        self._write_xyz(np.asarray(positions, dtype=np.float64).reshape(-1, 3) / ANG,
                        ("x", "y", "z"))

    def setVelocities(self, velocities):
        self._write_xyz(np.asarray(velocities, dtype=np.float64).reshape(-1, 3) / ANG,
                        ("vx", "vy", "vz"))

    def close(self):
        for c in self._conn:
            c.close()
        self._conn = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- internals ----
    @staticmethod
    def _read_schemas(conn):
        tables = {}
        for (name,) in conn.execute("SELECT name FROM sqlite_master WHERE type='table'"):
            tables[str(name)] = [str(e[1]) for e in conn.execute("PRAGMA table_info(%s)" % name)]
        return tables

    def _local(self):
        return zip(self._conn, self._tables, self._offset[:-1])

    def _write_xyz(self, v, cols):
        q = "UPDATE particle SET %s = ?1, %s = ?2, %s = ?3 WHERE id == ?4" % cols
        for conn, _t, off in self._local():
            n = conn.execute("SELECT COUNT(*) FROM particle").fetchone()[0]
            ids = [r[0] for r in conn.execute("SELECT id FROM particle ORDER BY id")]
            for k in range(n):
                a = v[off + k]
                conn.execute(q, (float(a[0]), float(a[1]), float(a[2]), ids[k]))
            conn.commit()

    def _create_topology(self):
        box = None
        if "global_cell" in self._tables[0]:
            cell = [row for row in self._conn[0].execute("SELECT x, y, z FROM global_cell")]
            box = np.array([cell[0][0], cell[1][1], cell[2][2]], dtype=np.float64) * ANG
        atoms, bonds, pos, vel = [], [], [], []
        for conn in self._conn:
            base = len(atoms)
            idmap = {}
            q = ("SELECT id, name, anum, resname, resid, chain, x, y, z, vx, vy, vz "
                 "FROM particle ORDER BY id")
            for (pid, name, anum, resname, resid, chain, x, y, z, vx, vy, vz) in conn.execute(q):
                idmap[pid] = len(atoms)
                atoms.append(DMSAtom(len(atoms), len(atoms) + 1, name, int(anum), resname,
                                     int(resid), chain))
                pos.append((x, y, z))
                vel.append((vx, vy, vz))
            self._natoms.append(len(atoms) - base)
            for p0, p1 in conn.execute("SELECT p0, p1 FROM bond"):
                bonds.append((idmap[p0], idmap[p1]))
        top = DMSTopology(atoms, np.asarray(bonds, dtype=np.int32).reshape(-1, 2), box)
        return (top, np.asarray(pos, dtype=np.float64) * ANG,
                np.asarray(vel, dtype=np.float64) * ANG)

    def _check_unsupported(self):
        opls = False
        for conn, tables, _off in self._local():
            for t in ("stretch_fbhw_term", "angle_fbhw_term", "improper_fbhw_term",
                      "posre_fbhw_term"):
                if t in tables:
                    raise NotImplementedError("Flat bottom potential terms are not implemeneted")
            row = conn.execute("SELECT * FROM nonbonded_info").fetchone()
            info = dict(zip(tables["nonbonded_info"], row))
            if info["vdw_funct"] != "vdw_12_6":
                raise NotImplementedError("Only Lennard-Jones van der Waals interactions are "
                                          "currently supported")
            rule = info["vdw_rule"]
            if rule not in ("arithmetic", "geometric", "arithmetic/geometric"):
                raise NotImplementedError("Unknown vdw combination rule")
            opls = opls or rule == "geometric"
            if "alchemical_particle" in tables:
                raise NotImplementedError("Alchemical particles are not supported")
            if "alchemical_stretch_harm" in tables:
                raise NotImplementedError("Alchemical bonds are not supported")
            if "polar_term" in tables and \
                    conn.execute("SELECT COUNT(*) FROM polar_term").fetchone()[0] != 0:
                raise NotImplementedError("Drude particles are not currently supported")
            if "cmap_term" in tables and \
                    conn.execute("SELECT COUNT(*) FROM cmap_term").fetchone()[0] != 0:
                raise NotImplementedError("CMAP terms are not supported")
            if "virtual_fdat3_term" in tables:
                raise NotImplementedError("OpenMM does not currently support fdat3-style "
                                          "virtual sites")
            if any(t.startswith("virtual_") and t.endswith("_term") and
                   conn.execute("SELECT COUNT(*) FROM %s" % t).fetchone()[0] != 0
                   for t in tables if t not in VSITE_TABLES and t != "virtual_term"):
                raise NotImplementedError("Virtual sites other than lc2, lc3 and out3 are "
                                          "not supported")
        return opls

    def _rows(self, q, table):
        """(offset, row) for every file holding `table`."""
        for conn, tables, off in self._local():
            if table in tables:
                for row in conn.execute(q):
                    yield off, row

    def createSystem(self, nonbondedMethod="NoCutoff", nonbondedCutoff=1.0, OPLS=False,
                     removeCMMotion=True, hydrogenMass=None):
        """The force-field terms of the files (see the module docstring).  nonbondedMethod
        is one of NONBONDED_METHODS (OpenMM's names); nonbondedCutoff in nm."""
        if nonbondedMethod not in NONBONDED_METHODS:
            raise ValueError("Illegal nonbonded method %r" % (nonbondedMethod,))
        # The file's vdw_rule is checked (an unknown rule raises) but, as in the
        # reference, only the OPLS argument selects geometric combining: its reader records
        # a "geometric" rule (desmonddmsfile75.py:1061-1063) and then builds the forces
        # with the createSystem argument (:409, :415), Lorentz-Berthelot by default.
        self._check_unsupported()
        opls = bool(OPLS)
        box = self.topology.box
        if box is None and nonbondedMethod in ("CutoffPeriodic", "Ewald", "PME", "LJPME"):
            raise ValueError("Illegal nonbonded method for a non-periodic system")
        n = self.topology.getNumAtoms()
        mass = np.array([m for _o, (m,) in self._rows("SELECT mass FROM particle ORDER BY id",
                                                       "particle")], dtype=np.float64)
        bond_r = {}  # (i, j) -> r0 nm, for the angle constraints
        bonds, br0, bk, cons, cr = [], [], [], [], []
        q = ("SELECT p0, p1, r0, fc, constrained FROM stretch_harm_term INNER JOIN "
             "stretch_harm_param ON stretch_harm_term.param=stretch_harm_param.id")
        for off, (p0, p1, r0, fc, constrained) in self._rows(q, "stretch_harm_term"):
            i, j = p0 + off, p1 + off
            if constrained:
                cons.append((i, j))
                cr.append(r0 * ANG)
            else:
                bonds.append((i, j))
                br0.append(r0 * ANG)
                bk.append(2.0 * fc * KCAL / ANG ** 2)
            bond_r[(i, j)] = bond_r[(j, i)] = r0 * ANG
        angles, at0, ak = [], [], []
        angle_cons = set()
        q = ("SELECT p0, p1, p2, theta0, fc, constrained FROM angle_harm_term INNER JOIN "
             "angle_harm_param ON angle_harm_term.param=angle_harm_param.id")
        for off, (p0, p1, p2, theta0, fc, constrained) in self._rows(q, "angle_harm_term"):
            i, j, k = p0 + off, p1 + off, p2 + off
            if constrained:
                l1, l2 = bond_r[(j, i)], bond_r[(j, k)]
                cons.append((i, k))
                cr.append(math.sqrt(l1 * l1 + l2 * l2 - 2 * l1 * l2 * math.cos(theta0 * DEG)))
                angle_cons.add((j, i, k))
                angle_cons.add((j, k, i))
            else:
                angles.append((i, j, k))
                at0.append(theta0 * DEG)
                ak.append(2.0 * fc * KCAL)
        for conn, tables, off in self._local():
            for term in sorted(t for t in tables if t.startswith("constraint_a") and
                               t.endswith("term")):
                param = term.replace("term", "param")
                q = ("SELECT p0, p1, r1 FROM %s INNER JOIN %s ON %s.param=%s.id"
                     % (term, param, term, param))
                for p0, p1, r1 in conn.execute(q):
                    i, j = p0 + off, p1 + off
                    if (i, j) not in bond_r:
                        cons.append((i, j))
                        cr.append(r1 * ANG)
                        bond_r[(i, j)] = bond_r[(j, i)] = r1 * ANG
            if "constraint_hoh_term" in tables:
                q = ("SELECT p0, p1, p2, r1, r2, theta FROM constraint_hoh_term INNER JOIN "
                     "constraint_hoh_param ON constraint_hoh_term.param=constraint_hoh_param.id")
                for p0, p1, p2, r1, r2, theta in conn.execute(q):
                    o, h1, h2 = p0 + off, p1 + off, p2 + off
                    if (o, h1, h2) not in angle_cons:
                        a, b = r1 * ANG, r2 * ANG
                        cons.append((h1, h2))
                        cr.append(math.sqrt(a * a + b * b - 2 * a * b * math.cos(theta * DEG)))
        tors, tn, tph, tk = [], [], [], []
        q = ("SELECT p0, p1, p2, p3, phi0, fc0, fc1, fc2, fc3, fc4, fc5, fc6 FROM "
             "dihedral_trig_term INNER JOIN dihedral_trig_param "
             "ON dihedral_trig_term.param=dihedral_trig_param.id")
        for off, row in self._rows(q, "dihedral_trig_term"):
            p = tuple(a + off for a in row[:4])
            for order, fc in enumerate(row[5:12]):
                if fc == 0:
                    continue
                if opls:  # f cos(x) = f/2 (1 + cos x) - f/2 (1 + cos(x - pi))
                    tors += [p, p]
                    tn += [order, order]
                    tph += [row[4] * DEG, row[4] * DEG + math.pi]
                    tk += [0.5 * fc * KCAL, -0.5 * fc * KCAL]
                else:
                    tors.append(p)
                    tn.append(order)
                    tph.append(row[4] * DEG)
                    tk.append(fc * KCAL)
        imps, it0, ik = [], [], []
        q = ("SELECT p0, p1, p2, p3, phi0, fc FROM improper_harm_term INNER JOIN "
             "improper_harm_param ON improper_harm_term.param=improper_harm_param.id")
        for off, (p0, p1, p2, p3, phi0, fc) in self._rows(q, "improper_harm_term"):
            imps.append((p0 + off, p1 + off, p2 + off, p3 + off))
            it0.append(phi0 * DEG)
            ik.append(fc * KCAL)
        fb_a, fb_x, fb_k, fb_t = [], [], [], []
        q = ("SELECT p0, x0, y0, z0, fc, tol FROM posre_harmflatbottom_term INNER JOIN "
             "posre_harmflatbottom_param "
             "ON posre_harmflatbottom_term.param=posre_harmflatbottom_param.id")
        for off, (p0, x0, y0, z0, fc, tol) in self._rows(q, "posre_harmflatbottom_term"):
            fb_a.append(p0 + off)
            fb_x.append((x0 * ANG, y0 * ANG, z0 * ANG))
            fb_k.append(0.5 * fc * KCAL / ANG ** 2)
            fb_t.append(tol * ANG)
        ph_a, ph_x, ph_k = [], [], []
        q = ("SELECT p0, x0, y0, z0, fcx, fcy, fcz FROM posre_harm_term INNER JOIN "
             "posre_harm_param ON posre_harm_term.param=posre_harm_param.id")
        for off, (p0, x0, y0, z0, fx, fy, fz) in self._rows(q, "posre_harm_term"):
            ph_a.append(p0 + off)
            ph_x.append((x0 * ANG, y0 * ANG, z0 * ANG))
            ph_k.append((0.5 * fx * KCAL / ANG ** 2, 0.5 * fy * KCAL / ANG ** 2,
                         0.5 * fz * KCAL / ANG ** 2))
        q = ("SELECT charge, sigma, epsilon FROM particle INNER JOIN nonbonded_param "
             "ON particle.nbtype=nonbonded_param.id ORDER BY particle.id")
        nb = np.array([row for _o, row in self._rows(q, "particle")], dtype=np.float64)
        charge, sigma, eps = nb[:, 0], nb[:, 1] * ANG, nb[:, 2] * KCAL
        excl = [(p0 + off, p1 + off) for off, (p0, p1) in
                self._rows("SELECT p0, p1 FROM exclusion", "exclusion")]
        ex_p, ex_q, ex_s, ex_e = [], [], [], []
        q = ("SELECT p0, p1, aij, bij, qij FROM pair_12_6_es_term INNER JOIN "
             "pair_12_6_es_param ON pair_12_6_es_term.param=pair_12_6_es_param.id")
        excl_set = set(excl) | set((b, a) for a, b in excl)
        for off, (p0, p1, a_ij, b_ij, q_ij) in self._rows(q, "pair_12_6_es_term"):
            i, j = p0 + off, p1 + off
            if (i, j) not in excl_set:
                raise NotImplementedError("All pair_12_6_es_terms must have a corresponding "
                                          "exclusion")
            a = a_ij * KCAL * ANG ** 12  # kJ/mol nm^12
            b = b_ij * KCAL * ANG ** 6  # kJ/mol nm^6
            if a == 0.0 or b == 0.0:
                se, ss = 0.0, 1.0
            else:
                se, ss = b * b / (4 * a), (a / b) ** (1.0 / 6.0)
            ex_p.append((i, j))
            ex_q.append(q_ij)
            ex_s.append(ss)
            ex_e.append(se)
        if hydrogenMass is not None:
            for a, b in self.topology.bonds_:
                if self.topology.atoms_[a].anum == 1:
                    a, b = b, a
                if self.topology.atoms_[b].anum == 1 and self.topology.atoms_[a].anum > 1:
                    transfer = hydrogenMass - mass[b]
                    mass[b] = hydrogenMass
                    mass[a] -= transfer

        # virtual sites, table by table as the reference adds them (:852-915)
        vk, va, vw = [], [], []
        q = ("SELECT p0, p1, p2, c1 FROM virtual_lc2_term INNER JOIN virtual_lc2_param "
             "ON virtual_lc2_term.param=virtual_lc2_param.id")
        for off, (p0, p1, p2, c1) in self._rows(q, "virtual_lc2_term"):
            vk.append(0)
            va.append((p0 + off, p1 + off, p2 + off, p1 + off))
            vw.append((1.0 - c1, c1, 0.0))
        q = ("SELECT p0, p1, p2, p3, c1, c2 FROM virtual_lc3_term INNER JOIN "
             "virtual_lc3_param ON virtual_lc3_term.param=virtual_lc3_param.id")
        for off, (p0, p1, p2, p3, c1, c2) in self._rows(q, "virtual_lc3_term"):
            vk.append(1)
            va.append((p0 + off, p1 + off, p2 + off, p3 + off))
            vw.append((1.0 - c1 - c2, c1, c2))
        q = ("SELECT p0, p1, p2, p3, c1, c2, c3 FROM virtual_out3_term INNER JOIN "
             "virtual_out3_param ON virtual_out3_term.param=virtual_out3_param.id")
        for off, (p0, p1, p2, p3, c1, c2, c3) in self._rows(q, "virtual_out3_term"):
            vk.append(2)
            va.append((p0 + off, p1 + off, p2 + off, p3 + off))
            vw.append((c1, c2, c3))

        def arr(v, shape, dtype=np.float64):
            return np.asarray(v, dtype=dtype).reshape(shape)
        return DMSForceField(
            mass=mass, charge=charge, sigma=sigma, epsilon=eps, opls=bool(opls),
            exclusions=arr(excl, (-1, 2), np.int32),
            exception_pairs=arr(ex_p, (-1, 2), np.int32), exception_qq=arr(ex_q, (-1,)),
            exception_sigma=arr(ex_s, (-1,)), exception_epsilon=arr(ex_e, (-1,)),
            bonds=arr(bonds, (-1, 2), np.int32), bond_r0=arr(br0, (-1,)), bond_k=arr(bk, (-1,)),
            angles=arr(angles, (-1, 3), np.int32), angle_t0=arr(at0, (-1,)),
            angle_k=arr(ak, (-1,)), constraints=arr(cons, (-1, 2), np.int32),
            constraint_r=arr(cr, (-1,)), torsions=arr(tors, (-1, 4), np.int32),
            torsion_n=arr(tn, (-1,), np.int32), torsion_phase=arr(tph, (-1,)),
            torsion_k=arr(tk, (-1,)), impropers=arr(imps, (-1, 4), np.int32),
            improper_t0=arr(it0, (-1,)), improper_k=arr(ik, (-1,)),
            posre_fb_atoms=arr(fb_a, (-1,), np.int32), posre_fb_x0=arr(fb_x, (-1, 3)),
            posre_fb_hk=arr(fb_k, (-1,)), posre_fb_tol=arr(fb_t, (-1,)),
            posre_h_atoms=arr(ph_a, (-1,), np.int32), posre_h_x0=arr(ph_x, (-1, 3)),
            posre_h_hk=arr(ph_k, (-1, 3)), box=None if box is None else box.copy(),
            nonbonded_method=nonbondedMethod, cutoff=float(nonbondedCutoff),
            remove_cm_motion=bool(removeCMMotion), vsite_kind=arr(vk, (-1,), np.int32),
            vsite_atoms=arr(va, (-1, 4), np.int32), vsite_weights=arr(vw, (-1, 3)))


def to_sdm_system(ff: DMSForceField, positions, velocities=None, displacement=None,
                  ligand=None, eps_rf: Optional[float] = None, cutoff: Optional[float] = None,
                  name: str = ""):
    """An SDMSystem of the engine from a DMS force field: the nonbonded force with the
    reaction field (dielectric eps_rf) in place of Ewald/PME, which the engine does not
    have; bonds, angles and exclusions as SDMSystem arrays; torsions, impropers,
    positional restraints, 1-4 exceptions and constraints as `terms`.  displacement:
    (n, 3) nm per atom (zeros by default); ligand: the displaced atom indices.

    CutoffNonPeriodic (the reference's vacuum example, example/test.py:63) is evaluated
    as OpenMM does it: no periodic images, the reaction field with OpenMM's default
    dielectric 78.3 (the reference's reader never sets another).  For the periodic
    methods eps_rf defaults to 78.5, the dielectric of the benchmark workloads."""
    from .systems import SDMSystem  # local: systems does not depend on this module
    periodic = ff.nonbonded_method != "CutoffNonPeriodic"
    if ff.nonbonded_method == "NoCutoff" or (periodic and ff.box is None):
        raise NotImplementedError("the engine evaluates cutoff systems only (CutoffPeriodic "
                                  "with the reaction field in place of Ewald/PME, or "
                                  "CutoffNonPeriodic)")
    if eps_rf is None:
        eps_rf = 78.5 if periodic else 78.3
    n = ff.n
    x = np.asarray(positions, dtype=np.float64).reshape(n, 3)
    v = np.zeros((n, 3)) if velocities is None else \
        np.asarray(velocities, dtype=np.float64).reshape(n, 3)
    disp = np.zeros((n, 3)) if displacement is None else \
        np.asarray(displacement, dtype=np.float64).reshape(n, 3)
    if ligand is None:
        ligand = np.nonzero(np.any(disp != 0.0, axis=1))[0]
    vsites = None
    if len(ff.vsite_kind):
        vsites = types.SimpleNamespace(kind=ff.vsite_kind.copy(), atoms=ff.vsite_atoms.copy(),
                                       weights=ff.vsite_weights.copy())
    return SDMSystem(
        box=ff.box.copy() if periodic else np.zeros(3),
        cutoff=float(ff.cutoff if cutoff is None else cutoff),
        eps_rf=float(eps_rf), charge=ff.charge.copy(), sigma=ff.sigma.copy(),
        epsilon=ff.epsilon.copy(), mass=ff.mass.copy(), exclusions=ff.exclusions.copy(),
        bonds=ff.bonds.copy(), bond_r0=ff.bond_r0.copy(), bond_k=ff.bond_k.copy(),
        angles=ff.angles.copy(), angle_t0=ff.angle_t0.copy(), angle_k=ff.angle_k.copy(),
        displacement=disp, positions=x, velocities=v, periodic=periodic,
        ligand=np.asarray(ligand, dtype=np.int32), name=name, terms=ff, virtual_sites=vsites,
        lj_rule=1 if ff.opls else 0)
