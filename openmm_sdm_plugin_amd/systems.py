"""Seeded synthetic SDM systems (water-like solvent, host cavity, guest ligands).

These are the five workloads the SDM hot path is measured and checked on:

  make_config1  3,000 atoms, 3.1 nm box: 990 waters + 30-atom ligand chain, d = (2.2,0,0)
  make_config2  22,750 atoms, 6.1 nm box: 7,500 waters + 200-atom host + 50-atom guest,
                d = (2.5,0,0)
  make_config3  22 replicas of config 2 (own coordinates, shared topology), lambda ladder
  make_config4  config 1 frames for the parameter-gradient (adjoint) workload
  make_config5  99,920 atoms, 10 nm box: 33,000 waters + 800-atom host + two 60-atom
                guests with d = (+3,0,0) and (-3,0,0), 8-replica lambda ladder

Choices the workload descriptions leave open (recorded here and in the docs):
  * water LJ: TIP3P oxygen sigma 0.315061 nm, epsilon 0.636386 kJ/mol; H epsilon 0
    (sigma 0.1 nm, irrelevant because epsilon is 0); masses O 15.9994, H 1.008;
    H-O-H angle constant 836.8 kJ/mol/rad^2 (bond constant 4.6e5 as specified).
  * solute atoms: mass 12.011; chain bonds 0.15 nm / 2.5e5 kJ/mol/nm^2, chain angles
    111 deg / 500 kJ/mol/rad^2; host shell bonds to near neighbours at their built
    length (1.0e5 kJ/mol/nm^2) plus angle terms (200 kJ/mol/rad^2).
  * exclusions: every bonded (1-2) pair and the end pairs of every angle term (1-3);
    no 1-4 scaling.
  * soft core: rational, umax 200, acore 0.0625, ubcore 100 kJ/mol in every config;
    friction 1 ps^-1, 300 K, 1 fs.
Coordinates are rounded to float32 (positions are FP32 in every workload).
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Optional

import numpy as np

# water model
_Q_O, _Q_H = -0.834, 0.417
_SIG_O, _EPS_O = 0.315061, 0.636386
_SIG_H, _EPS_H = 0.1, 0.0
_M_O, _M_H, _M_C = 15.9994, 1.008, 12.011
_R_OH, _K_OH = 0.09572, 4.6e5
_T_HOH, _K_HOH = math.radians(104.52), 836.8
_CHAIN_B, _CHAIN_KB = 0.15, 2.5e5
_CHAIN_T, _CHAIN_KT = math.radians(111.0), 500.0
_SHELL_KB, _SHELL_KT = 1.0e5, 200.0

BIAS_LINEAR, BIAS_QUADRATIC, BIAS_ILOGISTIC = 0, 1, 2
SOFTCORE_NONE, SOFTCORE_TANH, SOFTCORE_RATIONAL = 0, 1, 2


@dataclasses.dataclass
class SDMParams:
    """Per-replica alchemical and thermostat parameters (LangevinIntegratorSDM's state)."""

    bias_method: int = BIAS_ILOGISTIC
    softcore_method: int = SOFTCORE_RATIONAL
    lambda_: float = 0.5
    gamma: float = 0.0
    wbcoeff: float = 1.0
    w0coeff: float = 0.0
    lambda1: float = 0.5
    lambda2: float = 0.5
    alpha: float = 0.0
    u0: float = 0.0
    umax: float = 200.0
    acore: float = 0.0625
    ubcore: float = 100.0
    nonequilibrium: int = 0
    noneq_tmax: float = 1.0
    work: float = 0.0
    m_l1: float = 0.0
    m_l2: float = 0.0
    m_u0: float = 0.0
    m_w0: float = 0.0
    b_l1: float = 0.0
    b_l2: float = 0.0
    b_u0: float = 0.0
    b_w0: float = 0.0
    temperature: float = 300.0
    friction: float = 1.0
    step_size: float = 0.001
    seed: int = 12345
    m_alpha: float = 0.0  # alpha(lambda) = alpha + m_alpha*lambda while switching


@dataclasses.dataclass
class SDMSystem:
    """Topology + one set of coordinates.  Units nm, kJ/mol, e, amu, ps."""

    box: np.ndarray  # (3,)
    cutoff: float
    eps_rf: float
    charge: np.ndarray  # (n,)
    sigma: np.ndarray
    epsilon: np.ndarray
    mass: np.ndarray
    exclusions: np.ndarray  # (m,2) int32
    bonds: np.ndarray  # (b,2) int32
    bond_r0: np.ndarray
    bond_k: np.ndarray
    angles: np.ndarray  # (a,3) int32
    angle_t0: np.ndarray
    angle_k: np.ndarray
    displacement: np.ndarray  # (n,3) float64
    positions: np.ndarray  # (n,3) float64 holding float32 values
    velocities: np.ndarray  # (n,3)
    periodic: bool = True
    ligand: Optional[np.ndarray] = None  # displaced atom indices
    name: str = ""
    # group-1 restraints (restraints.CentroidRestraint / AlignmentRestraint)
    restraints: list = dataclasses.field(default_factory=list)
    # further force-field terms and constraints (a dms.DMSForceField), or None
    terms: Optional[object] = None
    # virtual sites (kind (v,), atoms (v,4): site, parents; weights (v,3)), or None
    virtual_sites: Optional[object] = None
    # LJ sigma combining: 0 Lorentz-Berthelot, 1 geometric (OPLS)
    lj_rule: int = 0

    @property
    def n(self) -> int:
        return int(self.charge.shape[0])

    def with_coordinates(self, positions, velocities=None) -> "SDMSystem":
        s = dataclasses.replace(self)
        s.positions = np.asarray(positions, dtype=np.float64).reshape(-1, 3)
        if velocities is not None:
            s.velocities = np.asarray(velocities, dtype=np.float64).reshape(-1, 3)
        return s


def _unit(v):
    return v / np.linalg.norm(v)


def _perp_basis(b):
    a = np.array([1.0, 0.0, 0.0]) if abs(b[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
    e1 = _unit(np.cross(b, a))
    e2 = np.cross(b, e1)
    return e1, e2


def _chain(rng, n, center, radius, min_nb=0.25):
    """Self-avoiding chain with fixed bond length and angle, confined to a sphere."""
    center = np.asarray(center, dtype=np.float64)
    for _attempt in range(2000):
        pos = [center + rng.uniform(-0.1, 0.1, 3)]
        pos.append(pos[0] + _CHAIN_B * _unit(rng.normal(size=3)))
        ok = True
        for i in range(2, n):
            b = _unit(pos[i - 2] - pos[i - 1])
            e1, e2 = _perp_basis(b)
            placed = False
            for _trial in range(60):
                phi = rng.uniform(0.0, 2.0 * math.pi)
                w = math.cos(_CHAIN_T) * b + math.sin(_CHAIN_T) * (
                    math.cos(phi) * e1 + math.sin(phi) * e2)
                new = pos[i - 1] + _CHAIN_B * w
                if np.linalg.norm(new - center) > radius:
                    continue
                if i >= 3:
                    prev = np.asarray(pos[: i - 2])
                    if np.min(np.linalg.norm(prev - new, axis=1)) < min_nb:
                        continue
                pos.append(new)
                placed = True
                break
            if not placed:
                ok = False
                break
        if ok:
            return np.asarray(pos)
    raise RuntimeError("could not grow ligand chain")


def _chain_topology(first, n):
    bonds = [(first + i, first + i + 1) for i in range(n - 1)]
    angles = [(first + i, first + i + 1, first + i + 2) for i in range(n - 2)]
    return bonds, angles


def _fibonacci_shell(n, radius, center):
    i = np.arange(n) + 0.5
    phi = np.arccos(1.0 - 2.0 * i / n)
    theta = math.pi * (1.0 + 5.0 ** 0.5) * i
    pts = np.stack([np.cos(theta) * np.sin(phi), np.sin(theta) * np.sin(phi), np.cos(phi)], 1)
    return np.asarray(center) + radius * pts


def _shell_topology(pos, first):
    """Bonds to near neighbours (<= 1.3 x nearest spacing, at most 5 per atom) and
    up to two angle terms per atom."""
    n = pos.shape[0]
    d = np.linalg.norm(pos[:, None, :] - pos[None, :, :], axis=2)
    np.fill_diagonal(d, np.inf)
    dnn = np.median(d.min(axis=1))
    nbrs = [[] for _ in range(n)]
    bonds = []
    order = np.argsort(d, axis=1)
    for i in range(n):
        for j in order[i, :5]:
            j = int(j)
            if d[i, j] > 1.3 * dnn:
                break
            if j > i and j not in nbrs[i] and len(nbrs[i]) < 5 and len(nbrs[j]) < 5:
                nbrs[i].append(j)
                nbrs[j].append(i)
                bonds.append((i, j))
    bond_r0 = [float(d[i, j]) for i, j in bonds]
    angles, angle_t0 = [], []
    for m in range(n):
        nb = nbrs[m]
        for k in range(min(2, len(nb) - 1)):
            a, c = nb[k], nb[k + 1]
            v1, v2 = pos[a] - pos[m], pos[c] - pos[m]
            cs = np.dot(v1, v2) / (np.linalg.norm(v1) * np.linalg.norm(v2))
            angles.append((a, m, c))
            angle_t0.append(float(np.arccos(np.clip(cs, -1.0, 1.0))))
    bonds = [(first + i, first + j) for i, j in bonds]
    angles = [(first + a, first + m, first + c) for a, m, c in angles]
    return bonds, bond_r0, angles, angle_t0


def _random_solute_params(rng, n):
    q = rng.uniform(-0.4, 0.4, n)
    q -= q.mean()
    sig = rng.uniform(0.25, 0.35, n)
    eps = rng.uniform(0.1, 0.8, n)
    return q, sig, eps


def _random_rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    a, b, c, d = q
    return np.array([
        [a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
        [2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)],
        [2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d],
    ])


def _water_sites(rng, box, n_water, solute_pos, clearance, cavities=()):
    """O sites on a jittered cubic lattice (spacing >= 0.28 nm), away from solutes."""
    box = np.asarray(box, dtype=np.float64)
    need = n_water
    for n_side in range(int(math.ceil(need ** (1.0 / 3.0))), 200):
        spacing = box / n_side
        if np.min(spacing) < 0.28:
            raise RuntimeError("box too small for the requested water count")
        g = np.stack(np.meshgrid(*[np.arange(n_side)] * 3, indexing="ij"), -1).reshape(-1, 3)
        sites = (g + 0.5) * spacing
        keep = np.ones(len(sites), dtype=bool)
        if solute_pos is not None and len(solute_pos):
            for start in range(0, len(solute_pos), 256):
                sp = solute_pos[start:start + 256]
                dd = sites[:, None, :] - sp[None, :, :]
                dd -= box * np.round(dd / box)
                keep &= np.min(np.linalg.norm(dd, axis=2), axis=1) > clearance
        for c, r in cavities:
            dd = sites - np.asarray(c)
            dd -= box * np.round(dd / box)
            keep &= np.linalg.norm(dd, axis=1) > r
        cand = np.nonzero(keep)[0]
        if len(cand) >= need:
            pick = np.sort(rng.choice(cand, size=need, replace=False))
            jit = max(0.0, (float(np.min(spacing)) - 0.28) / 2.0 - 1e-4)
            return sites[pick] + rng.uniform(-jit, jit, (need, 3))
    raise RuntimeError("could not place waters")


class _Builder:
    def __init__(self):
        self.q, self.sig, self.eps, self.mass = [], [], [], []
        self.pos = []
        self.bonds, self.bond_r0, self.bond_k = [], [], []
        self.angles, self.angle_t0, self.angle_k = [], [], []

    @property
    def n(self):
        return len(self.q)

    def add_atoms(self, pos, q, sig, eps, mass):
        self.pos.extend(list(np.asarray(pos)))
        self.q.extend(q)
        self.sig.extend(sig)
        self.eps.extend(eps)
        self.mass.extend(mass)

    def add_chain(self, rng, n, center, radius, params):
        first = self.n
        pos = _chain(rng, n, center, radius)
        q, sig, eps = params
        self.add_atoms(pos, q, sig, eps, [_M_C] * n)
        b, a = _chain_topology(first, n)
        self.bonds += b
        self.bond_r0 += [_CHAIN_B] * len(b)
        self.bond_k += [_CHAIN_KB] * len(b)
        self.angles += a
        self.angle_t0 += [_CHAIN_T] * len(a)
        self.angle_k += [_CHAIN_KT] * len(a)
        return np.arange(first, first + n)

    def add_shell(self, n, radius, center, params):
        first = self.n
        pos = _fibonacci_shell(n, radius, center)
        q, sig, eps = params
        self.add_atoms(pos, q, sig, eps, [_M_C] * n)
        b, r0, a, t0 = _shell_topology(pos, first)
        self.bonds += b
        self.bond_r0 += r0
        self.bond_k += [_SHELL_KB] * len(b)
        self.angles += a
        self.angle_t0 += t0
        self.angle_k += [_SHELL_KT] * len(a)
        return np.arange(first, first + n)

    def add_waters(self, rng, o_sites):
        canon = np.array([
            [0.0, 0.0, 0.0],
            [_R_OH * math.sin(_T_HOH / 2), 0.0, _R_OH * math.cos(_T_HOH / 2)],
            [-_R_OH * math.sin(_T_HOH / 2), 0.0, _R_OH * math.cos(_T_HOH / 2)],
        ])
        for o in o_sites:
            first = self.n
            rot = _random_rotation(rng)
            pos = o + canon @ rot.T
            self.add_atoms(pos, [_Q_O, _Q_H, _Q_H], [_SIG_O, _SIG_H, _SIG_H],
                           [_EPS_O, _EPS_H, _EPS_H], [_M_O, _M_H, _M_H])
            self.bonds += [(first, first + 1), (first, first + 2)]
            self.bond_r0 += [_R_OH, _R_OH]
            self.bond_k += [_K_OH, _K_OH]
            self.angles.append((first + 1, first, first + 2))
            self.angle_t0.append(_T_HOH)
            self.angle_k.append(_K_HOH)

    def build(self, box, cutoff, eps_rf, displaced, rng_v, temperature=300.0, name=""):
        n = self.n
        box = np.asarray(box, dtype=np.float64)
        excl = set()
        for i, j in self.bonds:
            excl.add((min(i, j), max(i, j)))
        for a, _m, c in self.angles:
            excl.add((min(a, c), max(a, c)))
        excl = np.array(sorted(excl), dtype=np.int32).reshape(-1, 2)
        displ = np.zeros((n, 3))
        lig = []
        for atoms, d in displaced:
            displ[atoms] = d
            lig.extend(list(atoms))
        # Molecules stay whole (bonded terms use unwrapped coordinates); every heavy
        # atom lies inside the box by construction, a hydrogen may sit just outside.
        pos = np.asarray(self.pos, dtype=np.float64)
        pos = pos.astype(np.float32).astype(np.float64)
        mass = np.asarray(self.mass, dtype=np.float64)
        kT = 1.380658e-23 * 6.0221367e23 / 1000.0 * temperature
        vel = rng_v.normal(size=(n, 3)) * np.sqrt(kT / mass)[:, None]
        return SDMSystem(
            box=box, cutoff=float(cutoff), eps_rf=float(eps_rf),
            charge=np.asarray(self.q, dtype=np.float64),
            sigma=np.asarray(self.sig, dtype=np.float64),
            epsilon=np.asarray(self.eps, dtype=np.float64),
            mass=mass, exclusions=excl,
            bonds=np.asarray(self.bonds, dtype=np.int32).reshape(-1, 2),
            bond_r0=np.asarray(self.bond_r0, dtype=np.float64),
            bond_k=np.asarray(self.bond_k, dtype=np.float64),
            angles=np.asarray(self.angles, dtype=np.int32).reshape(-1, 3),
            angle_t0=np.asarray(self.angle_t0, dtype=np.float64),
            angle_k=np.asarray(self.angle_k, dtype=np.float64),
            displacement=displ, positions=pos, velocities=vel,
            ligand=np.asarray(sorted(lig), dtype=np.int32), name=name)


def make_config1(frame: int = 0, topo_seed: int = 1) -> SDMSystem:
    """3,000 atoms: 990 waters + 30-atom ligand chain displaced by (2.2,0,0) nm in a
    3.1 nm box; rc 0.9 nm, eps_rf 78.5.  frame k uses coordinate seed 1000+k."""
    box = np.array([3.1, 3.1, 3.1])
    trng = np.random.default_rng(topo_seed)
    lig_params = _random_solute_params(trng, 30)
    rng = np.random.default_rng(1000 + frame)
    b = _Builder()
    center = box / 2.0
    lig = b.add_chain(rng, 30, center, 0.6, lig_params)
    sites = _water_sites(rng, box, 990, np.asarray(b.pos), 0.32)
    b.add_waters(rng, sites)
    return b.build(box, 0.9, 78.5, [(lig, (2.2, 0.0, 0.0))], rng, name="config1")


def params_config1() -> SDMParams:
    return SDMParams(bias_method=BIAS_ILOGISTIC, softcore_method=SOFTCORE_RATIONAL,
                     lambda_=0.5, lambda1=0.5, lambda2=0.5, alpha=0.0, u0=0.0, w0coeff=0.0,
                     umax=200.0, acore=0.0625, ubcore=100.0)


def make_config2(frame: int = 0, topo_seed: int = 2) -> SDMSystem:
    """22,750 atoms: 7,500 waters + 200-atom host shell + 50-atom guest displaced by
    (2.5,0,0) nm in a 6.1 nm box; rc 1.0 nm."""
    box = np.array([6.1, 6.1, 6.1])
    trng = np.random.default_rng(topo_seed)
    host_params = _random_solute_params(trng, 200)
    guest_params = _random_solute_params(trng, 50)
    rng = np.random.default_rng(2000 + frame)
    b = _Builder()
    center = box / 2.0
    b.add_shell(200, 1.0, center, host_params)
    guest = b.add_chain(rng, 50, center, 0.6, guest_params)
    sites = _water_sites(rng, box, 7500, np.asarray(b.pos), 0.32, cavities=[(center, 1.3)])
    b.add_waters(rng, sites)
    return b.build(box, 1.0, 78.5, [(guest, (2.5, 0.0, 0.0))], rng, name="config2")


def params_config2() -> SDMParams:
    return SDMParams(bias_method=BIAS_ILOGISTIC, softcore_method=SOFTCORE_RATIONAL,
                     lambda_=0.3, lambda1=0.3, lambda2=0.3, alpha=0.1, u0=110.0, w0coeff=0.0,
                     umax=200.0, acore=0.0625, ubcore=100.0)


def make_config3(n_replicas: int = 22) -> List[SDMSystem]:
    """22 replicas of config 2 with their own coordinates (frames 0..21), one topology."""
    return [make_config2(frame=k) for k in range(n_replicas)]


def params_config3(n_replicas: int = 22) -> List[SDMParams]:
    """Leg 1: 11 lambda 0 -> 0.5; leg 2: 11 lambda 0.5 -> 0; alpha 0.1, u0 110, w0 0."""
    half = n_replicas // 2
    lam = list(np.linspace(0.0, 0.5, half)) + list(np.linspace(0.5, 0.0, n_replicas - half))
    out = []
    for k, l in enumerate(lam):
        p = params_config2()
        p.lambda_ = p.lambda1 = p.lambda2 = float(l)
        p.seed = 777 + k
        out.append(p)
    return out


def make_config4(n_frames: int = 64) -> List[SDMSystem]:
    return [make_config1(frame=k) for k in range(n_frames)]


def params_config4() -> SDMParams:
    """Config 1's soft core with an ILogistic bias whose gradients in alpha, u0 and w0
    are all nonzero (lambda1 != lambda2, alpha > 0)."""
    return SDMParams(bias_method=BIAS_ILOGISTIC, softcore_method=SOFTCORE_RATIONAL,
                     lambda_=0.4, lambda1=0.2, lambda2=0.6, alpha=0.1, u0=110.0, w0coeff=0.0,
                     umax=200.0, acore=0.0625, ubcore=100.0)


def make_config5(frame: int = 0, topo_seed: int = 5) -> SDMSystem:
    """99,920 atoms: 33,000 waters + 800-atom host shell + guest A (60, d=(+3,0,0)) inside
    the host and guest B (60, d=(-3,0,0)) 3 nm away in solvent; 10 nm box, rc 1.2 nm."""
    box = np.array([10.0, 10.0, 10.0])
    trng = np.random.default_rng(topo_seed)
    host_params = _random_solute_params(trng, 800)
    ga_params = _random_solute_params(trng, 60)
    gb_params = _random_solute_params(trng, 60)
    rng = np.random.default_rng(5000 + frame)
    b = _Builder()
    center = box / 2.0 - np.array([1.5, 0.0, 0.0])
    b.add_shell(800, 1.9, center, host_params)
    ga = b.add_chain(rng, 60, center, 0.7, ga_params)
    gb = b.add_chain(rng, 60, center + np.array([3.0, 0.0, 0.0]), 0.7, gb_params)
    sites = _water_sites(rng, box, 33000, np.asarray(b.pos), 0.32, cavities=[(center, 2.2)])
    b.add_waters(rng, sites)
    return b.build(box, 1.2, 78.5, [(ga, (3.0, 0.0, 0.0)), (gb, (-3.0, 0.0, 0.0))], rng,
                   name="config5")


def params_config5(n_replicas: int = 8) -> List[SDMParams]:
    out = []
    for k, l in enumerate(np.linspace(0.1, 0.45, n_replicas)):
        p = params_config2()
        p.lambda_ = p.lambda1 = p.lambda2 = float(l)
        p.alpha, p.u0 = 0.1, 100.0
        p.seed = 555 + k
        out.append(p)
    return out


def make_small(n_water: int = 60, n_lig: int = 6, box: float = 1.9, cutoff: float = 0.8,
               frame: int = 0, displ=(0.9, 0.0, 0.0), topo_seed: int = 11) -> SDMSystem:
    """A few hundred atoms for brute-force checks."""
    boxv = np.array([box, box, box])
    trng = np.random.default_rng(topo_seed)
    lig_params = _random_solute_params(trng, n_lig)
    rng = np.random.default_rng(9000 + frame)
    b = _Builder()
    lig = b.add_chain(rng, n_lig, boxv / 2.0, 0.35, lig_params)
    sites = _water_sites(rng, boxv, n_water, np.asarray(b.pos), 0.32)
    b.add_waters(rng, sites)
    return b.build(boxv, cutoff, 78.5, [(lig, displ)], rng, name="small")
