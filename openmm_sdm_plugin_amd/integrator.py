"""Host-side mirror of the reference plugin's public API, over the HIP C ABI.

LangevinIntegratorSDM keeps the method names, argument meaning, defaults and error
behaviour of the reference's SWIG-wrapped class (python/SDMplugin.i:84-145,
openmmapi/include/LangevinIntegratorSDM.h:71-472, defaults
openmmapi/src/LangevinIntegratorSDM.cpp:48-85).  Context stands in for the OpenMM
Context the integrator is bound to (LangevinIntegratorSDM::initialize,
LangevinIntegratorSDM.cpp:88-106).  SDMForce exposes the same alchemical
parameters as a force evaluated without stepping; getPerturbationEnergy() is the
reference's getBindE() (the soft-cored u').  ReplicaLadder runs many
lambda-replicas of one topology in every launch.
"""
from __future__ import annotations

import copy
import math
import os
from typing import List, Optional, Sequence

import numpy as np

from . import _native
from .systems import SDMParams, SDMSystem


def _reference_defaults(temperature, friction, step_size) -> SDMParams:
    p = SDMParams(bias_method=0, softcore_method=0, lambda_=1.0, gamma=0.0, wbcoeff=1.0,
                  w0coeff=0.0, lambda1=1.0, lambda2=1.0, alpha=1.0, u0=0.0, umax=200.0,
                  acore=0.25, ubcore=0.0, nonequilibrium=0, noneq_tmax=1.0, work=0.0,
                  temperature=float(temperature), friction=float(friction),
                  step_size=float(step_size))
    p.seed = int.from_bytes(os.urandom(4), "little")
    return p


class LangevinIntegratorSDM:
    """Langevin integrator with the SDM alchemical hybrid force (reference API)."""

    LinearMethod = 0
    QuadraticMethod = 1
    ILogisticMethod = 2
    NoSoftCoreMethod = 0
    TanhMethod = 1
    RationalMethod = 2

    def __init__(self, temperature, frictionCoeff, stepSize, nParticles):
        self._p = _reference_defaults(temperature, frictionCoeff, stepSize)
        self._n = int(nParticles)
        self._displ = np.zeros((self._n, 3))
        self._ctx: Optional["Context"] = None
        self._dirty = True
        self._bind_e = 0.0
        self._pot = 0.0
        self._constraint_tol = 1e-5

    # -- parameter plumbing ---------------------------------------------------
    def _set(self, name, value):
        setattr(self._p, name, value)
        self._dirty = True

    def _sync_to_device(self):
        if self._ctx is not None and self._dirty:
            self._ctx._native.set_params(0, self._p)
            self._dirty = False

    # -- LangevinIntegratorSDM.h:81-91 ----------------------------------------
    def setLambda(self, v): self._set("lambda_", float(v))
    def getLambda(self): return self._p.lambda_
    def setBiasMethod(self, m): self._set("bias_method", int(m))
    def getBiasMethod(self): return self._p.bias_method

    def setSoftCoreMethod(self, m):
        self._set("softcore_method", int(m))

    def getSoftCoreMethod(self): return self._p.softcore_method
    def setGamma(self, v): self._set("gamma", float(v))
    def getGamma(self): return self._p.gamma
    def setWBcoeff(self, v): self._set("wbcoeff", float(v))
    def getWBcoeff(self): return self._p.wbcoeff
    def setW0coeff(self, v): self._set("w0coeff", float(v))
    def getW0coeff(self): return self._p.w0coeff
    def setLambda1(self, v): self._set("lambda1", float(v))
    def getLambda1(self): return self._p.lambda1
    def setLambda2(self, v): self._set("lambda2", float(v))
    def getLambda2(self): return self._p.lambda2
    def setAlpha(self, v): self._set("alpha", float(v))
    def getAlpha(self): return self._p.alpha
    def setU0(self, v): self._set("u0", float(v))
    def getU0(self): return self._p.u0
    def getTemperature(self): return self._p.temperature
    def setTemperature(self, v): self._set("temperature", float(v))
    def getFriction(self): return self._p.friction
    def setFriction(self, v): self._set("friction", float(v))
    def getStepSize(self): return self._p.step_size
    def setStepSize(self, v): self._set("step_size", float(v))
    def getConstraintTolerance(self): return self._constraint_tol

    def setConstraintTolerance(self, v):
        """Integrator::setConstraintTolerance: the relative tolerance the constraint
        solver is given every step (ReferenceStochasticDynamicsSDM.cpp:252)."""
        self._constraint_tol = float(v)
        if self._ctx is not None:
            self._ctx._native.set_constraint_tolerance(self._constraint_tol)
    def getRandomNumberSeed(self): return int(self._p.seed)
    def setRandomNumberSeed(self, s): self._set("seed", int(s) & 0xFFFFFFFFFFFFFFFF)
    def getBindE(self): return self._bind_e
    def setBindE(self, v): self._bind_e = float(v)
    def getPotEnergy(self): return self._pot
    def setPotEnergy(self, v): self._pot = float(v)
    def getUmax(self): return self._p.umax
    def setUmax(self, v): self._set("umax", float(v))
    def getAcore(self): return self._p.acore
    def setAcore(self, v): self._set("acore", float(v))
    def getUbcore(self): return self._p.ubcore
    def setUbcore(self, v): self._set("ubcore", float(v))
    def setNoneqtmax(self, v): self._set("noneq_tmax", float(v))
    def getNoneqtmax(self): return self._p.noneq_tmax
    def setNonEquilibrium(self, v): self._set("nonequilibrium", int(v))
    def getNonEquilibrium(self): return self._p.nonequilibrium
    def setNoneqWorkvalue(self, v): self._set("work", float(v))
    def getNoneqWorkvalue(self): return self._p.work
    def setlambda1Slope(self, v): self._set("m_l1", float(v))
    def getlambda1Slope(self): return self._p.m_l1
    def setlambda2Slope(self, v): self._set("m_l2", float(v))
    def getlambda2Slope(self): return self._p.m_l2
    def setu0Slope(self, v): self._set("m_u0", float(v))
    def getu0Slope(self): return self._p.m_u0
    def setw0Slope(self, v): self._set("m_w0", float(v))
    def getw0Slope(self): return self._p.m_w0
    # no counterpart in the reference (alpha is constant while switching there)
    def setalphaSlope(self, v): self._set("m_alpha", float(v))
    def getalphaSlope(self): return self._p.m_alpha
    def setlambda1intercept(self, v): self._set("b_l1", float(v))
    def getlambda1intercept(self): return self._p.b_l1
    def setlambda2intercept(self, v): self._set("b_l2", float(v))
    def getlambda2intercept(self): return self._p.b_l2
    def setu0intercept(self, v): self._set("b_u0", float(v))
    def getu0intercept(self): return self._p.b_u0
    def setw0intercept(self, v): self._set("b_w0", float(v))
    def getw0intercept(self): return self._p.b_w0

    def setDisplacement(self, atom, dx, dy, dz):
        self._displ[int(atom)] = (float(dx), float(dy), float(dz))

    def getDisplacement(self, atom):
        return tuple(float(v) for v in self._displ[int(atom)])

    def SoftCoreF(self, u, umax, a, ub):
        """LangevinIntegratorSDM::SoftCoreF (LangevinIntegratorSDM.cpp:125-149); returns
        (u', du'/du).  Host-side scalar helper of the public API."""
        m = self._p.softcore_method
        if u <= ub or m == self.NoSoftCoreMethod:
            return u, 1.0
        if m == self.TanhMethod:
            t = math.tanh((u - ub) / umax)
            return umax * t + ub, 1.0 - t * t
        if m == self.RationalMethod:
            gu = (u - ub) / (a * (umax - ub))
            zeta = 1.0 + 2.0 * gu * (gu + 1.0)
            zetap = zeta ** a
            fp = 4.0 * (2.0 * gu + 1.0) / zeta * zetap / (1.0 + zetap) ** 2
            return (umax - ub) * (zetap - 1.0) / (zetap + 1.0) + ub, fp
        raise ValueError("Unknown soft core method")

    # -- Integrator hooks (LangevinIntegratorSDM.cpp:88-120) -------------------
    def _initialize(self, ctx: "Context"):
        if self._ctx is not None and self._ctx is not ctx:
            raise RuntimeError("This Integrator is already bound to a context")
        self._ctx = ctx
        self._dirty = True

    def step(self, steps):
        """LangevinIntegratorSDM::step (LangevinIntegratorSDM.cpp:153-183)."""
        if self._ctx is None:
            raise RuntimeError("This Integrator is not bound to a context")
        if self._p.softcore_method not in (0, 1, 2):
            raise ValueError("Unknown soft core method")
        self._sync_to_device()
        nat = self._ctx._native
        nat.step(int(steps))
        e = nat.energies()[0]
        self._bind_e = e["bind_e"]
        self._pot = e["pot_energy"]
        self._last = e
        nat.get_params(0, self._p)
        self._dirty = False
        if e["flags"] & _native.FLAG_CAPACITY:
            raise RuntimeError("neighbour-list capacity exceeded")

    def computeKineticEnergy(self):
        if self._ctx is None:
            raise RuntimeError("This Integrator is not bound to a context")
        return float(self._ctx._native.kinetic_energy()[0])


class Context:
    """A system bound to a LangevinIntegratorSDM on the current HIP device."""

    def __init__(self, system: SDMSystem, integrator: LangevinIntegratorSDM):
        if system.n != integrator._n:
            raise ValueError("integrator particle count does not match the system")
        integrator._initialize(self)
        self.system = system
        self.integrator = integrator
        displ = integrator._displ
        if not np.any(displ) and system.displacement is not None:
            displ = np.asarray(system.displacement, dtype=np.float64)
            integrator._displ = displ.copy()
        self._native = _native.NativeContext(system, 1, displacement=displ)
        self._native.set_constraint_tolerance(integrator.getConstraintTolerance())
        integrator._sync_to_device()
        self.setPositions(system.positions)
        self.setVelocities(system.velocities)

    def setPositions(self, x):
        self._native.set_positions(0, np.asarray(x, dtype=np.float64))

    def setVelocities(self, v):
        self._native.set_velocities(0, np.asarray(v, dtype=np.float64))

    def setTime(self, t, step=0):
        self._native.set_time(0, t, step)

    def getPositions(self):
        return self._native.get_positions(0)

    def getVelocities(self):
        return self._native.get_velocities(0)

    def evaluate(self, components=False):
        """Energies and hybrid force at the current positions (no step)."""
        self.integrator._sync_to_device()
        if components:
            self._native.compute_components()
        else:
            self._native.compute()
        return self._native.energies()[0]

    def getForces(self, which=_native.FORCE_HYBRID):
        return self._native.forces(0, which)


class SDMForce:
    """The SDM alchemical force with the reference integrator's parameter surface
    (lambda1/lambda2/alpha/u0/w0, umax/acore/ubcore, bias and soft-core methods, the
    per-atom displacement map), evaluated without stepping."""

    def __init__(self, system: SDMSystem, n_replicas: int = 1):
        self.system = system
        self._displ = np.asarray(system.displacement, dtype=np.float64).copy()
        self._params = [_reference_defaults(300.0, 1.0, 0.001) for _ in range(n_replicas)]
        self._native = None
        self.R = n_replicas

    def _ensure(self):
        if self._native is None:
            self._native = _native.NativeContext(self.system, self.R, displacement=self._displ)
            for r, p in enumerate(self._params):
                self._native.set_params(r, p)
                self._native.set_positions(r, self.system.positions)

    def setParams(self, replica, params: SDMParams):
        self._params[replica] = copy.copy(params)
        if self._native is not None:
            self._native.set_params(replica, params)

    def setDisplacement(self, atom, dx, dy, dz):
        if self._native is not None:
            raise RuntimeError("the displacement map is fixed once the force is initialised")
        self._displ[int(atom)] = (dx, dy, dz)

    def setPositions(self, replica, x):
        self._ensure()
        self._native.set_positions(replica, x)

    def compute(self):
        self._ensure()
        self._native.compute()
        return self._native.energies()

    def getPerturbationEnergy(self, replica=0):
        return self.compute()[replica]["bind_e"]


class ReplicaLadder:
    """R lambda-replicas of one topology advanced together (every launch covers all)."""

    def __init__(self, system: SDMSystem, params: Sequence[SDMParams],
                 positions: Optional[Sequence[np.ndarray]] = None,
                 velocities: Optional[Sequence[np.ndarray]] = None):
        self.R = len(params)
        self.system = system
        self._native = _native.NativeContext(system, self.R)
        self.params: List[SDMParams] = [copy.copy(p) for p in params]
        for r, p in enumerate(self.params):
            self._native.set_params(r, p)
            self._native.set_positions(r, system.positions if positions is None else positions[r])
            self._native.set_velocities(r, system.velocities if velocities is None
                                        else velocities[r])

    @property
    def native(self):
        return self._native

    def step(self, n):
        self._native.step(n)

    def energies(self):
        return self._native.energies()

    def bind_energies(self):
        return np.array([e["bind_e"] for e in self.energies()])

    def set_params(self, r, p):
        self.params[r] = copy.copy(p)
        self._native.set_params(r, p)
