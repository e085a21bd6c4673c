"""ctypes bindings of the FP64 oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py.  The product package never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class orc_params(C.Structure):
    _fields_ = [
        ("bias_method", C.c_int), ("softcore_method", C.c_int),
        ("lambda_", C.c_double), ("gamma", C.c_double), ("wbcoeff", C.c_double),
        ("w0coeff", C.c_double), ("lambda1", C.c_double), ("lambda2", C.c_double),
        ("alpha", C.c_double), ("u0", C.c_double), ("umax", C.c_double),
        ("acore", C.c_double), ("ubcore", C.c_double), ("nonequilibrium", C.c_int),
        ("noneq_tmax", C.c_double), ("work", C.c_double), ("m_l1", C.c_double),
        ("m_l2", C.c_double), ("m_u0", C.c_double), ("m_w0", C.c_double),
        ("b_l1", C.c_double), ("b_l2", C.c_double), ("b_u0", C.c_double),
        ("b_w0", C.c_double), ("temperature", C.c_double), ("friction", C.c_double),
        ("step_size", C.c_double), ("m_alpha", C.c_double),
    ]


class orc_scalars(C.Structure):
    _fields_ = [("bind_e", C.c_double), ("pot_energy", C.c_double), ("ebias", C.c_double),
                ("bfp", C.c_double), ("fp", C.c_double), ("sp", C.c_double)]


class orc_nb(C.Structure):
    _fields_ = [("n", C.c_int), ("q", _dp), ("sigma", _dp), ("eps", _dp), ("n_excl", C.c_int),
                ("excl", _ip), ("box", C.c_double * 3), ("periodic", C.c_int),
                ("cutoff", C.c_double), ("eps_rf", C.c_double), ("geometric", C.c_int)]


class orc_bonded_t(C.Structure):
    _fields_ = [("n_bonds", C.c_int), ("bond_atoms", _ip), ("bond_r0", _dp), ("bond_k", _dp),
                ("n_angles", C.c_int), ("angle_atoms", _ip), ("angle_t0", _dp),
                ("angle_k", _dp)]


class orc_step_report(C.Structure):
    _fields_ = [("E1", C.c_double), ("E2", C.c_double), ("Eb", C.c_double), ("s", orc_scalars)]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load():
    if not os.path.exists(_LIB):
        build()
    lib = C.CDLL(_LIB)
    lib.orc_softcore.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, _dp, _dp]
    lib.orc_execute_scalars.argtypes = [C.POINTER(orc_params), C.c_double, C.c_double,
                                        C.c_double, C.c_double, C.POINTER(orc_scalars)]
    lib.orc_pair_energy.restype = C.c_double
    lib.orc_pair_energy.argtypes = [C.c_double] * 6 + [_dp]
    for f in ("orc_nonbonded_brute", "orc_nonbonded_cells", "orc_nonbonded"):
        getattr(lib, f).restype = C.c_double
        getattr(lib, f).argtypes = [C.POINTER(orc_nb), _dp, _dp, C.POINTER(C.c_longlong)]
    lib.orc_bonded.restype = C.c_double
    lib.orc_bonded.argtypes = [C.POINTER(orc_bonded_t), _dp, _dp]
    lib.orc_sdm_reference_eval.argtypes = [C.POINTER(orc_nb), C.POINTER(orc_bonded_t), _dp, _dp,
                                           _dp, _dp, _dp, _dp, _dp, _dp]
    lib.orc_sdm_direct.argtypes = [C.POINTER(orc_nb), _dp, _dp, _dp, _dp, _dp, _dp,
                                   C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]
    lib.orc_force_mix.argtypes = [C.c_int, C.c_double, _dp, _dp, _dp, _dp]
    lib.orc_du_dparams.argtypes = [C.POINTER(orc_nb), _dp, _dp, C.c_int, _ip, _dp, _dp, _dp]
    for f in ("orc_langevin_update", "orc_langevin_update_cl"):
        getattr(lib, f).argtypes = [C.c_int, _dp, _dp, _dp, _dp, C.c_double, C.c_double,
                                    C.c_double, _dp]
    lib.orc_kinetic_energy.restype = C.c_double
    lib.orc_kinetic_energy.argtypes = [C.c_int, _dp, _dp]
    lib.orc_philox4x32_10.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                      C.POINTER(C.c_uint32)]
    lib.orc_gauss3.argtypes = [C.c_uint64, C.c_uint32, C.c_uint64, C.c_uint32, _dp]
    lib.orc_ref_step.argtypes = [C.POINTER(orc_nb), C.POINTER(orc_bonded_t), _dp, _dp,
                                 C.POINTER(orc_params), _dp, _dp, _dp, C.POINTER(C.c_uint64),
                                 C.c_uint64, C.c_uint32, C.c_int, C.c_int,
                                 C.POINTER(orc_step_report)]
    lib.orc_ref_step_replicas.argtypes = [C.POINTER(orc_nb), C.POINTER(orc_bonded_t), _dp, _dp,
                                          C.POINTER(orc_params), _dp, _dp, _dp,
                                          C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.c_int,
                                          C.c_int, C.c_int, C.c_int]
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_dp)


def _i(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(_ip)


class Topology:
    """Holds the C structs (and keeps their numpy buffers alive) for one system."""

    def __init__(self, s, periodic=None):
        self._keep = []
        q, qp = _d(s.charge)
        sg, sgp = _d(s.sigma)
        ep, epp = _d(s.epsilon)
        ex, exp_ = _i(np.asarray(s.exclusions).reshape(-1))
        self._keep += [q, sg, ep, ex]
        nb = orc_nb()
        nb.n = s.n
        nb.q, nb.sigma, nb.eps = qp, sgp, epp
        nb.n_excl = int(np.asarray(s.exclusions).reshape(-1, 2).shape[0])
        nb.excl = exp_
        for k in range(3):
            nb.box[k] = float(s.box[k])
        nb.periodic = int(s.periodic if periodic is None else periodic)
        nb.cutoff = float(s.cutoff)
        nb.eps_rf = float(s.eps_rf)
        nb.geometric = int(getattr(s, "lj_rule", 0))
        self.nb = nb
        ba, bap = _i(np.asarray(s.bonds).reshape(-1))
        br, brp = _d(s.bond_r0)
        bk, bkp = _d(s.bond_k)
        aa, aap = _i(np.asarray(s.angles).reshape(-1))
        at, atp = _d(s.angle_t0)
        ak, akp = _d(s.angle_k)
        self._keep += [ba, br, bk, aa, at, ak]
        b = orc_bonded_t()
        b.n_bonds = int(np.asarray(s.bonds).reshape(-1, 2).shape[0])
        b.bond_atoms, b.bond_r0, b.bond_k = bap, brp, bkp
        b.n_angles = int(np.asarray(s.angles).reshape(-1, 3).shape[0])
        b.angle_atoms, b.angle_t0, b.angle_k = aap, atp, akp
        self.bonded = b
        dsp, dspp = _d(np.asarray(s.displacement).reshape(-1))
        m, mp = _d(s.mass)
        self._keep += [dsp, m]
        self.displ, self.displ_p = dsp, dspp
        self.mass, self.mass_p = m, mp
        self.n = s.n


def make_params(p) -> orc_params:
    o = orc_params()
    for name, _t in orc_params._fields_:
        setattr(o, name, getattr(p, name))
    return o


def softcore(method, u, umax, a, ub):
    up, fp = C.c_double(), C.c_double()
    rc = lib().orc_softcore(method, u, umax, a, ub, C.byref(up), C.byref(fp))
    if rc:
        raise ValueError("Unknown soft core method")
    return up.value, fp.value


def execute_scalars(params: orc_params, E1, E2, Eb, time):
    out = orc_scalars()
    rc = lib().orc_execute_scalars(C.byref(params), E1, E2, Eb, time, C.byref(out))
    if rc:
        raise ValueError("Unknown soft core method")
    return out


def pair_energy(r, qq, sig, eps, rc, eps_rf):
    f = C.c_double()
    e = lib().orc_pair_energy(r, qq, sig, eps, rc, eps_rf, C.byref(f))
    return e, f.value


def nonbonded(topo: Topology, x, method="cells"):
    x, xp = _d(np.asarray(x).reshape(-1))
    F = np.zeros(3 * topo.n)
    npairs = C.c_longlong()
    fn = {"cells": lib().orc_nonbonded_cells, "brute": lib().orc_nonbonded_brute}[method]
    E = fn(C.byref(topo.nb), xp, F.ctypes.data_as(_dp), C.byref(npairs))
    return E, F.reshape(-1, 3), npairs.value


def bonded(topo: Topology, x):
    x, xp = _d(np.asarray(x).reshape(-1))
    F = np.zeros(3 * topo.n)
    E = lib().orc_bonded(C.byref(topo.bonded), xp, F.ctypes.data_as(_dp))
    return E, F.reshape(-1, 3)


def reference_eval(topo: Topology, x, displ=None):
    x, xp = _d(np.asarray(x).reshape(-1))
    if displ is None:
        dp = topo.displ_p
    else:
        dd, dp = _d(np.asarray(displ).reshape(-1))
    n = topo.n
    E1, E2, Eb = C.c_double(), C.c_double(), C.c_double()
    F1, F2, Fb = np.zeros(3 * n), np.zeros(3 * n), np.zeros(3 * n)
    lib().orc_sdm_reference_eval(C.byref(topo.nb), C.byref(topo.bonded), xp, dp, C.byref(E1),
                                 C.byref(E2), C.byref(Eb), F1.ctypes.data_as(_dp),
                                 F2.ctypes.data_as(_dp), Fb.ctypes.data_as(_dp))
    return (E1.value, E2.value, Eb.value, F1.reshape(-1, 3), F2.reshape(-1, 3),
            Fb.reshape(-1, 3))


def sdm_direct(topo: Topology, x, displ=None):
    x, xp = _d(np.asarray(x).reshape(-1))
    if displ is None:
        dp = topo.displ_p
    else:
        dd, dp = _d(np.asarray(displ).reshape(-1))
    n = topo.n
    U0, u = C.c_double(), C.c_double()
    F0, dF = np.zeros(3 * n), np.zeros(3 * n)
    p0, pd = C.c_longlong(), C.c_longlong()
    lib().orc_sdm_direct(C.byref(topo.nb), xp, dp, C.byref(U0), C.byref(u),
                         F0.ctypes.data_as(_dp), dF.ctypes.data_as(_dp), C.byref(p0),
                         C.byref(pd))
    return U0.value, u.value, F0.reshape(-1, 3), dF.reshape(-1, 3), p0.value, pd.value


def du_dparams(topo: Topology, x, sel):
    x, xp = _d(np.asarray(x).reshape(-1))
    sel, selp = _i(sel)
    m = len(sel)
    dq, ds, de = np.zeros(m), np.zeros(m), np.zeros(m)
    lib().orc_du_dparams(C.byref(topo.nb), xp, topo.displ_p, m, selp, dq.ctypes.data_as(_dp),
                         ds.ctypes.data_as(_dp), de.ctypes.data_as(_dp))
    return dq, ds, de


def langevin_update(x, v, F, mass, dt, friction, temperature, xi, form="reference"):
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1).copy()
    v = np.ascontiguousarray(v, dtype=np.float64).reshape(-1).copy()
    F, Fp = _d(np.asarray(F).reshape(-1))
    m, mp = _d(mass)
    xi, xip = _d(np.asarray(xi).reshape(-1))
    fn = lib().orc_langevin_update if form == "reference" else lib().orc_langevin_update_cl
    fn(len(m), x.ctypes.data_as(_dp), v.ctypes.data_as(_dp), Fp, mp, dt, friction, temperature,
       xip)
    return x.reshape(-1, 3), v.reshape(-1, 3)


def kinetic_energy(v, mass):
    v, vp = _d(np.asarray(v).reshape(-1))
    m, mp = _d(mass)
    return lib().orc_kinetic_energy(len(m), vp, mp)


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return list(o)


def gauss3(seed, replica, step, atom):
    o = (C.c_double * 3)()
    lib().orc_gauss3(seed, replica, step, atom, o)
    return list(o)


def ref_step(topo: Topology, params: orc_params, x, v, time, step, seed, replica, nsteps,
             use_cl_langevin=True):
    """Advance (x, v) by nsteps reference steps; returns (x, v, time, step, report)."""
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1).copy()
    v = np.ascontiguousarray(v, dtype=np.float64).reshape(-1).copy()
    t = C.c_double(time)
    st = C.c_uint64(step)
    rep = orc_step_report()
    rc = lib().orc_ref_step(C.byref(topo.nb), C.byref(topo.bonded), topo.displ_p, topo.mass_p,
                            C.byref(params), x.ctypes.data_as(_dp), v.ctypes.data_as(_dp),
                            C.byref(t), C.byref(st), seed, replica, nsteps,
                            int(use_cl_langevin), C.byref(rep))
    if rc:
        raise ValueError("Unknown soft core method")
    return x.reshape(-1, 3), v.reshape(-1, 3), t.value, st.value, rep


def ref_step_replicas(topo: Topology, params, xs, vs, seeds, nsteps, threads,
                      use_cl_langevin=True):
    """R replicas (lists of orc_params, (n,3) positions and velocities, seeds), each
    advanced by nsteps reference steps from time 0, step 0, on `threads` OpenMP threads;
    returns (xs, vs) as (R, n, 3) arrays."""
    R = len(params)
    pa = (orc_params * R)(*params)
    x = np.ascontiguousarray(np.stack(xs), dtype=np.float64).copy()
    v = np.ascontiguousarray(np.stack(vs), dtype=np.float64).copy()
    t = np.zeros(R)
    st = (C.c_uint64 * R)()
    sd = (C.c_uint64 * R)(*[int(k) for k in seeds])
    rc = lib().orc_ref_step_replicas(C.byref(topo.nb), C.byref(topo.bonded), topo.displ_p,
                                     topo.mass_p, pa, x.ctypes.data_as(_dp),
                                     v.ctypes.data_as(_dp), t.ctypes.data_as(_dp), st, sd, R,
                                     nsteps, int(use_cl_langevin), int(threads))
    if rc:
        raise ValueError("Unknown soft core method")
    return x, v

