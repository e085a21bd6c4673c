"""ctypes binding of libsdm_hip.so (include/sdm_hip.h).

The HIP library is the only compute path: if it is missing, or no HIP device is
visible, every entry point raises instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsdm_hip.so")

SDM_OK = 0
SDM_ERR_INVALID = -1
SDM_ERR_HIP = -2
SDM_ERR_STATE = -3
SDM_ERR_NUMERIC = -4
SDM_ERR_CAPACITY = -5
SDM_ERR_UNSUPPORTED = -6

FLAG_NONFINITE, FLAG_OVERFLOW, FLAG_CAPACITY = 1, 2, 4
FORCE_HYBRID, FORCE_STATE1, FORCE_STATE2, FORCE_BONDED = 0, 1, 2, 3

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class SDMError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"sdm error {code}: {msg}")
        self.code = code


class sdm_system_desc(C.Structure):
    _fields_ = [
        ("n_atoms", C.c_int32), ("charge", _dp), ("sigma", _dp), ("epsilon", _dp), ("mass", _dp),
        ("n_exclusions", C.c_int32), ("exclusions", _ip), ("n_bonds", C.c_int32),
        ("bond_atoms", _ip), ("bond_r0", _dp), ("bond_k", _dp), ("n_angles", C.c_int32),
        ("angle_atoms", _ip), ("angle_theta0", _dp), ("angle_k", _dp), ("displacement", _dp),
        ("box", C.c_double * 3), ("periodic", C.c_int32), ("cutoff", C.c_double),
        ("eps_rf", C.c_double), ("lj_rule", C.c_int32),
    ]


class sdm_params(C.Structure):
    _fields_ = [
        ("bias_method", C.c_int32), ("softcore_method", C.c_int32), ("lambda_", C.c_double),
        ("gamma", C.c_double), ("wbcoeff", C.c_double), ("w0coeff", C.c_double),
        ("lambda1", C.c_double), ("lambda2", C.c_double), ("alpha", C.c_double),
        ("u0", C.c_double), ("umax", C.c_double), ("acore", C.c_double), ("ubcore", C.c_double),
        ("nonequilibrium", C.c_int32), ("noneq_tmax", C.c_double), ("work", C.c_double),
        ("m_l1", C.c_double), ("m_l2", C.c_double), ("m_u0", C.c_double), ("m_w0", C.c_double),
        ("b_l1", C.c_double), ("b_l2", C.c_double), ("b_u0", C.c_double), ("b_w0", C.c_double),
        ("temperature", C.c_double), ("friction", C.c_double), ("step_size", C.c_double),
        ("seed", C.c_uint64), ("m_alpha", C.c_double),
    ]


class sdm_energies(C.Structure):
    _fields_ = [
        ("U0", C.c_double), ("U1", C.c_double), ("u", C.c_double), ("Ub", C.c_double),
        ("bind_e", C.c_double), ("fp", C.c_double), ("W", C.c_double), ("Wp", C.c_double),
        ("sp", C.c_double), ("pot_energy", C.c_double), ("lambda_", C.c_double),
        ("lambda1", C.c_double), ("lambda2", C.c_double), ("u0", C.c_double), ("w0", C.c_double),
        ("work", C.c_double), ("time", C.c_double), ("step", C.c_int64), ("flags", C.c_int32),
        ("pad", C.c_int32),
    ]


class sdm_tuning(C.Structure):
    _fields_ = [("pair_waves_per_group", C.c_int32), ("cell_cells_per_dim", C.c_int32),
                ("use_graph", C.c_int32), ("pair_newton3", C.c_int32),
                ("pair_group", C.c_int32), ("sort_interval", C.c_int32),
                ("energy_every_step", C.c_int32), ("pair_split", C.c_int32),
                ("side_fork", C.c_int32), ("list_skin", C.c_float)]


class sdm_centroid_restraint(C.Structure):
    _fields_ = [("n_lig", C.c_int32), ("n_rcpt", C.c_int32),
                ("lig_atoms", C.POINTER(C.c_int32)), ("rcpt_atoms", C.POINTER(C.c_int32)),
                ("lig_weights", C.POINTER(C.c_double)), ("rcpt_weights", C.POINTER(C.c_double)),
                ("kfcm", C.c_double), ("tolcm", C.c_double), ("offset", C.c_double * 3),
                ("do_angles", C.c_int32), ("rcpt_ref", C.c_int32 * 3),
                ("lig_ref", C.c_int32 * 3), ("pad", C.c_int32), ("kf", C.c_double * 3),
                ("lo", C.c_double * 3), ("hi", C.c_double * 3), ("control", C.c_double)]


class sdm_alignment_restraint(C.Structure):
    _fields_ = [("liga_ref", C.c_int32 * 3), ("ligb_ref", C.c_int32 * 3),
                ("kfdispl", C.c_double), ("ktheta", C.c_double), ("kpsi", C.c_double),
                ("offset", C.c_double * 3)]


_I, _D = C.POINTER(C.c_int32), C.POINTER(C.c_double)


class sdm_forcefield_terms(C.Structure):
    _fields_ = [("n_torsions", C.c_int32), ("torsion_atoms", _I), ("torsion_n", _I),
                ("torsion_phase", _D), ("torsion_k", _D),
                ("n_impropers", C.c_int32), ("improper_atoms", _I), ("improper_phi0", _D),
                ("improper_k", _D),
                ("n_posre_fb", C.c_int32), ("posre_fb_atoms", _I), ("posre_fb_x0", _D),
                ("posre_fb_k", _D), ("posre_fb_tol", _D),
                ("n_posre_harm", C.c_int32), ("posre_harm_atoms", _I), ("posre_harm_x0", _D),
                ("posre_harm_k", _D),
                ("n_exceptions", C.c_int32), ("exception_atoms", _I), ("exception_qq", _D),
                ("exception_sigma", _D), ("exception_epsilon", _D),
                ("n_constraints", C.c_int32), ("constraint_atoms", _I),
                ("constraint_length", _D), ("constraint_tol", C.c_double)]


# every exported symbol of include/sdm_hip.h
EXPORTS = [
    "sdm_last_error", "sdm_create", "sdm_destroy", "sdm_num_replicas", "sdm_set_params",
    "sdm_get_params", "sdm_set_time", "sdm_set_nonbonded_params", "sdm_set_positions",
    "sdm_get_positions", "sdm_set_velocities", "sdm_get_velocities", "sdm_set_positions_device",
    "sdm_get_positions_device", "sdm_compute", "sdm_step", "sdm_get_energies", "sdm_clear_flags",
    "sdm_compute_components", "sdm_get_forces", "sdm_kinetic_energy", "sdm_set_u_history",
    "sdm_get_u_history", "sdm_count_pairs", "sdm_time_pair_kernel", "sdm_autotune",
    "sdm_get_tuning", "sdm_set_tuning", "sdm_du_dparams_frames", "sdm_set_restraints",
    "sdm_set_forcefield_terms", "sdm_set_cm_motion_removal", "sdm_set_constraint_tolerance",
    "sdm_du_dparams_frames_noalloc", "sdm_set_virtual_sites",
    "registerPlatforms", "registerKernelFactories", "registerSDMHIPKernelFactories",
]

_lib = None


def load(path: str = LIB_PATH):
    """Loads the HIP library (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is missing: build the HIP extension first (__graft_entry__.build())")
    lib = C.CDLL(path)
    vp = C.c_void_p
    lib.sdm_last_error.restype = C.c_char_p
    lib.sdm_create.argtypes = [C.POINTER(sdm_system_desc), C.c_int32, C.POINTER(vp)]
    lib.sdm_destroy.argtypes = [vp]
    lib.sdm_num_replicas.argtypes = [vp, _ip, _ip]
    lib.sdm_set_params.argtypes = [vp, C.c_int32, C.POINTER(sdm_params)]
    lib.sdm_get_params.argtypes = [vp, C.c_int32, C.POINTER(sdm_params)]
    lib.sdm_set_time.argtypes = [vp, C.c_int32, C.c_double, C.c_int64]
    lib.sdm_set_nonbonded_params.argtypes = [vp, _dp, _dp, _dp]
    for f in ("sdm_set_positions", "sdm_get_positions", "sdm_set_velocities", "sdm_get_velocities"):
        getattr(lib, f).argtypes = [vp, C.c_int32, _dp]
    lib.sdm_set_positions_device.argtypes = [vp, vp]
    lib.sdm_get_positions_device.argtypes = [vp, vp]
    lib.sdm_compute.argtypes = [vp]
    lib.sdm_step.argtypes = [vp, C.c_int32]
    lib.sdm_get_energies.argtypes = [vp, C.POINTER(sdm_energies)]
    lib.sdm_clear_flags.argtypes = [vp]
    lib.sdm_compute_components.argtypes = [vp]
    lib.sdm_get_forces.argtypes = [vp, C.c_int32, C.c_int32, _dp]
    lib.sdm_kinetic_energy.argtypes = [vp, _dp]
    lib.sdm_set_u_history.argtypes = [vp, C.c_int32, C.c_int32]
    lib.sdm_get_u_history.argtypes = [vp, C.c_int32, _dp, C.c_int32, _ip]
    lib.sdm_count_pairs.argtypes = [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    lib.sdm_time_pair_kernel.argtypes = [vp, C.c_int32, _dp]
    lib.sdm_autotune.argtypes = [vp, C.c_int32, C.POINTER(sdm_tuning)]
    lib.sdm_get_tuning.argtypes = [vp, C.POINTER(sdm_tuning)]
    lib.sdm_set_tuning.argtypes = [vp, C.POINTER(sdm_tuning)]
    lib.sdm_du_dparams_frames.argtypes = [vp, vp, C.c_int32, _ip, C.c_int32, vp, vp]
    lib.sdm_du_dparams_frames_noalloc.argtypes = [vp, vp, C.c_int32, _ip, C.c_int32, vp, vp]
    lib.sdm_set_restraints.argtypes = [vp, C.POINTER(sdm_centroid_restraint), C.c_int32,
                                       C.POINTER(sdm_alignment_restraint), C.c_int32]
    lib.sdm_set_forcefield_terms.argtypes = [vp, C.POINTER(sdm_forcefield_terms)]
    lib.sdm_set_cm_motion_removal.argtypes = [vp, C.c_int32]
    lib.sdm_set_constraint_tolerance.argtypes = [vp, C.c_double]
    lib.sdm_set_virtual_sites.argtypes = [vp, C.c_int32, _ip, _ip, _dp]
    for f in ("sdm_create", "sdm_destroy", "sdm_num_replicas", "sdm_set_params", "sdm_get_params",
              "sdm_set_time", "sdm_set_nonbonded_params", "sdm_set_positions",
              "sdm_get_positions", "sdm_set_velocities", "sdm_get_velocities",
              "sdm_set_positions_device", "sdm_get_positions_device", "sdm_compute", "sdm_step",
              "sdm_get_energies", "sdm_clear_flags", "sdm_compute_components", "sdm_get_forces",
              "sdm_kinetic_energy", "sdm_set_u_history", "sdm_get_u_history", "sdm_count_pairs",
              "sdm_time_pair_kernel", "sdm_autotune", "sdm_get_tuning", "sdm_set_tuning",
              "sdm_du_dparams_frames", "sdm_set_restraints", "sdm_set_forcefield_terms",
              "sdm_set_cm_motion_removal", "sdm_set_constraint_tolerance",
              "sdm_du_dparams_frames_noalloc", "sdm_set_virtual_sites"):
        getattr(lib, f).restype = C.c_int
    for f in ("registerPlatforms", "registerKernelFactories", "registerSDMHIPKernelFactories"):
        getattr(lib, f).restype = None
        getattr(lib, f).argtypes = []
    _lib = lib
    return lib


def check(rc):
    if rc != SDM_OK:
        msg = _lib.sdm_last_error().decode() if _lib is not None else ""
        raise SDMError(rc, msg)


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_dp)


def _i(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(_ip)


def params_to_c(p) -> sdm_params:
    o = sdm_params()
    for name, _t in sdm_params._fields_:
        setattr(o, name, getattr(p, name))
    return o


def params_from_c(o: sdm_params, p):
    for name, _t in sdm_params._fields_:
        setattr(p, name, getattr(o, name))
    return p


class NativeContext:
    """Owns one sdm_context (R replicas of one topology) on the current HIP device."""

    def __init__(self, system, n_replicas: int = 1, displacement=None):
        lib = load()
        self._lib = lib
        s = system
        self.n = s.n
        self.R = n_replicas
        keep = []

        def d(a):
            arr, p = _d(a)
            keep.append(arr)
            return p

        def i(a):
            arr, p = _i(a)
            keep.append(arr)
            return p

        desc = sdm_system_desc()
        desc.n_atoms = s.n
        desc.charge, desc.sigma, desc.epsilon, desc.mass = (d(s.charge), d(s.sigma),
                                                             d(s.epsilon), d(s.mass))
        ex = np.asarray(s.exclusions, dtype=np.int32).reshape(-1, 2)
        desc.n_exclusions = ex.shape[0]
        desc.exclusions = i(ex.reshape(-1))
        b = np.asarray(s.bonds, dtype=np.int32).reshape(-1, 2)
        desc.n_bonds = b.shape[0]
        desc.bond_atoms, desc.bond_r0, desc.bond_k = i(b.reshape(-1)), d(s.bond_r0), d(s.bond_k)
        a = np.asarray(s.angles, dtype=np.int32).reshape(-1, 3)
        desc.n_angles = a.shape[0]
        desc.angle_atoms, desc.angle_theta0, desc.angle_k = (i(a.reshape(-1)), d(s.angle_t0),
                                                             d(s.angle_k))
        disp = s.displacement if displacement is None else displacement
        desc.displacement = d(np.asarray(disp, dtype=np.float64).reshape(-1))
        for k in range(3):
            desc.box[k] = float(s.box[k])
        desc.periodic = int(s.periodic)
        desc.cutoff = float(s.cutoff)
        desc.eps_rf = float(s.eps_rf)
        desc.lj_rule = int(getattr(s, "lj_rule", 0))
        h = C.c_void_p()
        check(lib.sdm_create(C.byref(desc), n_replicas, C.byref(h)))
        self.h = h
        if getattr(s, "restraints", None):
            self.set_restraints(s.restraints)
        if getattr(s, "terms", None) is not None:
            self.set_forcefield_terms(s.terms)
            if getattr(s.terms, "remove_cm_motion", False):
                self.set_cm_motion_removal(1)
        if getattr(s, "virtual_sites", None) is not None:
            self.set_virtual_sites(s.virtual_sites)

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self._lib.sdm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- parameters / state ----
    def set_params(self, replica, p):
        c = params_to_c(p)
        check(self._lib.sdm_set_params(self.h, replica, C.byref(c)))

    def get_params(self, replica, into):
        c = sdm_params()
        check(self._lib.sdm_get_params(self.h, replica, C.byref(c)))
        return params_from_c(c, into)

    def set_time(self, replica, time, step):
        check(self._lib.sdm_set_time(self.h, replica, float(time), int(step)))

    def set_nonbonded_params(self, q, sig, eps):
        q, qp = _d(q)
        s, sp = _d(sig)
        e, ep = _d(eps)
        check(self._lib.sdm_set_nonbonded_params(self.h, qp, sp, ep))

    def set_restraints(self, restraints):
        """Replaces the restraints (a list of restraints.CentroidRestraint and
        restraints.AlignmentRestraint; [] removes them)."""
        cen = [r for r in restraints if hasattr(r, "lig_atoms")]
        ali = [r for r in restraints if hasattr(r, "liga_ref")]
        keep = []
        ca = (sdm_centroid_restraint * max(1, len(cen)))()
        for k, r in enumerate(cen):
            c = ca[k]
            for name in ("lig_atoms", "rcpt_atoms"):
                a = np.ascontiguousarray(r.__dict__[name], dtype=np.int32)
                keep.append(a)
                setattr(c, name, a.ctypes.data_as(C.POINTER(C.c_int32)))
            c.n_lig, c.n_rcpt = len(r.lig_atoms), len(r.rcpt_atoms)
            for name in ("lig_weights", "rcpt_weights"):
                w = r.__dict__[name]
                if w is not None:
                    w = np.ascontiguousarray(w, dtype=np.float64)
                    keep.append(w)
                    setattr(c, name, w.ctypes.data_as(C.POINTER(C.c_double)))
            c.kfcm, c.tolcm, c.control = float(r.kfcm), float(r.tolcm), float(r.control)
            c.do_angles = int(r.do_angles)
            for j in range(3):
                c.offset[j] = float(r.offset[j])
                c.rcpt_ref[j] = int(r.rcpt_ref[j])
                c.lig_ref[j] = int(r.lig_ref[j])
                c.kf[j], c.lo[j], c.hi[j] = float(r.kf[j]), float(r.lo[j]), float(r.hi[j])
        aa = (sdm_alignment_restraint * max(1, len(ali)))()
        for k, r in enumerate(ali):
            a = aa[k]
            for j in range(3):
                a.liga_ref[j], a.ligb_ref[j] = int(r.liga_ref[j]), int(r.ligb_ref[j])
                a.offset[j] = float(r.offset[j])
            a.kfdispl, a.ktheta, a.kpsi = float(r.kfdispl), float(r.ktheta), float(r.kpsi)
        check(self._lib.sdm_set_restraints(self.h, ca, len(cen), aa, len(ali)))

    def set_cm_motion_removal(self, frequency):
        check(self._lib.sdm_set_cm_motion_removal(self.h, int(frequency)))

    def set_constraint_tolerance(self, tol):
        check(self._lib.sdm_set_constraint_tolerance(self.h, float(tol)))

    def set_virtual_sites(self, vs):
        """Virtual sites (an object with kind (n,), atoms (n,4): site, p1, p2, p3 and
        weights (n,3), as sdm_set_virtual_sites takes them); None removes them."""
        if vs is None or len(vs.kind) == 0:
            check(self._lib.sdm_set_virtual_sites(self.h, 0, None, None, None))
            return
        kind = np.ascontiguousarray(vs.kind, dtype=np.int32)
        atoms = np.ascontiguousarray(np.asarray(vs.atoms).reshape(-1, 4), dtype=np.int32)
        w = np.ascontiguousarray(np.asarray(vs.weights).reshape(-1, 3), dtype=np.float64)
        check(self._lib.sdm_set_virtual_sites(self.h, len(kind), kind.ctypes.data_as(_I),
                                              atoms.ctypes.data_as(_I), w.ctypes.data_as(_D)))

    def set_forcefield_terms(self, ff, constraint_tol=0.0):
        """Torsions, impropers, positional restraints, 1-4 exceptions and constraints of a
        dms.DMSForceField (or any object with its array attributes); None removes them."""
        if ff is None:
            check(self._lib.sdm_set_forcefield_terms(self.h, None))
            return
        t = sdm_forcefield_terms()
        keep = []

        def put(field, arr, dtype, count_field=None, rows=None):
            a = np.ascontiguousarray(arr, dtype=dtype)
            keep.append(a)
            setattr(t, field, a.ctypes.data_as(_I if dtype == np.int32 else _D))
            if count_field:
                setattr(t, count_field, int(a.shape[0] if rows is None else rows))
        put("torsion_atoms", ff.torsions, np.int32, "n_torsions")
        put("torsion_n", ff.torsion_n, np.int32)
        put("torsion_phase", ff.torsion_phase, np.float64)
        put("torsion_k", ff.torsion_k, np.float64)
        put("improper_atoms", ff.impropers, np.int32, "n_impropers")
        put("improper_phi0", ff.improper_t0, np.float64)
        put("improper_k", ff.improper_k, np.float64)
        put("posre_fb_atoms", ff.posre_fb_atoms, np.int32, "n_posre_fb")
        put("posre_fb_x0", ff.posre_fb_x0, np.float64)
        put("posre_fb_k", ff.posre_fb_hk, np.float64)
        put("posre_fb_tol", ff.posre_fb_tol, np.float64)
        put("posre_harm_atoms", ff.posre_h_atoms, np.int32, "n_posre_harm")
        put("posre_harm_x0", ff.posre_h_x0, np.float64)
        put("posre_harm_k", ff.posre_h_hk, np.float64)
        put("exception_atoms", ff.exception_pairs, np.int32, "n_exceptions")
        put("exception_qq", ff.exception_qq, np.float64)
        put("exception_sigma", ff.exception_sigma, np.float64)
        put("exception_epsilon", ff.exception_epsilon, np.float64)
        put("constraint_atoms", ff.constraints, np.int32, "n_constraints")
        put("constraint_length", ff.constraint_r, np.float64)
        t.constraint_tol = float(constraint_tol)
        check(self._lib.sdm_set_forcefield_terms(self.h, C.byref(t)))

    def set_positions(self, replica, x):
        x, xp = _d(np.asarray(x).reshape(-1))
        check(self._lib.sdm_set_positions(self.h, replica, xp))

    def get_positions(self, replica):
        out = np.zeros(3 * self.n)
        check(self._lib.sdm_get_positions(self.h, replica, out.ctypes.data_as(_dp)))
        return out.reshape(-1, 3)

    def set_velocities(self, replica, v):
        v, vp = _d(np.asarray(v).reshape(-1))
        check(self._lib.sdm_set_velocities(self.h, replica, vp))

    def get_velocities(self, replica):
        out = np.zeros(3 * self.n)
        check(self._lib.sdm_get_velocities(self.h, replica, out.ctypes.data_as(_dp)))
        return out.reshape(-1, 3)

    def set_positions_device(self, ptr):
        check(self._lib.sdm_set_positions_device(self.h, C.c_void_p(ptr)))

    def get_positions_device(self, ptr):
        check(self._lib.sdm_get_positions_device(self.h, C.c_void_p(ptr)))

    # ---- evaluation ----
    def compute(self):
        check(self._lib.sdm_compute(self.h))

    def compute_components(self):
        check(self._lib.sdm_compute_components(self.h))

    def step(self, nsteps):
        check(self._lib.sdm_step(self.h, int(nsteps)))

    def energies(self):
        out = (sdm_energies * self.R)()
        check(self._lib.sdm_get_energies(self.h, out))
        return [dict((n, getattr(e, n)) for n, _ in sdm_energies._fields_) for e in out]

    def clear_flags(self):
        check(self._lib.sdm_clear_flags(self.h))

    def forces(self, replica, which=FORCE_HYBRID):
        out = np.zeros(3 * self.n)
        check(self._lib.sdm_get_forces(self.h, replica, which, out.ctypes.data_as(_dp)))
        return out.reshape(-1, 3)

    def kinetic_energy(self):
        out = np.zeros(self.R)
        check(self._lib.sdm_kinetic_energy(self.h, out.ctypes.data_as(_dp)))
        return out

    def set_u_history(self, every, capacity):
        check(self._lib.sdm_set_u_history(self.h, every, capacity))

    def u_history(self, replica, capacity):
        out = np.zeros(capacity)
        cnt = C.c_int32()
        check(self._lib.sdm_get_u_history(self.h, replica, out.ctypes.data_as(_dp), capacity,
                                          C.byref(cnt)))
        return out[: cnt.value]

    def count_pairs(self):
        p0 = (C.c_int64 * self.R)()
        pd = (C.c_int64 * self.R)()
        check(self._lib.sdm_count_pairs(self.h, p0, pd))
        return np.array(list(p0), dtype=np.int64), np.array(list(pd), dtype=np.int64)

    def time_pair_kernel(self, nsteps):
        ms = C.c_double()
        check(self._lib.sdm_time_pair_kernel(self.h, nsteps, C.byref(ms)))
        return ms.value

    def autotune(self, steps):
        t = sdm_tuning()
        check(self._lib.sdm_autotune(self.h, steps, C.byref(t)))
        return dict((n, getattr(t, n)) for n, _ in sdm_tuning._fields_)

    def get_tuning(self):
        t = sdm_tuning()
        check(self._lib.sdm_get_tuning(self.h, C.byref(t)))
        return dict((n, getattr(t, n)) for n, _ in sdm_tuning._fields_)

    def set_tuning(self, **kw):
        t = sdm_tuning()
        check(self._lib.sdm_get_tuning(self.h, C.byref(t)))
        for k, v in kw.items():
            setattr(t, k, float(v) if k == "list_skin" else int(v))
        check(self._lib.sdm_set_tuning(self.h, C.byref(t)))

    def du_dparams_frames_noalloc(self, frames_ptr, nframes, sel, grad_ptr, u_ptr):
        sel, selp = _i(sel)
        check(self._lib.sdm_du_dparams_frames_noalloc(
            self.h, C.c_void_p(frames_ptr), int(nframes), selp, len(sel),
            C.c_void_p(grad_ptr), C.c_void_p(u_ptr)))

    def du_dparams_frames(self, frames_ptr, nframes, sel, grad_ptr, u_ptr):
        sel, selp = _i(sel)
        check(self._lib.sdm_du_dparams_frames(self.h, C.c_void_p(frames_ptr), int(nframes), selp,
                                              len(sel), C.c_void_p(grad_ptr), C.c_void_p(u_ptr)))
