"""CPU checks of the C ABI: the library loads, exports every symbol of include/sdm_hip.h,
and rejects bad input before touching a device."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from openmm_sdm_plugin_amd import _native, systems

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    text = open(os.path.join(ROOT, "include", "sdm_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sdm_\w+|register\w+)\s*\(", text)))


def test_header_lists_match_binding():
    assert header_symbols() == sorted(_native.EXPORTS)


def test_library_exports_every_symbol():
    lib = _native.load()
    for name in header_symbols():
        assert hasattr(lib, name), name


def test_struct_sizes_match_header_layout():
    # sdm_params: 2 ints, 11 doubles, int(+pad), 15 doubles... checked against the C
    # layout rule: every double is 8-aligned.
    assert C.sizeof(_native.sdm_params) == 8 + 8 * 11 + 8 + 8 * 13 + 8 + 8
    assert C.sizeof(_native.sdm_energies) == 8 * 17 + 8 + 8
    assert C.sizeof(_native.sdm_tuning) == 40
    # sdm_system_desc: ints padded to the following pointers/doubles; lj_rule last
    assert C.sizeof(_native.sdm_system_desc) == 184
    assert _native.sdm_system_desc.lj_rule.offset == 176
    # 2 ints, 4 pointers, 5 doubles, 8 ints, 10 doubles
    assert C.sizeof(_native.sdm_centroid_restraint) == 8 + 32 + 40 + 32 + 80
    assert C.sizeof(_native.sdm_alignment_restraint) == 24 + 48


def _desc(s, **over):
    keep = []

    def d(a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        keep.append(a)
        return a.ctypes.data_as(_native._dp)

    def i(a):
        a = np.ascontiguousarray(a, dtype=np.int32)
        keep.append(a)
        return a.ctypes.data_as(_native._ip)

    desc = _native.sdm_system_desc()
    desc.n_atoms = s.n
    desc.charge, desc.sigma, desc.epsilon, desc.mass = d(s.charge), d(s.sigma), d(s.epsilon), d(s.mass)
    desc.n_exclusions = len(s.exclusions)
    desc.exclusions = i(np.asarray(s.exclusions).reshape(-1))
    desc.n_bonds = len(s.bonds)
    desc.bond_atoms, desc.bond_r0, desc.bond_k = i(np.asarray(s.bonds).reshape(-1)), d(s.bond_r0), d(s.bond_k)
    desc.n_angles = len(s.angles)
    desc.angle_atoms, desc.angle_theta0, desc.angle_k = (i(np.asarray(s.angles).reshape(-1)),
                                                         d(s.angle_t0), d(s.angle_k))
    desc.displacement = d(np.asarray(s.displacement).reshape(-1))
    for k in range(3):
        desc.box[k] = s.box[k]
    desc.periodic = 1
    desc.cutoff = s.cutoff
    desc.eps_rf = s.eps_rf
    for k, v in over.items():
        setattr(desc, k, v)
    return desc, keep


@pytest.mark.parametrize("field,value,code", [
    ("n_atoms", 0, _native.SDM_ERR_INVALID),
    ("cutoff", 2.0, _native.SDM_ERR_INVALID),
    ("cutoff", -1.0, _native.SDM_ERR_INVALID),
    ("eps_rf", 0.0, _native.SDM_ERR_INVALID),
])
def test_create_rejects_bad_system(field, value, code):
    lib = _native.load()
    s = systems.make_small()
    desc, keep = _desc(s, **{field: value})
    h = C.c_void_p()
    rc = lib.sdm_create(C.byref(desc), 1, C.byref(h))
    assert rc == code
    assert lib.sdm_last_error().decode() != ""
    assert not h.value


def test_create_rejects_bad_exclusion_and_replicas():
    lib = _native.load()
    s = systems.make_small()
    bad = np.array(s.exclusions, dtype=np.int32)
    bad[0, 0] = s.n + 5
    s2 = s.with_coordinates(s.positions)
    s2.exclusions = bad
    desc, keep = _desc(s2)
    h = C.c_void_p()
    assert lib.sdm_create(C.byref(desc), 1, C.byref(h)) == _native.SDM_ERR_INVALID
    desc, keep = _desc(s)
    assert lib.sdm_create(C.byref(desc), 0, C.byref(h)) == _native.SDM_ERR_INVALID


def test_too_many_exclusions_is_unsupported():
    lib = _native.load()
    s = systems.make_small()
    s2 = s.with_coordinates(s.positions)
    s2.exclusions = np.array([(0, j) for j in range(1, 34)], dtype=np.int32)
    desc, keep = _desc(s2)
    h = C.c_void_p()
    assert lib.sdm_create(C.byref(desc), 1, C.byref(h)) == _native.SDM_ERR_UNSUPPORTED
    assert "exclusions" in lib.sdm_last_error().decode()


def test_null_context_calls_fail_cleanly():
    lib = _native.load()
    assert lib.sdm_step(None, 1) == _native.SDM_ERR_INVALID
    assert lib.sdm_compute(None) == _native.SDM_ERR_INVALID
    assert lib.sdm_destroy(None) == _native.SDM_OK


def test_plugin_registration_symbols_are_inert():
    lib = _native.load()
    lib.registerPlatforms()
    lib.registerKernelFactories()
    lib.registerSDMHIPKernelFactories()
