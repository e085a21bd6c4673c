"""Writes the reference's example systems as read by openmm_sdm_plugin_amd.dms, for the
GPU tests (the GPU machines do not have the reference):

  tests/golden/temoa_g1_g4.npz  example/temoa-g1-g4.dms (20,446 atoms, periodic), read
                                with nonbondedMethod PME as example/test_explicit.py does
  tests/golden/oa_g6_g3.npz     example/oa-g6-g3-align-restr_0_displaced.dms (230 atoms,
                                vacuum), read with CutoffNonPeriodic and a 15 nm cutoff as
                                example/test.py:63 does
  tests/golden/oa_g6_g3_opls.npz the same file read with createSystem(OPLS=True): its
                                nonbonded_info declares the geometric rule, which the
                                reference's reader applies only when asked this way

Run from the repository root:

    python tests/golden/make_dms_fixture.py /root/reference/example
"""
import dataclasses
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from openmm_sdm_plugin_amd.dms import DesmondDMSFile  # noqa: E402

CASES = (("temoa-g1-g4.dms", "temoa_g1_g4.npz", "PME", 1.0, False),
         ("oa-g6-g3-align-restr_0_displaced.dms", "oa_g6_g3.npz", "CutoffNonPeriodic", 15.0,
          False),
         ("oa-g6-g3-align-restr_0_displaced.dms", "oa_g6_g3_opls.npz", "CutoffNonPeriodic",
          15.0, True))


def write(path, dest_name, method, cutoff, opls):
    d = DesmondDMSFile(path)
    ff = d.createSystem(nonbondedMethod=method, nonbondedCutoff=cutoff, OPLS=opls)
    out = {k: v for k, v in dataclasses.asdict(ff).items() if isinstance(v, np.ndarray)}
    out["positions"] = d.getPositions()
    out["velocities"] = d.getVelocities()
    out["resid"] = np.array([a.resid for a in d.topology.atoms()], dtype=np.int32)
    out["scalars"] = np.array([ff.cutoff, float(ff.opls), float(ff.remove_cm_motion)])
    dest = os.path.join(ROOT, "tests", "golden", dest_name)
    np.savez_compressed(dest, **out)
    print(dest, os.path.getsize(dest))
    d.close()


def main(folder, only=None):
    for name, dest, method, cutoff, opls in CASES:
        if only is None or dest in only:
            write(os.path.join(folder, name), dest, method, cutoff, opls)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/example",
         sys.argv[2:] or None)
