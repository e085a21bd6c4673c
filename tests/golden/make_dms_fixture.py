"""Writes tests/golden/temoa_g1_g4.npz: the reference's example system
example/temoa-g1-g4.dms (20,446 atoms) as read by openmm_sdm_plugin_amd.dms, for the
GPU tests (the GPU machines do not have the reference).  Run from the repository root:

    python tests/golden/make_dms_fixture.py /root/reference/example/temoa-g1-g4.dms
"""
import dataclasses
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from openmm_sdm_plugin_amd.dms import DesmondDMSFile  # noqa: E402


def main(path):
    d = DesmondDMSFile(path)
    ff = d.createSystem(nonbondedMethod="PME", nonbondedCutoff=1.0)
    out = {k: v for k, v in dataclasses.asdict(ff).items() if isinstance(v, np.ndarray)}
    out["positions"] = d.getPositions()
    out["velocities"] = d.getVelocities()
    out["resid"] = np.array([a.resid for a in d.topology.atoms()], dtype=np.int32)
    out["scalars"] = np.array([ff.cutoff, float(ff.opls), float(ff.remove_cm_motion)])
    dest = os.path.join(ROOT, "tests", "golden", "temoa_g1_g4.npz")
    np.savez_compressed(dest, **out)
    print(dest, os.path.getsize(dest))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/example/temoa-g1-g4.dms")
