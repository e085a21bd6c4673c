"""The four-site test system shared by the virtual-site GPU tests: the small water box with
every water's charge on a massless site M = (1 - 2a) O + a H1 + a H2 (a = 0.128), and two
guest sites (a two-particle average and an out-of-plane site) displaced with the guest."""
import dataclasses
import types

import numpy as np

from openmm_sdm_plugin_amd import systems
from oracle import vsites as OV

A_M = 0.128


def four_site(frame=1):
    s = systems.make_small(frame=frame)
    n0 = s.n
    lig = [int(a) for a in s.ligand]
    ligset = set(lig)
    waters = [(o, h1, h2) for h1, o, h2 in np.asarray(s.angles).reshape(-1, 3).tolist()
              if o not in ligset]
    q = np.array(s.charge, dtype=np.float64)
    new_q, kind, atoms, w, excl = [], [], [], [], []
    for k, (o, h1, h2) in enumerate(waters):
        m = n0 + k
        new_q.append(q[o])
        q[o] = 0.0
        kind.append(OV.AVERAGE3)
        atoms.append((m, o, h1, h2))
        w.append((1 - 2 * A_M, A_M, A_M))
        excl += [(m, o), (m, h1), (m, h2)]
    # guest sites: charge taken from their first parent, so the guest stays neutral
    g2, g3 = n0 + len(waters), n0 + len(waters) + 1
    kind += [OV.AVERAGE2, OV.OUT_OF_PLANE]
    atoms += [(g2, lig[0], lig[1], lig[0]), (g3, lig[1], lig[2], lig[3])]
    w += [(0.7, 0.3, 0.0), (0.2, 0.3, 1.5)]
    new_q += [0.12, -0.08]
    q[lig[0]] -= 0.12
    q[lig[1]] += 0.08
    for g in (g2, g3):
        excl += [(g, a) for a in lig]
    excl.append((g2, g3))
    nv = len(kind)
    vs = types.SimpleNamespace(kind=np.array(kind, dtype=np.int32),
                               atoms=np.array(atoms, dtype=np.int32),
                               weights=np.array(w, dtype=np.float64))
    disp = np.asarray(s.displacement, dtype=np.float64)
    d_new = np.zeros((nv, 3))
    d_new[-2:] = disp[lig[0]]
    x = np.concatenate([np.asarray(s.positions), np.zeros((nv, 3))])
    x = OV.positions(x, vs)
    s2 = dataclasses.replace(
        s, charge=np.concatenate([q, new_q]),
        sigma=np.concatenate([s.sigma, np.full(nv, 0.1)]),
        epsilon=np.concatenate([s.epsilon, np.zeros(nv)]),
        mass=np.concatenate([s.mass, np.zeros(nv)]),
        exclusions=np.concatenate([np.asarray(s.exclusions).reshape(-1, 2),
                                   np.array(excl, dtype=np.int32)]).astype(np.int32),
        displacement=np.concatenate([disp, d_new]), positions=x,
        velocities=np.concatenate([np.asarray(s.velocities), np.zeros((nv, 3))]),
        ligand=np.array(lig + [g2, g3], dtype=np.int32), virtual_sites=vs, name="small-4site")
    return s2
