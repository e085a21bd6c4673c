"""Error scale of FP32 pair arithmetic, shared by the GPU parity tests."""
import numpy as np


def pair_abs_sum(s, x):
    """Sum of |E_LJ| + |E_Coulomb-RF| over the in-cutoff, non-excluded pairs (FP64, brute
    force; small systems only).  FP32 pair arithmetic on exact FP32 positions has an
    error of a few units of 2^-24 times this sum, whatever the (possibly strongly
    cancelling) total: the 186-atom test system has U0 ~ 30 kJ/mol from a sum of
    |pair energies| ~ 4.5e4, where an FP32 emulation of the kernel's arithmetic alone
    is off by 3.5e-4, above 1e-5 relative."""
    L = np.asarray(s.box)
    d = x[:, None, :] - x[None, :, :]
    d -= L * np.round(d / L)
    i, j = np.triu_indices(s.n, 1)
    r = np.sqrt((d[i, j] ** 2).sum(-1))
    ex = set(map(tuple, np.sort(np.asarray(s.exclusions).reshape(-1, 2), axis=1).tolist()))
    keep = (r < s.cutoff) & np.array([(a, b) not in ex for a, b in zip(i.tolist(), j.tolist())])
    i, j, r = i[keep], j[keep], r[keep]
    sig = 0.5 * (s.sigma[i] + s.sigma[j])
    eps = np.sqrt(s.epsilon[i] * s.epsilon[j])
    krf = (s.eps_rf - 1) / ((2 * s.eps_rf + 1) * s.cutoff ** 3)
    crf = 3 * s.eps_rf / ((2 * s.eps_rf + 1) * s.cutoff)
    sr6 = (sig / r) ** 6
    lj = 4 * eps * sr6 * (sr6 - 1)
    c = 138.935456 * s.charge[i] * s.charge[j] * (1 / r + krf * r * r - crf)
    return float(np.sum(np.abs(lj) + np.abs(c)))
