"""The allocation-free adjoint entry point (sdm_du_dparams_frames_noalloc), called through
the C ABI, gives the results of sdm_du_dparams_frames bit for bit: two calls in a row,
the second with fewer frames and fewer selected atoms than the first (the static
staging array then holds the first call's tail, which must not be read), and in a
context on which the allocating entry point has never run.  A selection larger than
SDM_NOALLOC_MAX_SEL is refused.  The first frame's u and gradient are also checked
against the FP64 oracle."""
import math

import numpy as np
import pytest
import torch

from openmm_sdm_plugin_amd import _native, systems
from oracle import oracle as O

pytestmark = pytest.mark.gpu

MAX_SEL = 16384  # SDM_NOALLOC_MAX_SEL in include/sdm_hip.h


def _frames(n):
    frames = systems.make_config4(n)
    x = np.stack([f.positions for f in frames])
    return frames, torch.tensor(x, dtype=torch.float64, device="cuda").contiguous()


def _run(ctx, fn, X, nframes, sel):
    grad = torch.full((nframes, len(sel), 3), float("nan"), dtype=torch.float64, device="cuda")
    u = torch.full((nframes,), float("nan"), dtype=torch.float64, device="cuda")
    fn(X.data_ptr(), nframes, sel, grad.data_ptr(), u.data_ptr())
    torch.cuda.synchronize()
    return grad.cpu().numpy(), u.cpu().numpy()


def test_noalloc_adjoint_matches_allocating_entry_point(hip_device):
    frames, X = _frames(4)
    s = frames[0]
    print(f"case noalloc adjoint n={s.n} frames=4,2", flush=True)
    ctx = _native.NativeContext(s, 1)
    big = np.concatenate([s.ligand, np.arange(0, 40)]).astype(np.int32)  # 70 atoms
    small = np.asarray(s.ligand[5:17][::-1], dtype=np.int32)            # 12 atoms
    # two calls in a row of the allocation-free entry point, the second smaller
    g1, u1 = _run(ctx, ctx.du_dparams_frames_noalloc, X, 4, big)
    g2, u2 = _run(ctx, ctx.du_dparams_frames_noalloc, X, 2, small)
    # the allocating entry point on the same inputs
    r1, v1 = _run(ctx, ctx.du_dparams_frames, X, 4, big)
    r2, v2 = _run(ctx, ctx.du_dparams_frames, X, 2, small)
    assert np.array_equal(g1, r1) and np.array_equal(u1, v1)
    assert np.array_equal(g2, r2) and np.array_equal(u2, v2)
    assert np.all(np.isfinite(g1)) and np.all(np.isfinite(g2))
    # and the oracle, on the first frame
    topo = O.Topology(s)
    _, uref, *_ = O.sdm_direct(topo, s.positions)
    assert math.isclose(float(u1[0]), uref, rel_tol=1e-10, abs_tol=1e-8)
    dq, ds, de = O.du_dparams(topo, s.positions, small)
    ref = np.stack([dq, ds, de], 1)
    assert np.allclose(g2[0], ref, rtol=1e-9, atol=1e-9 * np.max(np.abs(ref)))


def test_noalloc_adjoint_refuses_oversized_selection(hip_device):
    frames, X = _frames(1)
    ctx = _native.NativeContext(frames[0], 1)
    sel = np.zeros(MAX_SEL + 1, dtype=np.int32)
    grad = torch.zeros((1, 1), dtype=torch.float64, device="cuda")
    with pytest.raises(_native.SDMError) as err:
        ctx.du_dparams_frames_noalloc(X.data_ptr(), 1, sel, grad.data_ptr(), 0)
    assert err.value.code == _native.SDM_ERR_CAPACITY
    assert "SDM_NOALLOC_MAX_SEL" in str(err.value)
