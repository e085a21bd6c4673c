"""GPU checks of the hand-written adjoint (config 4: 3,000-atom frames): u per frame and
du/d(q, sigma, eps) against the FP64 oracle, and the gradients of mean W against FP64
central finite differences of the oracle (h = 1e-4 relative, tolerance 1e-4 relative)."""
import math

import numpy as np
import pytest
import torch

from openmm_sdm_plugin_amd import adjoint, systems
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def frames_tensor(frames):
    x = np.stack([f.positions for f in frames])
    return torch.tensor(x, dtype=torch.float64, device="cuda").contiguous()


@pytest.fixture(scope="module")
def c4(hip_device):
    frames = systems.make_config4(4)
    s = frames[0]
    p = systems.params_config4()
    model = adjoint.AdjointModel(s, s.ligand, p)
    return frames, s, p, model


def test_u_and_grad_match_oracle(c4):
    frames, s, p, model = c4
    X = frames_tensor(frames)
    sel = torch.tensor(s.ligand, dtype=torch.long)
    q = torch.tensor(s.charge[s.ligand], dtype=torch.float64, device="cuda", requires_grad=True)
    sg = torch.tensor(s.sigma[s.ligand], dtype=torch.float64, device="cuda", requires_grad=True)
    ep = torch.tensor(s.epsilon[s.ligand], dtype=torch.float64, device="cuda", requires_grad=True)
    u = adjoint.PerturbationEnergy.apply(q, sg, ep, X, model)
    wts = torch.linspace(0.5, 2.0, len(frames), dtype=torch.float64, device="cuda")
    (u * wts).sum().backward()
    topo = O.Topology(s)
    gq = np.zeros(len(sel))
    gs = np.zeros(len(sel))
    ge = np.zeros(len(sel))
    for k, f in enumerate(frames):
        _, uref, *_ = O.sdm_direct(topo, f.positions)
        assert math.isclose(float(u[k]), uref, rel_tol=1e-10, abs_tol=1e-8)
        dq, ds, de = O.du_dparams(topo, f.positions, s.ligand)
        w = float(wts[k])
        gq += w * dq
        gs += w * ds
        ge += w * de
    for got, ref in ((q.grad, gq), (sg.grad, gs), (ep.grad, ge)):
        got = got.cpu().numpy()
        assert np.allclose(got, ref, rtol=1e-9, atol=1e-9 * np.max(np.abs(ref)))


def test_geometric_rule_u_and_grad_match_oracle(hip_device):
    """The OPLS (geometric sigma) rule: sigma_ij = sqrt(s_i s_j), so d sigma_ij / d s_i =
    s_j / (2 sigma_ij); u and its parameter gradient against the oracle (whose derivative
    is pinned by central differences, tests/test_oracle_pins.py)."""
    frames = [systems.make_small(frame=k) for k in range(3)]
    for f in frames:
        f.lj_rule = 1
    s = frames[0]
    print(f"case geometric adjoint small n={s.n} frames={len(frames)}", flush=True)
    p = systems.params_config1()
    model = adjoint.AdjointModel(s, s.ligand, p)
    X = frames_tensor(frames)
    q = torch.tensor(s.charge[s.ligand], dtype=torch.float64, device="cuda", requires_grad=True)
    sg = torch.tensor(s.sigma[s.ligand], dtype=torch.float64, device="cuda", requires_grad=True)
    ep = torch.tensor(s.epsilon[s.ligand], dtype=torch.float64, device="cuda", requires_grad=True)
    u = adjoint.PerturbationEnergy.apply(q, sg, ep, X, model)
    u.sum().backward()
    topo = O.Topology(s)
    gs = np.zeros(len(s.ligand))
    for k, f in enumerate(frames):
        _, uref, *_ = O.sdm_direct(topo, f.positions)
        assert math.isclose(float(u[k]), uref, rel_tol=1e-10, abs_tol=1e-8)
        gs += O.du_dparams(topo, f.positions, s.ligand)[1]
    got = sg.grad.cpu().numpy()
    assert np.allclose(got, gs, rtol=1e-9, atol=1e-9 * np.max(np.abs(gs)))
    # the arithmetic rule's gradient differs
    s0 = s.with_coordinates(s.positions)
    s0.lj_rule = 0
    g0 = sum(O.du_dparams(O.Topology(s0), f.positions, s.ligand)[1] for f in frames)
    assert np.max(np.abs(g0 - gs)) > 1e-3 * np.max(np.abs(gs))


def _oracle_mean_W(s, frames, p, q=None, sig=None, eps=None, alpha=None, u0=None, w0=None):
    s2 = s.with_coordinates(s.positions)
    s2.charge = s.charge.copy() if q is None else q
    s2.sigma = s.sigma.copy() if sig is None else sig
    s2.epsilon = s.epsilon.copy() if eps is None else eps
    topo = O.Topology(s2)
    pp = systems.SDMParams(**{k: getattr(p, k) for k in p.__dataclass_fields__})
    if alpha is not None:
        pp.alpha = alpha
    if u0 is not None:
        pp.u0 = u0
    if w0 is not None:
        pp.w0coeff = w0
    Ws = []
    for f in frames:
        _, u, *_ = O.sdm_direct(topo, f.positions)
        Ws.append(O.execute_scalars(O.make_params(pp), 0.0, u, 0.0, 0.0).ebias)
    return float(np.mean(Ws))


def test_mean_W_gradient_vs_central_fd(c4):
    frames, s, p, model = c4
    X = frames_tensor(frames)
    lig = s.ligand
    q = torch.tensor(s.charge[lig], dtype=torch.float64, device="cuda", requires_grad=True)
    sg = torch.tensor(s.sigma[lig], dtype=torch.float64, device="cuda", requires_grad=True)
    ep = torch.tensor(s.epsilon[lig], dtype=torch.float64, device="cuda", requires_grad=True)
    al = torch.tensor(p.alpha, dtype=torch.float64, device="cuda", requires_grad=True)
    u0 = torch.tensor(p.u0, dtype=torch.float64, device="cuda", requires_grad=True)
    w0 = torch.tensor(p.w0coeff, dtype=torch.float64, device="cuda", requires_grad=True)
    W, U = model.mean_W_and_u(q, sg, ep, al, u0, w0, X)
    W.backward()
    checks = []
    for k in (0, 7, 19):
        a = int(lig[k])
        for name, arr, grad in (("charge", s.charge, q.grad), ("sigma", s.sigma, sg.grad),
                                ("epsilon", s.epsilon, ep.grad)):
            h = 1e-4 * max(abs(arr[a]), 0.05)
            vals = []
            for sgn in (1, -1):
                mod = arr.copy()
                mod[a] += sgn * h
                vals.append(_oracle_mean_W(s, frames, p, **{{"charge": "q", "sigma": "sig",
                                                            "epsilon": "eps"}[name]: mod}))
            checks.append((name, k, (vals[0] - vals[1]) / (2 * h), float(grad[k])))
    for name, t in (("alpha", al), ("u0", u0), ("w0", w0)):
        base = getattr(p, {"alpha": "alpha", "u0": "u0", "w0": "w0coeff"}[name])
        h = 1e-4 * max(abs(base), 1.0)
        vp = _oracle_mean_W(s, frames, p, **{name: base + h})
        vm = _oracle_mean_W(s, frames, p, **{name: base - h})
        checks.append((name, -1, (vp - vm) / (2 * h), float(t.grad)))
    scale = max(abs(c[2]) for c in checks)
    for name, k, fd, got in checks:
        assert abs(fd - got) <= 1e-4 * max(abs(fd), 1e-3 * scale), (name, k, fd, got)


def test_full_config4_runs_and_matches_oracle_on_end_frames(hip_device):
    frames = systems.make_config4(64)
    s = frames[0]
    model = adjoint.AdjointModel(s, s.ligand, systems.params_config4())
    X = frames_tensor(frames)
    lig = s.ligand
    q = torch.tensor(s.charge[lig], dtype=torch.float64, device="cuda", requires_grad=True)
    u = adjoint.PerturbationEnergy.apply(q, torch.tensor(s.sigma[lig], device="cuda"),
                                         torch.tensor(s.epsilon[lig], device="cuda"), X, model)
    gw = torch.zeros(64, dtype=torch.float64, device="cuda")
    gw[0], gw[63] = 1.0, 1.0
    u.backward(gw)
    topo = O.Topology(s)
    ref = np.zeros(len(lig))
    for k in (0, 63):
        _, uref, *_ = O.sdm_direct(topo, frames[k].positions)
        assert math.isclose(float(u[k]), uref, rel_tol=1e-10, abs_tol=1e-8)
        ref += O.du_dparams(topo, frames[k].positions, lig)[0]
    assert np.allclose(q.grad.cpu().numpy(), ref, rtol=1e-9, atol=1e-9 * np.max(np.abs(ref)))
