"""CPU checks of the differentiable soft-core/bias chain used by the adjoint: values
against the oracle's SoftCoreF/execute() restatement, derivatives by torch.gradcheck."""
import math

import pytest
import torch

from openmm_sdm_plugin_amd import adjoint, systems
from oracle import oracle as O


@pytest.mark.parametrize("method", [0, 1, 2])
def test_softcore_matches_oracle(method):
    for u in (-20.0, 50.0, 120.0, 400.0, 1e6):
        up = adjoint.softcore(torch.tensor(u, dtype=torch.float64), method, 200.0, 0.0625, 100.0)
        ref, _ = O.softcore(method, u, 200.0, 0.0625, 100.0)
        assert math.isclose(float(up), ref, rel_tol=1e-13, abs_tol=1e-12)


@pytest.mark.parametrize("bias_method", [0, 1, 2])
def test_bias_matches_oracle(bias_method):
    p = systems.params_config4()
    p.bias_method = bias_method
    p.softcore_method = 0
    p.gamma, p.wbcoeff = 0.02, 0.3
    for u in (-40.0, 30.0, 110.0, 180.0):
        W = adjoint.bias(torch.tensor(u, dtype=torch.float64), p)
        sc = O.execute_scalars(O.make_params(p), 0.0, u, 0.0, 0.0)
        assert math.isclose(float(W), sc.ebias, rel_tol=1e-13, abs_tol=1e-12)


def test_chain_gradients_gradcheck():
    p = systems.params_config4()

    def f(u, alpha, u0, w0):
        up = adjoint.softcore(u, p.softcore_method, p.umax, p.acore, p.ubcore)
        return adjoint.bias(up, p, alpha, u0, w0)

    u = torch.tensor([50.0, 105.0, 150.0, 900.0], dtype=torch.float64, requires_grad=True)
    a = torch.tensor(0.1, dtype=torch.float64, requires_grad=True)
    u0 = torch.tensor(110.0, dtype=torch.float64, requires_grad=True)
    w0 = torch.tensor(1.0, dtype=torch.float64, requires_grad=True)
    assert torch.autograd.gradcheck(f, (u, a, u0, w0), eps=1e-6, atol=1e-7, rtol=1e-5)


def test_dW_du_is_oracle_sp():
    p = systems.params_config4()
    for uval in (60.0, 130.0, 400.0):
        u = torch.tensor(uval, dtype=torch.float64, requires_grad=True)
        up = adjoint.softcore(u, p.softcore_method, p.umax, p.acore, p.ubcore)
        W = adjoint.bias(up, p)
        W.backward()
        sc = O.execute_scalars(O.make_params(p), 0.0, uval, 0.0, 0.0)
        assert math.isclose(float(u.grad), sc.sp, rel_tol=1e-12)
