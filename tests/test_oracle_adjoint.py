"""Pins the numpy discrete adjoint (oracle/adjoint.py) against torch.autograd
through the FP64 stepper, and checks a dot-product identity."""

import numpy as np
import pytest
import torch

from oracle import adjoint, fdtd_spec as S, helpers as H


def problem(axis, seed, X=8, Y=7, Z=10, tt=30):
  rng = np.random.default_rng(seed)
  eps = 1 + 3 * rng.random((3, X - 2, Y - 2, Z - 2))
  sf = {0: rng.standard_normal((2, 1, Y, Z)), 1: rng.standard_normal((2, X, 1, Z)),
        2: rng.standard_normal((2, 2, X, Y, 1))}[axis]
  pos = {0: 4, 1: 2, 2: 5}[axis]
  kw = dict(dt=0.5, source_field=sf, source_waveform=rng.standard_normal((tt, 2)),
            source_position=pos, absorption_mask=0.2 * rng.random((3, X, Y)),
            pml_kappa=1 + rng.random((Z, 2)), pml_sigma=H.pml_sigma((3, 2), Z, 0.5, 3.0),
            pml_alpha=0.1 * rng.random((Z, 2)), pml_widths=(3, 2),
            output_steps=(tt - 12, tt, 4), offset=(1, 1, 1))
  return eps, kw, rng


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_adjoint_matches_autograd(axis):
  eps, kw, rng = problem(axis, seed=axis)
  nn = len(range(*kw["output_steps"]))
  g = rng.standard_normal((nn, 3) + eps.shape[1:])
  t = torch.tensor(eps, requires_grad=True)
  out = S.fdtdz(t, xp=torch, **kw)
  (out * torch.as_tensor(g)).sum().backward()
  ref = t.grad.numpy()
  got = adjoint.grad_epsilon(eps, g_out=g, **kw)
  np.testing.assert_allclose(got, ref, rtol=0, atol=1e-10 * np.abs(ref).max())


def test_adjoint_kappa_inf():
  eps, kw, rng = problem(1, seed=9)
  Z = kw["pml_sigma"].shape[0]
  kw.update(pml_kappa=np.full((Z, 2), np.inf), pml_sigma=np.zeros((Z, 2)),
            pml_alpha=np.zeros((Z, 2)))
  nn = len(range(*kw["output_steps"]))
  g = rng.standard_normal((nn, 3) + eps.shape[1:])
  t = torch.tensor(eps, requires_grad=True)
  (S.fdtdz(t, xp=torch, **kw) * torch.as_tensor(g)).sum().backward()
  got = adjoint.grad_epsilon(eps, g_out=g, **kw)
  np.testing.assert_allclose(got, t.grad.numpy(), rtol=0, atol=1e-10 * np.abs(got).max())


def test_central_difference():
  eps, kw, rng = problem(0, seed=5)
  nn = len(range(*kw["output_steps"]))
  g = rng.standard_normal((nn, 3) + eps.shape[1:])
  grad = adjoint.grad_epsilon(eps, g_out=g, **kw)
  L = lambda e: float((S.fdtdz(e, **kw) * g).sum())
  for idx in [(0, 1, 1, 2), (1, 3, 2, 4), (2, 5, 4, 7)]:
    h = 1e-5
    ep, em = eps.copy(), eps.copy()
    ep[idx] += h
    em[idx] -= h
    fd = (L(ep) - L(em)) / (2 * h)
    assert fd == pytest.approx(grad[idx], rel=1e-6, abs=1e-9 * np.abs(grad).max())
