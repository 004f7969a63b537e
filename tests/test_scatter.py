"""scatter(): amplitude solve on CPU, S-values and reciprocity gradient on GPU."""

import numpy as np
import pytest
import torch

from oracle import scatter_ref
from pjz_amd import _field as Fd
from pjz_amd import _scatter as Sc

P = Fd.SimParams(omega_range=(2 * np.pi / 21, 2 * np.pi / 19), tt=600, absorption_padding=6,
                 absorption_coeff=4e-3, pml_widths=(6, 6), pml_alpha_coeff=0.05,
                 use_reduced_precision=False)


def test_amplitudes_pure_waves():
  # time factor exp(i w t) (output phases [cos; -sin]): a forward wave is
  # exp(-i beta x) -> amplitudes (1, 0), a backward one exp(i beta x) -> (0, 1)
  beta = np.array([0.3, 0.7])
  for direction, expect in ((-1, (1, 0)), (1, (0, 1))):
    x = np.array([1.0, 2.0])
    vals = np.exp(direction * 1j * beta[:, None] * x[None, :])
    ref = scatter_ref.amplitudes(beta, vals, x)
    np.testing.assert_allclose(ref, np.tile(expect, (2, 1)), atol=1e-12)
    got = Sc.amplitudes(torch.as_tensor(beta, dtype=torch.complex128), torch.as_tensor(vals),
                        torch.as_tensor(x))
    np.testing.assert_allclose(got.numpy(), ref, atol=1e-12)


def test_overlap_matches_oracle_on_cpu():
  rng = np.random.default_rng(0)
  out = rng.standard_normal((2, 3, 10, 8, 6)) + 1j * rng.standard_normal((2, 3, 10, 8, 6))
  mode = rng.standard_normal((2, 2, 1, 8, 6)) + 1j * rng.standard_normal((2, 2, 1, 8, 6))
  beta = np.array([0.4, 0.5])
  for fwd in (None, True, False):
    ref = scatter_ref.overlap(mode, beta, 4, fwd, out)
    got = Sc.overlap(torch.as_tensor(mode), torch.as_tensor(beta), 4, fwd, torch.as_tensor(out))
    np.testing.assert_allclose(got.numpy(), ref, rtol=1e-10, atol=1e-12)


def ports(seed=0, xx=16, yy=12, zz=16):
  rng = np.random.default_rng(seed)
  eps = (1 + 3 * rng.random((3, xx, yy, zz))).astype(np.float32)
  ww = 2
  m1 = (rng.standard_normal((ww, 2, 1, yy, zz))).astype(np.float32)
  m2 = (rng.standard_normal((ww, 2, 1, yy, zz))).astype(np.float32)
  betas = [np.array([0.5, 0.55]), np.array([0.5, 0.55])]
  return eps, (m1, m2), betas, (4, 11), (True, False)


@pytest.mark.gpu
def test_scatter_and_gradient_match_oracle(gpu):
  eps, modes, betas, pos, fwd = ports()
  omega = np.array([2 * np.pi / 20.5, 2 * np.pi / 19.5])
  ref_s, ref_grads = scatter_ref.scatter(eps, omega, modes, betas, pos, fwd, P)
  e = torch.tensor(eps, device=gpu, requires_grad=True)
  svals = Sc.scatter(e, omega, [torch.as_tensor(m, device=gpu) for m in modes],
                     [torch.as_tensor(b) for b in betas], pos, fwd, P)
  got = np.array([[s.detach().cpu().numpy() for s in row] for row in svals])
  ref = np.array(ref_s)
  err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
  print(f"svals rel err {err:.2e}")
  assert err <= 1e-3
  rng = np.random.default_rng(1)
  g = rng.standard_normal(ref.shape) + 1j * rng.standard_normal(ref.shape)
  loss = sum(torch.sum(torch.real(svals[i][j] * torch.as_tensor(np.conj(g[i, j]), device=gpu)))
             for i in range(2) for j in range(2))
  loss.backward()
  # The custom backward returns sum Re(g_ij * f_i f_j / a_i) for the cotangent
  # torch hands it; for this loss that cotangent is g itself.
  ref_grad = scatter_ref.scatter_vjp(ref_grads, g)
  got_grad = e.grad.cpu().numpy()
  gerr = np.linalg.norm(got_grad - ref_grad) / np.linalg.norm(ref_grad)
  print(f"scatter grad rel err {gerr:.2e}")
  assert gerr <= 1e-3
