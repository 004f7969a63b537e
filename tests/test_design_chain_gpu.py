"""The inverse-design chain on the GPU, end to end: latent -> density() ->
epsilon() -> fdtdz() -> objective, differentiated back to the latent through
torch autograd (density filter and projection, sub-pixel render) and the
adjoint kernels (fdtdz, checkpointed), against central finite differences
of the same float32 chain along a random direction (Richardson-extrapolated
from steps h and h/2).

density() binarises filtered pixels away from the design boundary, which is
piecewise: the latent is a block pattern whose filtered values keep a margin
from the threshold, and the step h is set so that no pixel crosses it (the
boundary set is asserted unchanged), so the difference and the gradient see
the same branch.
"""

import numpy as np
import pytest
import torch

from oracle import helpers as H
from pjz_amd import _density as D
from pjz_amd import density, epsilon, fdtdz

pytestmark = pytest.mark.gpu


def test_latent_to_objective_gradient():
  dev = torch.device("cuda", 0)
  X, Y, Z, tt, m, radius, eta = 24, 20, 32, 120, 1, 3.0, 0.5
  s = -torch.ones((2 * m * X, 2 * m * Y))
  s[10:34, 8:30] = 1.0
  s[18:26, 14:22] = -1.0
  u0 = (0.5 + 0.8 * s).to(dev)
  rng = np.random.default_rng(7)
  v = torch.tensor(rng.standard_normal(u0.shape), dtype=torch.float32, device=dev)
  margin = float((D.blur(u0, radius) - eta).abs().min())
  h = 0.5 * margin / float(D.blur(v, radius).abs().max())
  sets = [D.boundary_cells(D.blur(u0 + t * h * v, radius) - eta) for t in (0, 1, -1)]
  assert margin > 0.05 and all(torch.equal(sets[0], b) for b in sets[1:])

  t = np.arange(tt)
  wave = np.stack([np.sin(0.35 * t), np.cos(0.35 * t)], -1) * np.exp(-((t - 35) / 12.0) ** 2)[:, None]
  src = rng.standard_normal((2, 1, Y, Z))
  common = dict(dt=0.5, source_field=torch.tensor(src, dtype=torch.float32, device=dev),
                source_waveform=torch.tensor(wave, dtype=torch.float32, device=dev),
                source_position=6,
                absorption_mask=torch.tensor(H.absorption_mask(X, Y, 4, 5e-3), dtype=torch.float32,
                                             device=dev),
                pml_kappa=torch.ones((Z, 2), device=dev),
                pml_sigma=torch.tensor(H.pml_sigma((4, 4), Z, 0.5, 3.0), dtype=torch.float32,
                                       device=dev),
                pml_alpha=torch.full((Z, 2), 0.05, device=dev), pml_widths=(4, 4),
                output_steps=(tt - 20, tt, 5), use_reduced_precision=False,
                launch_params=dict(ckpt_every=16), offset=(0, 0, 0))
  gen = torch.Generator(device=dev).manual_seed(3)
  g = torch.randn((4, 3, X, Y, Z), generator=gen, device=dev)
  pos = torch.tensor([10.0, 18.0], device=dev)  # cladding | core | cladding in z

  def objective(u):
    d, _ = density(u, radius, 0.6, eta=eta)
    core = 2.085 + (12.25 - 2.085) * d
    clad = torch.full_like(core, 2.085)
    eps = epsilon(torch.stack([clad, core, clad]), pos, m, Z)
    return (fdtdz(epsilon=eps, **common) * g).sum()

  u = u0.clone().requires_grad_(True)
  objective(u).backward()
  analytic = float((u.grad * v).sum())
  with torch.no_grad():
    def central(step):
      return (float(objective(u0 + step * v)) - float(objective(u0 - step * v))) / (2 * step)
    fd1, fd2 = central(h), central(h / 2)
  # Richardson: the O(h^2) term of the central difference cancels.
  fd = (4 * fd2 - fd1) / 3
  print(f"h {h:.3e}: adjoint {analytic:.6e}, central differences {fd1:.6e} (h), "
        f"{fd2:.6e} (h/2), extrapolated {fd:.6e}", flush=True)
  assert abs(fd) > 0
  assert abs(analytic - fd) <= 5e-3 * abs(fd)
