"""GPU gradient of fdtdz() with respect to epsilon (adjoint kernels through
torch.autograd) against the FP64 discrete adjoint oracle, which is itself
pinned to torch.autograd through the FP64 stepper and to central differences.

Tolerance: relative L2 error <= 1e-3 of the gradient.
"""

import numpy as np
import pytest
import torch

from oracle import adjoint
from oracle import helpers as H
from pjz_amd import fdtdz_torch as F
from pjz_amd import workloads

pytestmark = pytest.mark.gpu
TOL = 1e-3


def rel_l2(a, b):
  return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def problem(axis, seed, X=20, Y=18, Z=24, tt=90, widths=(5, 4)):
  rng = np.random.default_rng(seed)
  eps = (1 + 3 * rng.random((3, X - 4, Y - 4, Z - 6))).astype(np.float32)
  sf = {0: rng.standard_normal((2, 1, Y, Z)), 1: rng.standard_normal((2, X, 1, Z)),
        2: rng.standard_normal((2, 2, X, Y, 1))}[axis].astype(np.float32)
  pos = {0: 6, 1: 8, 2: 11}[axis]
  t = np.arange(tt)
  wave = np.stack([np.sin(0.3 * t), np.cos(0.3 * t)], -1) * np.exp(-((t - 25) / 10.0) ** 2)[:, None]
  kw = dict(dt=0.5, source_field=sf, source_waveform=wave.astype(np.float32),
            source_position=pos,
            absorption_mask=H.absorption_mask(X, Y, 4, 5e-3).astype(np.float32),
            pml_kappa=(1 + 0.5 * rng.random((Z, 2))).astype(np.float32),
            pml_sigma=H.pml_sigma(widths, Z, 0.5, 3.0).astype(np.float32),
            pml_alpha=np.full((Z, 2), 0.05, np.float32), pml_widths=widths,
            output_steps=(tt - 30, tt, 6), use_reduced_precision=False, launch_params=None,
            offset=(2, 2, 3))
  nn = len(range(*kw["output_steps"]))
  g = rng.standard_normal((nn, 3) + eps.shape[1:]).astype(np.float32)
  return eps, kw, g


def gpu_grad(eps, kw, g, **launch):
  e = torch.tensor(eps, device="cuda", requires_grad=True)
  kw = dict(kw, launch_params=launch or None)
  out = F.fdtdz(e, **kw)
  (out * torch.as_tensor(g, device="cuda")).sum().backward()
  torch.cuda.synchronize()
  return e.grad.cpu().numpy(), out.detach().cpu().numpy()


@pytest.mark.parametrize("kernel", ["reference", "stream", "stream2"])
@pytest.mark.parametrize("axis", [0, 1, 2])
def test_gradient_matches_oracle(kernel, axis):
  print(f"case kernel={kernel} axis={axis}", flush=True)
  eps, kw, g = problem(axis, seed=axis)
  ref, ref_out = adjoint.grad_epsilon(eps, g_out=g, return_out=True, **kw)
  got, out = gpu_grad(eps, kw, g, kernel=kernel)
  assert rel_l2(out, ref_out) <= 1e-4
  err = rel_l2(got, ref)
  print(f"grad rel L2 {err:.2e}")
  assert err <= TOL


@pytest.mark.parametrize("tile_y", [4, 8])
@pytest.mark.parametrize("axis", [0, 2])
def test_two_step_reverse_kernel_tiles(axis, tile_y):
  """The two-step reverse kernel (default with the two-step forward) in both
  row tilings, z = 64 and 128 lanes layouts, against the oracle."""
  for Z in (24, 80):
    eps, kw, g = problem(axis, seed=40 + axis, Z=Z)
    ref = adjoint.grad_epsilon(eps, g_out=g, **kw)
    got, _ = gpu_grad(eps, kw, g, kernel="stream2", tile_y=tile_y, ckpt_every=30)
    assert rel_l2(got, ref) <= TOL, (Z, tile_y)


@pytest.mark.parametrize("kernel", ["stream", "stream2"])
@pytest.mark.parametrize("ckpt", [1, 7, 88, 1000])
def test_gradient_independent_of_checkpoint_spacing(ckpt, kernel):
  eps, kw, g = problem(2, seed=5)
  ref = adjoint.grad_epsilon(eps, g_out=g, **kw)
  got, _ = gpu_grad(eps, kw, g, ckpt_every=ckpt, kernel=kernel)
  assert rel_l2(got, ref) <= TOL


def test_gradient_kappa_inf():
  eps, kw, g = problem(1, seed=6)
  Z = kw["pml_sigma"].shape[0]
  kw.update(pml_kappa=np.full((Z, 2), np.inf, np.float32),
            pml_sigma=np.zeros((Z, 2), np.float32), pml_alpha=np.zeros((Z, 2), np.float32))
  ref = adjoint.grad_epsilon(eps, g_out=g, **kw)
  got, _ = gpu_grad(eps, kw, g)
  assert rel_l2(got, ref) <= TOL


def test_gradient_layered_64():
  """64x64x64 layered stack, 400 steps, loss on the last snapshots."""
  kw = workloads.layered_64(steps=400)
  kw["output_steps"] = (370, 400, 10)
  eps = kw.pop("epsilon")
  rng = np.random.default_rng(1)
  nn = len(range(*kw["output_steps"]))
  g = rng.standard_normal((nn, 3) + eps.shape[1:]).astype(np.float32)
  ref = adjoint.grad_epsilon(eps, g_out=g, **kw)
  got, _ = gpu_grad(eps, kw, g)
  err = rel_l2(got, ref)
  print(f"layered_64 grad rel L2 {err:.2e}")
  assert err <= TOL


def test_no_grad_path_allocates_no_checkpoints():
  eps, kw, g = problem(0, seed=7)
  with torch.no_grad():
    e = torch.tensor(eps, device="cuda", requires_grad=True)
    out = F.fdtdz(e, **kw)
  assert not out.requires_grad


def test_gradient_full_size_directional():
  """The gradient benchmark's domain (256x256x128) over 800 steps, with its
  overlap objective: the adjoint's directional derivative <dL/deps, v>
  against central differences of the float32 forward, Richardson-extrapolated
  from steps h and h/2 (the oracle cannot run this size in a test).  v raises
  the permittivity of the waveguide core between the source and the monitor
  (randomly modulated), so the objective's change is coherent and well above
  float32 rounding."""
  dev = torch.device("cuda", 0)
  kw, mode, omegas = workloads.grad_256(steps=800)
  eps0 = torch.as_tensor(kw.pop("epsilon"), device=dev)
  for k, val in list(kw.items()):
    if isinstance(val, np.ndarray):
      kw[k] = torch.as_tensor(val, device=dev)
  kw["launch_params"] = dict(ckpt_every=64)

  def loss(e):
    return workloads.overlap_objective(F.fdtdz(e, **kw), mode, omegas, kw["output_steps"])

  print("grad_256 800 steps: directional derivative", flush=True)
  e = eps0.clone().requires_grad_(True)
  loss(e).backward()
  gen = torch.Generator(device=dev).manual_seed(5)
  v = torch.zeros_like(eps0)
  core = (slice(None), slice(60, 200), slice(104, 152), slice(48, 64))
  v[core] = 0.5 + torch.rand(v[core].shape, generator=gen, device=dev)
  analytic = float((e.grad * v).sum())
  h = 1e-2
  with torch.no_grad():
    def central(step):
      return (float(loss(eps0 + step * v)) - float(loss(eps0 - step * v))) / (2 * step)
    fd1, fd2 = central(h), central(h / 2)
  fd = (4 * fd2 - fd1) / 3
  print(f"adjoint {analytic:.6e}, central differences {fd1:.6e} (h), {fd2:.6e} (h/2), "
        f"extrapolated {fd:.6e}", flush=True)
  assert abs(fd) > 0
  assert abs(analytic - fd) <= 5e-3 * abs(fd)


@pytest.mark.parametrize("every,inner", [(24, 8), (24, 7), (30, 29), (1000, 9)])
def test_two_level_checkpoints_match_one_level(every, inner):
  """Two levels of checkpoints (the backward recomputes each outer segment
  keeping inner states, then each inner sub-segment capturing E changes):
  the same gradient as one level, and the oracle's."""
  print(f"case every={every} inner={inner}", flush=True)
  eps, kw, g = problem(0, seed=9)
  ref = adjoint.grad_epsilon(eps, g_out=g, **kw)
  one, _ = gpu_grad(eps, kw, g, ckpt_every=every)
  two, _ = gpu_grad(eps, kw, g, ckpt_every=every, ckpt_inner=inner)
  assert rel_l2(two, ref) <= TOL
  # Reversed steps pair up within a recomputed segment only, so the inner
  # spacing changes which steps share a launch (and the rounding): 1e-5.
  assert rel_l2(two, one) <= 1e-5


def test_memory_budget_selects_two_levels():
  """A budget below the one-level workspace makes fdtdz() choose two levels
  of checkpoints by itself; the gradient still matches the oracle."""
  eps, kw, g = problem(1, seed=10)
  prob = F.make_problem(eps.shape, kw["dt"], kw["source_field"].shape, kw["source_position"],
                        kw["absorption_mask"].shape, kw["pml_kappa"].shape, kw["pml_widths"],
                        kw["output_steps"], kw["source_waveform"].shape[0],
                        kw["use_reduced_precision"], None, kw["offset"])
  one = F.default_ckpt_every(F.steps_of(prob))
  budget = F._lib.workspace_bytes(F._with_ckpt(prob, one, 0)) - 1
  every, inner = F.plan_checkpoints(prob, budget)
  assert inner > 0
  ref = adjoint.grad_epsilon(eps, g_out=g, **kw)
  got, _ = gpu_grad(eps, kw, g, memory_budget=budget)
  assert rel_l2(got, ref) <= TOL
