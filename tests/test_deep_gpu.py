"""Three and four steps per launch ("stream3", "stream4") against the FP64
oracle (relative L2 <= 1e-4) and against the two-step kernel on the same
inputs (<= 1e-6: the same per-cell arithmetic, another schedule).

Covered: every source axis at heights up to 64 (one cell per lane, the
only heights these kernels take), both depths; odd shapes; x chunks shorter
than the depth (ghost planes that reach the domain start); step counts that
leave a remainder for the two-step / one-step tail; snapshots on every level
of a launch; z as batch; and the gradient, whose segment recomputes run the
deep kernel with captures.  The autotuner, which times the depths among its
candidates, is checked in test_fdtdz_gpu.py.
"""

import numpy as np
import pytest
import torch

import test_fdtdz_gpu as T
from oracle import adjoint
from oracle import c_oracle as C
from oracle import helpers as H
from pjz_amd import fdtdz_torch as F

pytestmark = pytest.mark.gpu

DEEP = [("stream3", 32), ("stream4", 32), ("stream3", 64), ("stream4", 64), ("stream3", 17),
        ("stream4", 50)]


def run(kw, kernel, **lp):
  kw = dict(kw, launch_params=dict(kernel=kernel, **lp))
  out = F.fdtdz(**kw)
  torch.cuda.synchronize()
  return out.cpu().numpy()


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("kernel,Z", DEEP)
def test_deep_matches_oracle_and_two_step(kernel, Z, axis):
  print(f"case kernel={kernel} Z={Z} axis={axis}", flush=True)
  kw = T.random_problem(axis, X=24, Y=20, Z=Z, tt=120, seed=Z + axis)
  ref = C.fdtdz(**kw)
  got = run(kw, kernel)
  assert got.shape == ref.shape
  assert T.rel_l2(got, ref) <= T.TOL
  assert T.rel_l2(got, run(kw, "stream2")) <= 1e-6


@pytest.mark.parametrize("kernel", ["stream3", "stream4"])
@pytest.mark.parametrize("shape", [(17, 13, 9), (10, 9, 64), (5, 31, 40), (33, 4, 17)])
def test_deep_odd_shapes(kernel, shape):
  X, Y, Z = shape
  print(f"case kernel={kernel} shape={shape}", flush=True)
  rng = np.random.default_rng(X * Y + Z)
  kw = T.random_problem(2, X=20, Y=20, Z=20, tt=50, seed=X * Y, widths=(2, 3))
  kw.update(epsilon=(1 + 4 * rng.random((3, X, Y, Z))).astype(np.float32),
            source_field=rng.standard_normal((2, 2, X, Y, 1)).astype(np.float32),
            source_position=Z // 2, offset=(0, 0, 0),
            absorption_mask=(0.05 * rng.random((3, X, Y))).astype(np.float32),
            pml_kappa=np.ones((Z, 2), np.float32),
            pml_sigma=H.pml_sigma((2, 3), Z, 0.5, 3.0).astype(np.float32),
            pml_alpha=np.full((Z, 2), 0.05, np.float32))
  assert T.rel_l2(run(kw, kernel), C.fdtdz(**kw)) <= T.TOL


@pytest.mark.parametrize("tile_x", [1, 2, 3, 5, 9])
@pytest.mark.parametrize("kernel,Z", [("stream3", 32), ("stream4", 32), ("stream3", 64)])
def test_deep_short_x_chunks(kernel, Z, tile_x):
  """Chunks shorter than the depth: a chunk's ghost planes reach back past
  several chunks and, near x = 0, to the domain start."""
  print(f"case kernel={kernel} Z={Z} tile_x={tile_x}", flush=True)
  kw = T.random_problem(0, X=21, Y=15, Z=Z, tt=80, seed=tile_x)
  assert T.rel_l2(run(kw, kernel, tile_x=tile_x), C.fdtdz(**kw)) <= T.TOL


@pytest.mark.parametrize("tt", [49, 50, 51, 52])
@pytest.mark.parametrize("kernel", ["stream3", "stream4"])
def test_deep_step_remainders_and_every_level_snapshots(kernel, tt):
  """Snapshots after every step of the last 13 (each level of a launch
  writes some), with run lengths leaving 0..3 steps for the tail."""
  print(f"case kernel={kernel} tt={tt}", flush=True)
  kw = T.random_problem(1, X=18, Y=22, Z=48, tt=tt, seed=tt)
  kw["output_steps"] = (tt - 13, tt, 1)
  ref = C.fdtdz(**kw)
  got = run(kw, kernel)
  assert got.shape == ref.shape == (13, 3) + kw["epsilon"].shape[1:]
  assert T.rel_l2(got, ref) <= T.TOL
  # every snapshot on its own (a wrong level's store would show in one)
  for i in range(13):
    assert T.rel_l2(got[i], ref[i]) <= T.TOL


@pytest.mark.parametrize("kernel", ["stream3", "stream4"])
def test_deep_z_as_batch(kernel):
  kw = T.random_problem(1, seed=7)
  Z = kw["pml_sigma"].shape[0]
  kw["pml_kappa"] = np.full((Z, 2), np.inf, np.float32)
  kw["pml_sigma"] = np.zeros((Z, 2), np.float32)
  kw["pml_alpha"] = np.zeros((Z, 2), np.float32)
  assert T.rel_l2(run(kw, kernel), C.fdtdz(**kw)) <= T.TOL


@pytest.mark.parametrize("kernel,Z", [("stream3", 32), ("stream4", 40), ("stream3", 64)])
def test_deep_gradient_matches_adjoint_oracle(kernel, Z):
  """Segment recomputes on the deep kernel capture every level's E change."""
  kw = T.random_problem(0, X=20, Y=17, Z=Z, tt=70, seed=Z, widths=(6, 5))
  eps = kw.pop("epsilon")
  rng = np.random.default_rng(3)
  nn = len(range(*kw["output_steps"]))
  g = rng.standard_normal((nn, 3) + eps.shape[1:]).astype(np.float32)
  ref = adjoint.grad_epsilon(eps, g_out=g, **kw)
  e = torch.tensor(eps, device="cuda", requires_grad=True)
  out = F.fdtdz(e, **dict(kw, launch_params=dict(kernel=kernel, ckpt_every=16)))
  (out * torch.as_tensor(g, device="cuda")).sum().backward()
  got = e.grad.cpu().numpy()
  err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
  print(f"gradient {kernel} Z={Z}: rel L2 {err:.2e}", flush=True)
  assert err <= 1e-3


def test_deep_native_shape_short_run():
  """512x512x64 (one cell per lane) for 24 steps, every depth against the
  two-step kernel, and depth 3 against the oracle."""
  from pjz_amd import workloads
  kw = workloads.native_512(steps=24, zz=64)
  ref = run(kw, "stream2")
  for kernel in ("stream3", "stream4"):
    got = run(kw, kernel)
    err = T.rel_l2(got, ref)
    print(f"native 512x512x64 {kernel} vs stream2: {err:.2e}", flush=True)
    assert err <= 1e-6
  assert T.rel_l2(ref, C.fdtdz(**kw)) <= T.TOL


def test_auto_takes_three_steps_on_large_low_grids():
  """The default kernel choice: three steps per launch at 512x512x64, two at
  64x64x64 (too few tiles) and at 512x512x128 (two cells per lane)."""
  from pjz_amd import _lib, workloads

  def stepper(kw):
    shape = {k: tuple(v.shape) for k, v in kw.items() if hasattr(v, "shape")}
    return _lib.stepper(F.make_problem(
        shape["epsilon"], kw["dt"], shape["source_field"], kw["source_position"],
        shape["absorption_mask"], shape["pml_kappa"], kw["pml_widths"], kw["output_steps"],
        shape["source_waveform"][0], kw["use_reduced_precision"], kw["launch_params"],
        kw["offset"]))

  assert stepper(workloads.native_512(steps=30, zz=64)) == _lib.KERNEL_STREAM3
  assert stepper(workloads.layered_64(steps=30)) == _lib.KERNEL_STREAM2
  assert stepper(workloads.native_512(steps=30)) == _lib.KERNEL_STREAM2
  kw = workloads.native_512(steps=30, zz=64)
  assert np.array_equal(run(kw, "auto"), run(kw, "stream3"))
