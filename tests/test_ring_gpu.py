"""The two-step kernel with its inputs staged through an LDS ring
(k_stream2_ring: two cells per lane, heights in (64, 128] that are multiples
of 4, at most 32 PML cells; the default two-step kernel on those shapes with
8-row tiles) against the FP64 oracle (relative L2 <= 1e-4) and against the
register-prefetch two-step kernel with 12-row tiles (<= 1e-6: the same
per-cell arithmetic, another tiling).

The gradient tests also cover the reverse kernel's LDS-ring form
(k_adj_stream2_ring, the same shapes).

Covered: every source axis; heights 68, 100, 116 and 128; PML widths odd
(psi stored cell by cell), even (psi in pairs), summing to 32 (the ring's
limit) and zero; x chunks of 1-3 planes (ghost planes before the chunk start
and ring copies clamped at the chunk end); step counts leaving a one-step
tail; snapshots after every step; z as batch; the gradient (segment
recomputes capture every level's change); domains one or two planes or rows
thin, forward and gradient; and 512x512x128 for 24 steps.
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


def run(kw, **lp):
  kw = dict(kw, launch_params=dict(kernel="stream2", **lp))
  out = F.fdtdz(**kw)
  torch.cuda.synchronize()
  return out.cpu().numpy()


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("Z,widths", [(68, (6, 5)), (100, (10, 10)), (116, (16, 16)),
                                      (128, (10, 10)), (128, (0, 0)), (100, (3, 8))])
def test_ring_matches_oracle_and_register_kernel(Z, widths, axis):
  print(f"case Z={Z} widths={widths} axis={axis}", flush=True)
  kw = T.random_problem(axis, X=22, Y=19, Z=Z, tt=90, seed=Z + axis, widths=widths)
  ref = C.fdtdz(**kw)
  got = run(kw)
  assert got.shape == ref.shape
  assert T.rel_l2(got, ref) <= T.TOL
  assert T.rel_l2(got, run(kw, tile_y=12)) <= 1e-6


@pytest.mark.parametrize("tile_x", [1, 2, 3])
def test_ring_short_x_chunks(tile_x):
  print(f"case tile_x={tile_x}", flush=True)
  kw = T.random_problem(0, X=17, Y=21, Z=100, tt=60, seed=tile_x, widths=(6, 5))
  got = run(kw, tile_x=tile_x)
  assert T.rel_l2(got, C.fdtdz(**kw)) <= T.TOL
  assert T.rel_l2(got, run(kw, tile_y=12, tile_x=tile_x)) <= 1e-6


@pytest.mark.parametrize("tt", [49, 50])
def test_ring_step_remainder_and_every_step_snapshots(tt):
  print(f"case tt={tt}", flush=True)
  kw = T.random_problem(2, X=18, Y=22, Z=128, tt=tt, seed=tt, widths=(10, 10))
  kw["output_steps"] = (tt - 13, tt, 1)
  ref = C.fdtdz(**kw)
  got = run(kw)
  assert got.shape == ref.shape
  for i in range(13):
    assert T.rel_l2(got[i], ref[i]) <= T.TOL


def test_ring_z_as_batch():
  kw = T.random_problem(1, X=20, Y=18, Z=100, seed=7)
  Z = kw["pml_sigma"].shape[0]
  kw["pml_kappa"] = np.full((Z, 2), np.inf, np.float32)
  kw["pml_sigma"] = np.zeros((Z, 2), np.float32)
  kw["pml_alpha"] = np.zeros((Z, 2), np.float32)
  assert T.rel_l2(run(kw), C.fdtdz(**kw)) <= T.TOL


@pytest.mark.parametrize("axis", [0, 2])
@pytest.mark.parametrize("Z,widths", [(128, (10, 10)), (100, (6, 5)), (116, (16, 16)),
                                      (68, (0, 0)), (100, (3, 8))])
def test_ring_gradient_matches_adjoint_oracle(Z, widths, axis):
  """The forward's segment recomputes and the reverse pairs both run the
  LDS-ring kernels here (k_stream2_ring with captures, k_adj_stream2_ring)."""
  print(f"case Z={Z} widths={widths} axis={axis}", flush=True)
  kw = T.random_problem(axis, X=20, Y=17, Z=Z, tt=60, seed=Z + axis, widths=widths)
  eps = kw.pop("epsilon")
  rng = np.random.default_rng(3)
  nn = len(range(*kw["output_steps"]))
  g = rng.standard_normal((nn, 3) + eps.shape[1:]).astype(np.float32)
  ref = adjoint.grad_epsilon(eps, g_out=g, **kw)
  e = torch.tensor(eps, device="cuda", requires_grad=True)
  out = F.fdtdz(e, **dict(kw, launch_params=dict(kernel="stream2", ckpt_every=16)))
  (out * torch.as_tensor(g, device="cuda")).sum().backward()
  got = e.grad.cpu().numpy()
  err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
  print(f"gradient Z={Z}: rel L2 {err:.2e}", flush=True)
  assert err <= 1e-3


def test_ring_native_shape_short_run():
  """512x512x128 for 24 steps against the 12-row register kernel and the
  oracle."""
  from pjz_amd import workloads
  kw = workloads.native_512(steps=24)
  got = run(kw)
  err = T.rel_l2(got, run(kw, tile_y=12))
  print(f"native 512x512x128 ring vs 12-row kernel: {err:.2e}", flush=True)
  assert err <= 1e-6
  assert T.rel_l2(got, C.fdtdz(**kw)) <= T.TOL


def thin_problem(X, Y, Z, widths=(6, 5), tt=40):
  """A z-source problem on the whole (X, Y, Z) grid, any X, Y >= 1."""
  rng = np.random.default_rng(X * Y + Z)
  kw = T.random_problem(2, X=20, Y=20, Z=Z, tt=tt, seed=Z, widths=widths)
  kw.update(epsilon=(1 + 4 * rng.random((3, X, Y, Z))).astype(np.float32),
            source_field=rng.standard_normal((2, 2, X, Y, 1)).astype(np.float32),
            source_position=Z // 2, offset=(0, 0, 0),
            absorption_mask=(0.05 * rng.random((3, X, Y))).astype(np.float32),
            output_steps=(tt - 10, tt, 3))
  return kw


@pytest.mark.parametrize("shape", [(1, 9, 128), (2, 3, 68), (5, 1, 100), (3, 2, 96),
                                   (40, 1, 128)])
def test_ring_thin_domains(shape):
  """One or two planes or rows: ghost rows and planes past every edge, the
  top row's copy of the row above clamped, ring copies repeating the last
  plane from the first."""
  print(f"case shape={shape}", flush=True)
  kw = thin_problem(*shape)
  assert T.rel_l2(run(kw), C.fdtdz(**kw)) <= T.TOL


@pytest.mark.parametrize("shape", [(2, 3, 68), (5, 1, 100)])
def test_ring_thin_domains_gradient(shape):
  kw = thin_problem(*shape, tt=30)
  eps = kw.pop("epsilon")
  rng = np.random.default_rng(4)
  nn = len(range(*kw["output_steps"]))
  g = rng.standard_normal((nn, 3) + eps.shape[1:]).astype(np.float32)
  ref = adjoint.grad_epsilon(eps, g_out=g, **kw)
  e = torch.tensor(eps, device="cuda", requires_grad=True)
  out = F.fdtdz(e, **dict(kw, launch_params=dict(kernel="stream2", ckpt_every=8)))
  (out * torch.as_tensor(g, device="cuda")).sum().backward()
  got = e.grad.cpu().numpy()
  err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
  print(f"gradient {shape}: rel L2 {err:.2e}", flush=True)
  assert err <= 1e-3
