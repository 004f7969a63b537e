"""Three-field density (pjz_amd.density), torch port of pjz's
``src/pjz/_density.py``.

The reference has no tests for it, so each stage is pinned against something
independent of the code under test: the cone's closed forms (radius 1 is the
identity; radius 1.5 has centre/edge/corner weights 1.5, 0.5, 1.5 - sqrt 2
before normalisation), the blur against scipy.signal.convolve on an
edge-padded array, the boundary mask / projection / inflection / loss against
per-pixel loops written from their definitions, and invariants (a constant
variable is a fixed point of the blur; a binary variable far from the
thresholds has zero loss).
"""

import math

import numpy as np
import pytest
import scipy.ndimage
import scipy.signal
import torch

from pjz_amd import _density as D


def test_cone_closed_forms():
  np.testing.assert_allclose(D.cone(1.0, torch.float64).numpy(),
                             [[0, 0, 0], [0, 1, 0], [0, 0, 0]])
  corner = 1.5 - math.sqrt(2)
  w = np.array([[corner, 0.5, corner], [0.5, 1.5, 0.5], [corner, 0.5, corner]])
  np.testing.assert_allclose(D.cone(1.5, torch.float64).numpy(), w / w.sum(), rtol=1e-12)
  for r in (2.0, 3.7, 6.5):
    k = D.cone(r, torch.float64).numpy()
    assert k.shape == (2 * math.ceil(r - 0.5) + 1,) * 2
    assert k.sum() == pytest.approx(1.0) and np.allclose(k, k.T) and np.allclose(k, k[::-1])


@pytest.mark.parametrize("radius", [1.0, 2.5, 4.0])
def test_blur_matches_scipy(radius):
  rng = np.random.default_rng(1)
  u = rng.random((17, 23))
  k = D.cone(radius, torch.float64).numpy()
  r = k.shape[0] // 2
  ref = scipy.signal.convolve(np.pad(u, r, mode="edge"), k, mode="valid")
  got = D.blur(torch.tensor(u), radius).numpy()
  np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-14)
  np.testing.assert_allclose(D.blur(torch.full((9, 7), 0.3, dtype=torch.float64), radius).numpy(),
                             0.3, rtol=1e-12)


def loop_boundary(v):
  xx, yy = v.shape
  b = v >= 0
  out = np.zeros_like(b)
  for i in range(xx):
    for j in range(yy):
      for di, dj in ((1, 0), (-1, 0), (0, 1), (0, -1)):
        ii, jj = min(max(i + di, 0), xx - 1), min(max(j + dj, 0), yy - 1)
        out[i, j] |= b[ii, jj] != b[i, j]
  return out


def loop_inflection(u, c):
  xx, yy = u.shape
  out = np.zeros_like(u)
  for i in range(xx):
    for j in range(yy):
      gx = u[min(i + 1, xx - 1), j] - u[max(i - 1, 0), j]
      gy = u[i, min(j + 1, yy - 1)] - u[i, max(j - 1, 0)]
      out[i, j] = np.exp(-c * (gx * gx + gy * gy))
  return out


def test_projection_and_loss_against_loops():
  rng = np.random.default_rng(2)
  u = rng.random((12, 15))
  uf = D.blur(torch.tensor(u), 2.0).numpy()
  bnd = loop_boundary(uf - 0.5)
  np.testing.assert_array_equal(D.boundary_cells(torch.tensor(uf - 0.5)).numpy(), bnd)
  proj = np.where(bnd, uf, (uf > 0.5).astype(float) + 0.5 * (uf == 0.5))
  np.testing.assert_allclose(D.project(torch.tensor(uf), 0.5).numpy(), proj, rtol=1e-12)
  infl = loop_inflection(uf, 1.7)
  np.testing.assert_allclose(D.inflection(torch.tensor(uf), 1.7).numpy(), infl, rtol=1e-12)
  loss = np.zeros_like(uf)
  for i in range(uf.shape[0]):
    for j in range(uf.shape[1]):
      lo, hi = min(0.0, uf[i, j] - 0.75), min(0.0, 0.25 - uf[i, j])
      loss[i, j] = infl[i, j] * (proj[i, j] * lo * lo + (1 - proj[i, j]) * hi * hi)
  np.testing.assert_allclose(
      D.geom_loss(torch.tensor(proj), torch.tensor(uf), 1.7, 0.25, 0.75).numpy(), loss,
      rtol=1e-12, atol=1e-300)
  dens, l2 = D.density(torch.tensor(u), 2.0, 0.3, c=1.7)
  np.testing.assert_allclose(dens.numpy(), 0.3 * proj + 0.7 * uf, rtol=1e-12)
  np.testing.assert_allclose(l2.numpy(), loss, rtol=1e-12, atol=1e-300)


def test_binary_far_from_thresholds():
  """A large solid square: alpha = 1 leaves only the blurred rim non-binary,
  and pixels well inside either phase carry no loss."""
  u = np.zeros((40, 40))
  u[10:30, 10:30] = 1
  dens, loss = D.density(torch.tensor(u), 3.0, 1.0)
  dens, loss = dens.numpy(), loss.numpy()
  assert dens[20, 20] == 1 and dens[2, 2] == 0 and loss[20, 20] == 0 and loss[2, 2] == 0
  assert set(np.unique(dens[~D.boundary_cells(D.blur(torch.tensor(u), 3.0) - 0.5).numpy()])) \
      <= {0.0, 1.0}


def test_differentiable():
  u = torch.rand((16, 16), dtype=torch.float64, requires_grad=True)
  dens, loss = D.density(u, 2.5, 0.5)
  (dens.sum() + loss.sum()).backward()
  assert u.grad.shape == (16, 16) and torch.isfinite(u.grad).all()


@pytest.mark.gpu
def test_density_on_the_gpu(gpu):
  """density() on the device against references independent of it: the blur
  by scipy.signal.convolve on the edge-padded variable, boundary cells,
  inflection and loss by the per-pixel loops above, all in float64."""
  rng = np.random.default_rng(4)
  u = rng.random((64, 48))
  radius, alpha, c = 3.5, 0.6, 1.0
  d1, l1 = D.density(torch.tensor(u, dtype=torch.float32, device=gpu), radius, alpha, c=c)
  assert d1.device.type == "cuda"
  k = D.cone(radius, torch.float64).numpy()
  r = k.shape[0] // 2
  uf = scipy.signal.convolve(np.pad(u, r, mode="edge"), k, mode="valid")
  bnd = loop_boundary(uf - 0.5)
  proj = np.where(bnd, uf, (uf > 0.5).astype(float) + 0.5 * (uf == 0.5))
  infl = loop_inflection(uf, c)
  lo, hi = np.minimum(0.0, uf - 0.75), np.minimum(0.0, 0.25 - uf)
  loss = infl * (proj * lo * lo + (1 - proj) * hi * hi)
  # Pixels whose blurred value, or a neighbour's, sits within float32
  # rounding of eta can binarise differently; compare everything else.
  near = np.abs(uf - 0.5) <= 1e-5
  ok = ~(near | np.roll(near, 1, 0) | np.roll(near, -1, 0) | np.roll(near, 1, 1) |
         np.roll(near, -1, 1))
  assert ok.mean() > 0.95
  np.testing.assert_allclose(d1.cpu().numpy()[ok], (alpha * proj + (1 - alpha) * uf)[ok],
                             rtol=1e-5, atol=1e-6)
  np.testing.assert_allclose(l1.cpu().numpy()[ok], loss[ok], rtol=1e-4, atol=1e-7)


def _near_threshold(uf, eta=0.5, tol=1e-5):
  """Pixels whose blurred value, or a 4-neighbour's, is within float32
  rounding of eta (their projection branch may differ between precisions)."""
  near = np.abs(uf - eta) <= tol
  return near | np.roll(near, 1, 0) | np.roll(near, -1, 0) | np.roll(near, 1, 1) | \
      np.roll(near, -1, 1)


@pytest.mark.gpu
@pytest.mark.parametrize("shape,radius", [((64, 48), 3.5), ((37, 29), 1.0), ((20, 33), 2.5),
                                          ((40, 40), 6.0), ((1, 17), 2.0), ((2, 9), 3.0),
                                          ((90, 80), 26.0)])
def test_density_kernel_gradient(gpu, shape, radius):
  """The HIP reverse pass against torch.autograd through the float64 torch
  restatement on the CPU (whose forward the tests above pin), for random
  cotangents of both outputs; pixels near the threshold excluded."""
  rng = np.random.default_rng(int(10 * radius) + shape[0])
  u = rng.random(shape)
  if radius > 10:
    # A wide cone averages noise to ~0.5 everywhere; a ramp keeps most
    # blurred pixels away from the threshold.
    u = np.linspace(0, 1, shape[0])[:, None] + 0.2 * (u - 0.5)
  alpha, c = 0.7, 2.0
  gd, gl = rng.standard_normal(shape), rng.standard_normal(shape)
  uc = torch.tensor(u, requires_grad=True)
  d0, l0 = D.density(uc, radius, alpha, c=c)
  (d0 * torch.tensor(gd)).sum().backward(retain_graph=True)
  gu_d = uc.grad.clone()
  uc.grad = None
  (l0 * torch.tensor(gl)).sum().backward()
  ref = gu_d.numpy() + uc.grad.numpy()
  ug = torch.tensor(u, dtype=torch.float32, device=gpu, requires_grad=True)
  d1, l1 = D.density(ug, radius, alpha, c=c)
  ((d1 * torch.tensor(gd, dtype=torch.float32, device=gpu)).sum() +
   (l1 * torch.tensor(gl, dtype=torch.float32, device=gpu)).sum()).backward()
  got = ug.grad.cpu().numpy()
  uf = D.blur(torch.tensor(u), radius).numpy()
  bad = _near_threshold(uf)
  # A near-threshold pixel's branch feeds the gradient of every pixel its
  # blur reads: compare where no such pixel lies within the cone.
  r = D.cone(radius, torch.float64).shape[0] // 2 + 1
  far = ~(scipy.ndimage.maximum_filter(bad.astype(np.uint8), size=2 * r + 1) > 0)
  assert far.mean() > 0.5
  err = np.linalg.norm((got - ref)[far]) / np.linalg.norm(ref[far])
  print(f"density gradient {shape} r={radius}: rel L2 {err:.2e}", flush=True)
  assert err <= 1e-5


@pytest.mark.gpu
def test_density_kernel_rejects_bad_radius(gpu):
  u = torch.rand((8, 8), device=gpu)
  with pytest.raises(ValueError, match="radius"):
    D.density(u, 41.0, 0.5)
  with pytest.raises(ValueError, match="float32"):
    D.density(u.double(), 2.0, 0.5)


@pytest.mark.gpu
@pytest.mark.parametrize("radius", [1.0, 24.4, 30.0])
def test_density_kernel_blur_both_tilings(gpu, radius):
  """The blur of both tilings (four outputs a lane up to half-width 24, one
  above) against scipy on the edge-padded variable, with ragged edges."""
  rng = np.random.default_rng(int(radius * 10))
  u = rng.random((37, 70))
  k = D.cone(radius, torch.float64).numpy()
  r = k.shape[0] // 2
  ref = scipy.signal.convolve(np.pad(u, r, mode="edge"), k, mode="valid")
  d, _ = D.density(torch.tensor(u, dtype=torch.float32, device=gpu), radius, 0.0)
  np.testing.assert_allclose(d.cpu().numpy(), ref, rtol=1e-5, atol=1e-6)
