"""Permittivity render (pjz_amd.render / epsilon), torch port of pjz's
``src/pjz/_epsilon.py``.

Every case and expected value is the reference's own (tests/test_layers.py),
read from tests/golden/reference_vectors.json:
threshold positions for each component and axis, sub-pixel interpolation as an
interface sweeps a cell, z interfaces, and the corner cases whose values are
closed-form harmonic/arithmetic mixes.  Run in float64 on the CPU and, for a
subset, in float32 on the GPU.
"""

import json
import pathlib

import numpy as np
import pytest
import torch

from pjz_amd import _epsilon as E

GOLDEN = json.loads((pathlib.Path(__file__).parent / "golden" / "reference_vectors.json").read_text())


def zgrid(zz, dtype=torch.float64):
  k = torch.arange(zz, dtype=dtype)[:, None]
  return k + torch.tensor([[-0.5, 0.0]], dtype=dtype), k + torch.tensor([[0.5, 1.0]], dtype=dtype)


def two_layers(xx, yy, m, thresh, vals=(1, 3), dtype=torch.float64):
  layer = vals[0] * torch.ones((2, 2 * m * xx, 2 * m * yy), dtype=dtype)
  layer[1, thresh[0]:, thresh[1]:] = vals[1]
  return layer


SIMPLE = [(tuple(c["thresh"]), tuple(c["index"]), c["expected"])
          for c in GOLDEN["render_simple"]["cases"]]


@pytest.mark.parametrize("thresh,index,expected", SIMPLE)
def test_simple(thresh, index, expected):
  xx, yy, zz, m = 4, 5, 6, 1
  out = E.render(two_layers(xx, yy, m, thresh), torch.tensor([thresh[2]]), *zgrid(zz), m)
  assert out.shape == (3, xx, yy, zz)
  assert float(out[index]) == pytest.approx(expected)


X_THRESHOLD = [(c["m"], c["thresh"], c["component"], c["index"], c["expected"])
               for c in GOLDEN["render_x_threshold"]["cases"]]


@pytest.mark.parametrize("m,thresh,component,index,expected", X_THRESHOLD)
def test_x_threshold(m, thresh, component, index, expected):
  xx, yy, zz = 4, 3, 2
  layer = torch.ones((1, 2 * m * xx, 2 * m * yy), dtype=torch.float64)
  layer[0, thresh:, :] = 3
  out = E.render(layer, torch.zeros(0, dtype=torch.float64), *zgrid(zz), m)
  assert float(out[component, index, 0, 0]) == pytest.approx(expected)


Z_THRESHOLD = [(c["thresh"], c["component"], c["index"], c["expected"])
               for c in GOLDEN["render_z_threshold"]["cases"]]


@pytest.mark.parametrize("thresh,component,index,expected", Z_THRESHOLD)
def test_z_threshold(thresh, component, index, expected):
  layer = torch.ones((2, 2, 2), dtype=torch.float64)
  layer[1] = 3
  out = E.render(layer, torch.tensor([thresh], dtype=torch.float64), *zgrid(1), m=1)
  assert float(out[component, 0, 0, index]) == pytest.approx(expected)


# Corner mixes, harmonic and arithmetic: 1/(0.5 (5/6) + 0.5 (2/3)) and
# 1/((1/3)(11/12) + (2/3)(4/5)) among them.
CORNER = [(tuple(c["thresh"]), tuple(c["index"]), c["expected"])
          for c in GOLDEN["render_corner"]["cases"]]


@pytest.mark.parametrize("thresh,index,expected", CORNER)
def test_corner(thresh, index, expected):
  out = E.render(two_layers(2, 2, 1, thresh), torch.tensor([thresh[2]], dtype=torch.float64),
                 *zgrid(2), m=1)
  assert float(out[index]) == pytest.approx(expected)


def test_differentiable():
  layers = torch.ones((2, 4, 4), dtype=torch.float64, requires_grad=True)
  pos = torch.tensor([1.0], dtype=torch.float64, requires_grad=True)
  E.render(layers, pos, *zgrid(2), 1).sum().backward()
  assert layers.grad.shape == (2, 4, 4) and pos.grad.shape == (1,)
  assert torch.isfinite(layers.grad).all() and torch.isfinite(pos.grad).all()


def test_epsilon_uniform_and_simple_average():
  """A uniform stack renders to its value everywhere; the simple average of
  two layers split at z = 1.25 gives the volume-weighted mean in the cell it
  cuts (component z sits on [1, 2], the others on [0.5, 1.5])."""
  layers = torch.full((2, 6, 4), 2.5, dtype=torch.float64)
  np.testing.assert_allclose(E.epsilon(layers, [1.25], 1, 3).numpy(), 2.5)
  layers[1] = 4.0
  out = E.epsilon(layers, [1.25], 1, 3, use_simple_averaging=True)
  assert float(out[0, 0, 0, 1]) == pytest.approx(0.75 * 2.5 + 0.25 * 4.0)
  assert float(out[2, 0, 0, 1]) == pytest.approx(0.25 * 2.5 + 0.75 * 4.0)


def test_gradient_matches_finite_differences():
  """d(sum of rendered eps)/d(layer pixel) and /d(interface) against
  central differences in float64."""
  rng = np.random.default_rng(3)
  layers = torch.tensor(1 + 3 * rng.random((2, 8, 6)), requires_grad=True)
  pos = torch.tensor([1.3], dtype=torch.float64, requires_grad=True)
  wts = torch.tensor(rng.standard_normal((3, 4, 3, 3)))
  f = lambda l, p: float((E.render(l, p, *zgrid(3), 1) * wts).sum())
  (E.render(layers, pos, *zgrid(3), 1) * wts).sum().backward()
  h = 1e-6
  for idx in [(0, 3, 2), (1, 5, 0), (1, 0, 5)]:
    lp, lm = layers.detach().clone(), layers.detach().clone()
    lp[idx] += h
    lm[idx] -= h
    fd = (f(lp, pos.detach()) - f(lm, pos.detach())) / (2 * h)
    assert float(layers.grad[idx]) == pytest.approx(fd, rel=1e-5, abs=1e-8)
  fd = (f(layers.detach(), pos.detach() + h) - f(layers.detach(), pos.detach() - h)) / (2 * h)
  assert float(pos.grad[0]) == pytest.approx(fd, rel=1e-5, abs=1e-8)


@pytest.mark.gpu
def test_render_on_the_gpu(gpu):
  rng = np.random.default_rng(5)
  layers = 1 + 11 * rng.random((3, 24, 16))
  pos = [1.3, 2.6]
  cpu = E.epsilon(torch.tensor(layers), pos, 2, 5).numpy()
  dev = E.epsilon(torch.tensor(layers, dtype=torch.float32, device=gpu), pos, 2, 5)
  assert dev.device.type == "cuda"
  np.testing.assert_allclose(dev.cpu().numpy(), cpu, rtol=1e-5)
  for thresh, index, expected in CORNER[:11]:
    out = E.render(two_layers(2, 2, 1, thresh, dtype=torch.float32).to(gpu),
                   torch.tensor([thresh[2]], device=gpu),
                   *(g.to(gpu) for g in zgrid(2, torch.float32)), m=1)
    assert float(out[index]) == pytest.approx(expected, rel=1e-5)


@pytest.mark.gpu
def test_render_kernel_reference_cases(gpu):
  """Every literal case of the reference's tests through the HIP render
  (float32 on the GPU): the simple, x-threshold, z-threshold and corner
  families."""
  f32 = torch.float32
  zg = lambda zz: tuple(g.to(gpu) for g in zgrid(zz, f32))
  for thresh, index, expected in SIMPLE:
    out = E.render(two_layers(4, 5, 1, thresh, dtype=f32).to(gpu),
                   torch.tensor([float(thresh[2])], device=gpu), *zg(6), 1)
    assert float(out[index]) == pytest.approx(expected, rel=1e-5), (thresh, index)
  for m, thresh, component, index, expected in X_THRESHOLD:
    layer = torch.ones((1, 2 * m * 4, 2 * m * 3), dtype=f32)
    layer[0, thresh:, :] = 3
    out = E.render(layer.to(gpu), torch.zeros(0, device=gpu), *zg(2), m)
    assert float(out[component, index, 0, 0]) == pytest.approx(expected, rel=1e-5)
  for thresh, component, index, expected in Z_THRESHOLD:
    layer = torch.ones((2, 2, 2), dtype=f32)
    layer[1] = 3
    out = E.render(layer.to(gpu), torch.tensor([float(thresh)], device=gpu), *zg(1), m=1)
    assert float(out[component, 0, 0, index]) == pytest.approx(expected, rel=1e-5)
  for thresh, index, expected in CORNER:
    out = E.render(two_layers(2, 2, 1, thresh, dtype=f32).to(gpu),
                   torch.tensor([float(thresh[2])], device=gpu), *zg(2), m=1)
    assert float(out[index]) == pytest.approx(expected, rel=1e-5), (thresh, index)


@pytest.mark.gpu
@pytest.mark.parametrize("pos,m,simple", [([1.3, 2.6], 2, False), ([2.0, 3.5], 1, False),
                                          ([1.3, 2.6], 2, True), ([], 3, False)])
def test_render_kernel_gradient(gpu, pos, m, simple):
  """The HIP reverse pass against torch.autograd through the float64 torch
  restatement on the CPU, for a random cotangent: interfaces inside cells,
  interfaces on cell bounds of both z grids (the clip's half gradients at
  ties), the simple average, a single layer."""
  rng = np.random.default_rng(len(pos) + m)
  ll = len(pos) + 1
  xx, yy, zz = 5, 4, 6
  layers = 1 + 11 * rng.random((ll, 2 * m * xx, 2 * m * yy))
  g = rng.standard_normal((3, xx, yy, zz))
  lc = torch.tensor(layers, requires_grad=True)
  pc = torch.tensor(pos, dtype=torch.float64, requires_grad=True)
  out = E.render(lc, pc, *zgrid(zz), m, use_simple_averaging=simple)
  (out * torch.tensor(g)).sum().backward()
  lg = torch.tensor(layers, dtype=torch.float32, device=gpu, requires_grad=True)
  pg = torch.tensor(pos, dtype=torch.float32, device=gpu, requires_grad=True)
  og = E.render(lg, pg, *(t.to(gpu) for t in zgrid(zz, torch.float32)), m,
                use_simple_averaging=simple)
  np.testing.assert_allclose(og.detach().cpu().numpy(), out.detach().numpy(), rtol=1e-5)
  (og * torch.tensor(g, dtype=torch.float32, device=gpu)).sum().backward()
  got, ref = lg.grad.cpu().numpy(), lc.grad.numpy()
  assert np.linalg.norm(got - ref) <= 1e-5 * np.linalg.norm(ref)
  if pos:
    np.testing.assert_allclose(pg.grad.cpu().numpy(), pc.grad.numpy(), rtol=1e-4, atol=1e-5)
