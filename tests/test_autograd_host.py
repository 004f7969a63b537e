"""Host logic of the autograd Functions around the kernels (no GPU).

The launches are replaced by stand-ins that record their calls, so these tests
check what the Functions do with the workspace and the gradient's placement:
a second backward through one graph reruns the forward to rebuild the
released checkpoints and returns the same gradient, and the gradient comes
back on epsilon's device and in its dtype.
"""

import numpy as np
import torch

from pjz_amd import fdtdz_torch as F


def _problem(steps=10):
  return F.make_problem((3, 4, 4, 4), 0.5, (2, 1, 4, 4), 2, (3, 4, 4), (4, 2), (0, 0),
                        (steps - 2, steps, 1), steps, False, None, (0, 0, 0))


class _Recorder:

  def __init__(self):
    self.forwards = 0
    self.backward_ws = []
    self.batch_calls = []

  def forward(self, prob, dev, device):
    self.forwards += 1
    out = torch.ones((F._lib.num_outputs(prob), 3, prob.ex, prob.ey, prob.ez))
    return out, object()

  def backward(self, prob, dev, device, ws, grad_out):
    self.backward_ws.append(ws)
    return grad_out.sum(dim=0).to(torch.float32) * 2.0

  def batch_forward(self, probs, devs, device):
    self.forwards += 1
    self.batch_calls.append(("forward", [p.tt for p in probs], [p.ckpt_every for p in probs]))
    outs = [torch.ones((F._lib.num_outputs(p), 3, p.ex, p.ey, p.ez)) for p in probs]
    return outs, object()

  def batch_backward(self, probs, devs, device, ws, grad_outs):
    self.backward_ws.append(ws)
    self.batch_calls.append(("backward", [p.tt for p in probs], [p.ckpt_every for p in probs]))
    return [g.sum(dim=0).to(torch.float32) * (3.0 + p.tt) for g, p in zip(grad_outs, probs)]


def test_single_backward_twice_reruns_forward(monkeypatch):
  rec = _Recorder()
  monkeypatch.setattr(F, "_forward", rec.forward)
  monkeypatch.setattr(F, "_backward", rec.backward)
  eps = torch.ones((3, 4, 4, 4), dtype=torch.float64, requires_grad=True)
  out = F.FdtdzFunction.apply(eps, _problem(), None, torch.device("cpu"))
  w = torch.arange(out.numel(), dtype=torch.float32).reshape(out.shape)
  loss = (out * w).sum()
  g1, = torch.autograd.grad(loss, eps, retain_graph=True)
  g2, = torch.autograd.grad(loss, eps)
  assert rec.forwards == 2  # the second backward rebuilt the workspace
  assert all(ws is not None for ws in rec.backward_ws)
  assert g1.dtype == torch.float64 and g1.device == eps.device
  np.testing.assert_array_equal(g1.numpy(), g2.numpy())
  np.testing.assert_array_equal(g1.numpy(), (2.0 * w.sum(dim=0)).numpy())


def _batch(steps, ckpt=4):
  probs = (F._lib.Problem * len(steps))()
  for i, s in enumerate(steps):
    probs[i] = _problem(s)
    probs[i].ckpt_every = ckpt
  return probs


def test_batch_backward_twice_reruns_forward(monkeypatch):
  rec = _Recorder()
  monkeypatch.setattr(F, "_batch_forward", rec.batch_forward)
  monkeypatch.setattr(F, "_batch_backward", rec.batch_backward)
  probs = _batch([10, 12])
  eps = [torch.ones((3, 4, 4, 4), dtype=torch.float64, requires_grad=True) for _ in probs]
  outs = F.FdtdzBatchFunction.apply(probs, [None, None], torch.device("cpu"), [[0, 1]], *eps)
  loss = sum((o * (i + 1)).sum() for i, o in enumerate(outs))
  g1 = torch.autograd.grad(loss, eps, retain_graph=True)
  g2 = torch.autograd.grad(loss, eps)
  assert rec.forwards == 2
  assert all(ws is not None for ws in rec.backward_ws)
  for a, b, e in zip(g1, g2, eps):
    assert a.dtype == e.dtype and a.device == e.device
    np.testing.assert_array_equal(a.numpy(), b.numpy())


def test_batch_waves(monkeypatch):
  """Three waves: the forward runs every wave, only the last with
  checkpoints; the backward takes the waves last to first and reruns the
  forward (with checkpoints) of every wave but the last; each simulation's
  gradient comes from its own wave."""
  rec = _Recorder()
  monkeypatch.setattr(F, "_batch_forward", rec.batch_forward)
  monkeypatch.setattr(F, "_batch_backward", rec.batch_backward)
  steps = [10, 12, 14, 16, 18]
  probs = _batch(steps)
  eps = [torch.ones((3, 4, 4, 4), requires_grad=True) for _ in probs]
  waves = [[0, 1], [2], [3, 4]]
  outs = F.FdtdzBatchFunction.apply(probs, [None] * 5, torch.device("cpu"), waves, *eps)
  loss = sum(o.sum() for o in outs)
  grads = torch.autograd.grad(loss, eps)
  assert rec.batch_calls == [
      ("forward", [10, 12], [0, 0]), ("forward", [14], [0]), ("forward", [16, 18], [4, 4]),
      ("backward", [16, 18], [4, 4]),
      ("forward", [14], [4]), ("backward", [14], [4]),
      ("forward", [10, 12], [4, 4]), ("backward", [10, 12], [4, 4])]
  for g, s in zip(grads, steps):
    assert torch.all(g == 2 * (3.0 + s))  # two snapshots of ones, scaled by its own tt


def test_plan_waves_fits_the_budget():
  """Waves cover the batch in order and each wave's workspace (from the
  library's size query, no GPU needed) fits the budget."""
  sizes = [(40, 30, 64), (64, 64, 128), (32, 32, 64), (80, 40, 128), (50, 50, 64)]
  probs = (F._lib.Problem * len(sizes))()
  for i, (X, Y, Z) in enumerate(sizes):
    probs[i] = F.make_problem((3, X, Y, Z), 0.5, (2, 1, Y, Z), 4, (3, X, Y), (Z, 2), (4, 4),
                              (90, 100, 5), 100, False, None, (0, 0, 0))
    probs[i].ckpt_every = 14
  single = [F._batch_workspace_bytes(F._subset(probs, [i])) for i in range(len(sizes))]
  budget = int(1.5 * max(single))
  waves = F.plan_waves(probs, budget)
  assert [i for w in waves for i in w] == list(range(len(sizes)))
  assert len(waves) > 1
  for w in waves:
    assert F._batch_workspace_bytes(F._subset(probs, w)) <= budget
  assert F.plan_waves(probs, 10 * sum(single)) == [list(range(len(sizes)))]
  try:
    F.plan_waves(probs, max(single) - 1)
  except ValueError as e:
    assert "exceeds the memory budget" in str(e)
  else:
    raise AssertionError("an item larger than the budget must raise")


def test_plan_checkpoints_two_levels_when_one_does_not_fit():
  """Single calls: one level every ~sqrt(2 steps) steps when it fits the
  budget; else the two-level spacing with the smallest workspace (the
  library's size query, no GPU), which fits a budget one level cannot meet;
  and a clear error below that."""
  p = F.make_problem((3, 512, 512, 128), 0.5, (2, 1, 512, 128), 40, (3, 512, 512), (128, 2),
                     (16, 16), (9976, 10000, 3), 10000, True, None, (0, 0, 0))
  one = F.default_ckpt_every(F.steps_of(p))
  w_one = F._lib.workspace_bytes(F._with_ckpt(p, one, 0))
  assert F.plan_checkpoints(p, w_one) == (one, 0)
  s1, s2 = F.plan_checkpoints(p, w_one - 1)
  assert 0 < s2 < s1 < F.steps_of(p)
  w_two = F._lib.workspace_bytes(F._with_ckpt(p, s1, s2))
  assert w_two < 0.6 * w_one
  assert F.plan_checkpoints(p, w_two) == (s1, s2)
  try:
    F.plan_checkpoints(p, w_two - 1)
  except ValueError as e:
    assert "memory budget" in str(e)
  else:
    raise AssertionError("a budget below the smallest workspace must raise")


def test_ckpt_inner_is_checked():
  p = F.make_problem((3, 40, 30, 64), 0.5, (2, 1, 30, 64), 4, (3, 40, 30), (64, 2), (4, 4),
                     (90, 100, 5), 100, False, dict(ckpt_inner=5), (0, 0, 0))
  try:
    F._lib.workspace_bytes(p)
  except ValueError as e:
    assert "ckpt_inner" in str(e)
  else:
    raise AssertionError("ckpt_inner without ckpt_every must raise")
  p.ckpt_every = 20
  assert F._lib.workspace_bytes(p) < F._lib.workspace_bytes(F._with_ckpt(p, 20, 0))
