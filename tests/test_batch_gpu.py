"""Batched forward (pjz_fdtd_batch_forward) on the GPU.

Each simulation of an uneven batch must match the FP64 oracle (relative L2
<= 1e-4) and the same simulation run alone (bit for bit where both run the
same kernel schedule: every cell sees the same arithmetic whatever tile or
launch it is computed in; <= 1e-6 where a last odd step or the launch depth
differs).
"""

import numpy as np
import pytest
import torch

from oracle import c_oracle as C
from oracle import helpers as H
from pjz_amd import fdtdz_torch as F
from pjz_amd import workloads

pytestmark = pytest.mark.gpu


def sim(seed, X, Y, Z, tt, axis):
  rng = np.random.default_rng(seed)
  eps = (1 + 3 * rng.random((3, X - 4, Y - 2, Z - 6))).astype(np.float32)
  sf = {0: rng.standard_normal((2, 1, Y, Z)), 1: rng.standard_normal((2, X, 1, Z)),
        2: rng.standard_normal((2, 2, X, Y, 1))}[axis].astype(np.float32)
  t = np.arange(tt)
  wave = np.stack([np.sin(0.31 * t), np.cos(0.31 * t)], -1) * np.exp(-((t - 30) / 12.0) ** 2)[:, None]
  widths = (4, 5)
  return dict(epsilon=eps, dt=0.5, source_field=sf, source_waveform=wave.astype(np.float32),
              source_position={0: 6, 1: 4, 2: Z // 2}[axis],
              absorption_mask=H.absorption_mask(X, Y, 4, 5e-3).astype(np.float32),
              pml_kappa=(1 + rng.random((Z, 2))).astype(np.float32),
              pml_sigma=H.pml_sigma(widths, Z, 0.5, 3.0).astype(np.float32),
              pml_alpha=np.full((Z, 2), 0.05, np.float32), pml_widths=widths,
              output_steps=(tt - 20, tt, 5), use_reduced_precision=False, launch_params=None,
              offset=(2, 1, 3))


CASES = [(1, 24, 18, 40, 70, 0), (2, 37, 21, 64, 61, 1), (3, 20, 29, 96, 70, 2),
         (4, 140, 17, 128, 44, 0), (5, 19, 30, 24, 53, 2)]


# tile_y 0: rows per tile from the queue model (8 rows for batches this
# small); 13: the 16-wave tiles large batches of zz <= 64 simulations get.
@pytest.mark.parametrize("tile_y", [0, 13])
def test_uneven_batch_matches_oracle_and_single_runs(tile_y):
  calls = [dict(sim(*c), launch_params=dict(tile_y=tile_y)) for c in CASES]
  outs = F.fdtdz_batch(calls)
  torch.cuda.synchronize()
  for case, kw, out in zip(CASES, calls, outs):
    print(f"case {case} tile_y {tile_y}", flush=True)
    got = out.cpu().numpy()
    ref = C.fdtdz(**kw)
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    assert err <= 1e-4, (case, err)
    single = F.fdtdz(**dict(kw, launch_params=dict(kernel="stream2"))).cpu().numpy()
    if workloads.time_steps(kw) % 2 == 0:
      np.testing.assert_array_equal(got, single)
    else:
      # An odd last step runs as level 1 of a pair in the batch but in the
      # one-step kernel alone: same formulas, different instruction order.
      assert np.linalg.norm(got - single) <= 1e-6 * np.linalg.norm(single)


def test_forward_batch_three_steps_per_launch():
  """A forward-only batch whose zz <= 64 simulations fill the device runs
  that class three steps a launch (the zz = 128 one in pairs): uneven step
  counts (remainders 0, 1, 2), snapshots every step at the end, against the
  oracle and each simulation alone on the three-step kernel."""
  cases = [(11, 150, 140, 64, 60, 0), (12, 130, 150, 40, 50, 1), (13, 141, 122, 64, 49, 2),
           (14, 90, 70, 128, 44, 0)]
  calls = []
  for c in cases:
    kw = sim(*c)
    tt = c[4]
    kw["output_steps"] = (tt - 7, tt, 1)
    calls.append(kw)
  outs = F.fdtdz_batch(calls)
  torch.cuda.synchronize()
  for case, kw, out in zip(cases, calls, outs):
    print(f"case {case}", flush=True)
    got = out.cpu().numpy()
    ref = C.fdtdz(**kw)
    assert np.linalg.norm(got - ref) <= 1e-4 * np.linalg.norm(ref), case
    kernel = "stream3" if case[3] <= 64 else "stream2"
    alone = F.fdtdz(**dict(kw, launch_params=dict(kernel=kernel))).cpu().numpy()
    print(f"  vs {kernel} alone: bit for bit {np.array_equal(got, alone)}", flush=True)
    assert np.linalg.norm(got - alone) <= 1e-6 * np.linalg.norm(alone), case


@pytest.mark.parametrize("tile_y", [0, 8, 13])
def test_batch64_builder_small(tile_y):
  calls = [dict(kw, launch_params=dict(tile_y=tile_y))
           for kw in workloads.batch64(steps=40, n=6, sizes=(32, 48), heights=(64, 128))]
  outs = F.fdtdz_batch(calls)
  for kw, out in zip(calls, outs):
    single = F.fdtdz(**dict(kw, launch_params=dict(kernel="stream2")))
    if workloads.time_steps(kw) % 2 == 0:
      assert torch.equal(out, single)
    else:  # odd last step: level 1 of a pair here, the one-step kernel alone
      diff = float(torch.linalg.vector_norm(out - single))
      assert diff <= 1e-6 * float(torch.linalg.vector_norm(single)) + 1e-30


def test_batch_rejects_unsupported_height():
  kw = sim(9, 20, 20, 129, 30, 0)
  with pytest.raises(ValueError, match="zz"):
    F.fdtdz_batch([kw])


def grad_case(seed, X, Y, Z, tt, axis):
  kw = sim(seed, X, Y, Z, tt, axis)
  rng = np.random.default_rng(100 + seed)
  nn = len(range(*kw["output_steps"]))
  g = rng.standard_normal((nn,) + kw["epsilon"].shape).astype(np.float32)
  return kw, g


# Uneven sizes, both cells-per-lane classes, odd and even step counts, every
# source axis; step counts 61/70/53 put segment ends inside the last pair.
GRAD_CASES = [(11, 24, 18, 40, 70, 0), (12, 30, 21, 64, 61, 1), (13, 20, 23, 96, 70, 2),
              (14, 19, 30, 24, 53, 2)]


@pytest.mark.parametrize("ckpt,tile_y", [(8, 0), (22, 0), (22, 13)])
def test_batch_gradient_matches_single_and_oracle(ckpt, tile_y):
  cases = [grad_case(*c) for c in GRAD_CASES]
  eps = [torch.tensor(kw["epsilon"], device="cuda", requires_grad=True) for kw, _ in cases]
  calls = [dict(kw, epsilon=e, launch_params=dict(ckpt_every=ckpt, tile_y=tile_y))
           for (kw, _), e in zip(cases, eps)]
  outs = F.fdtdz_batch(calls)
  loss = sum((o * torch.as_tensor(g, device="cuda")).sum() for o, (_, g) in zip(outs, cases))
  loss.backward()
  torch.cuda.synchronize()
  for case, e, (kw, g), out in zip(GRAD_CASES, eps, cases, outs):
    print(f"case {case} ckpt {ckpt}", flush=True)
    # The same simulation alone, differentiated by the single-call adjoint.
    e1 = torch.tensor(kw["epsilon"], device="cuda", requires_grad=True)
    o1 = F.fdtdz(**dict(kw, epsilon=e1, launch_params=dict(kernel="stream2", ckpt_every=ckpt)))
    (o1 * torch.as_tensor(g, device="cuda")).sum().backward()
    got, single = e.grad.cpu().numpy(), e1.grad.cpu().numpy()
    assert np.linalg.norm(got - single) <= 1e-5 * np.linalg.norm(single), case
    assert np.linalg.norm(out.detach().cpu().numpy() - o1.detach().cpu().numpy()) <= \
        1e-6 * np.linalg.norm(o1.detach().cpu().numpy())
  # Two cases against the FP64 discrete adjoint oracle.
  from oracle import adjoint
  for i in (0, 3):
    kw, g = cases[i]
    ref = adjoint.grad_epsilon(kw["epsilon"], g_out=g,
                               **{k: v for k, v in kw.items() if k != "epsilon"})
    got = eps[i].grad.cpu().numpy()
    assert np.linalg.norm(got - ref) <= 1e-3 * np.linalg.norm(ref), GRAD_CASES[i]


def test_batch_gradient_only_where_asked():
  (kw0, g0), (kw1, g1) = grad_case(21, 20, 18, 40, 50, 0), grad_case(22, 22, 20, 64, 50, 1)
  e0 = torch.tensor(kw0["epsilon"], device="cuda", requires_grad=True)
  outs = F.fdtdz_batch([dict(kw0, epsilon=e0), kw1])
  (outs[0] * torch.as_tensor(g0, device="cuda")).sum().backward()
  ref = F.fdtdz(**dict(kw0, launch_params=dict(kernel="stream2"))).cpu().numpy()
  assert np.linalg.norm(outs[0].detach().cpu().numpy() - ref) <= 1e-6 * np.linalg.norm(ref)
  assert e0.grad is not None and torch.isfinite(e0.grad).all()


@pytest.mark.parametrize("graph_steps", [2, 6, 64])
def test_graph_chunks_match_direct_launches(graph_steps):
  """Replayed hipGraph chunks of step pairs give the same fields and
  gradients, bit for bit, as one launch per pair (forward and the batched
  backward's segment recomputes)."""
  cases = [grad_case(*c) for c in GRAD_CASES]

  def run(gs):
    eps = [torch.tensor(kw["epsilon"], device="cuda", requires_grad=True) for kw, _ in cases]
    calls = [dict(kw, epsilon=e, launch_params=dict(ckpt_every=22, graph_steps=gs))
             for (kw, _), e in zip(cases, eps)]
    outs = F.fdtdz_batch(calls)
    sum((o * torch.as_tensor(g, device="cuda")).sum() for o, (_, g) in zip(outs, cases)).backward()
    torch.cuda.synchronize()
    return [o.detach().cpu().numpy() for o in outs], [e.grad.cpu().numpy() for e in eps]

  outs0, grads0 = run(0)
  outs1, grads1 = run(graph_steps)
  for a, b in zip(outs0, outs1):
    np.testing.assert_array_equal(a, b)
  for a, b in zip(grads0, grads1):
    # grad_eps sums edge-replicated cells with atomics: order may differ.
    assert np.linalg.norm(a - b) <= 1e-6 * np.linalg.norm(a)


def test_value_and_grad_in_waves_equals_one_batch():
  """fdtdz_batch_value_and_grad with a budget that forces one simulation
  per wave gives the objective and gradients of one batch over all calls;
  so does fdtdz_batch itself under that budget (extra forward per wave)."""
  cases = [grad_case(*c) for c in GRAD_CASES]
  gs = [torch.as_tensor(g, device="cuda") for _, g in cases]

  def run(mode):
    eps = [torch.tensor(kw["epsilon"], device="cuda", requires_grad=True) for kw, _ in cases]
    calls = [dict(kw, epsilon=e, launch_params=dict(ckpt_every=22)) for (kw, _), e in
             zip(cases, eps)]
    probs = F._batch_problems(calls)[0]
    for p in probs:
      p.ckpt_every = 22
    one = max(F._batch_workspace_bytes(F._subset(probs, [i])) for i in range(len(probs)))
    if mode == "all":
      outs = F.fdtdz_batch(calls)
      loss = sum((o * g).sum() for o, g in zip(outs, gs))
      loss.backward()
    elif mode == "waves":
      assert len(F.plan_waves(probs, one)) == len(probs)
      outs = F.fdtdz_batch(calls, memory_budget=one)
      loss = sum((o * g).sum() for o, g in zip(outs, gs))
      loss.backward()
    else:
      loss = F.fdtdz_batch_value_and_grad(calls, lambda i, o: (o * gs[i]).sum(),
                                          memory_budget=one)
    torch.cuda.synchronize()
    return float(loss), [e.grad.cpu().numpy() for e in eps]

  l0, g0 = run("all")
  for mode in ("waves", "value_and_grad"):
    l1, g1 = run(mode)
    assert abs(l1 - l0) <= 1e-6 * abs(l0), mode
    for a, b in zip(g1, g0):
      assert np.linalg.norm(a - b) <= 1e-6 * np.linalg.norm(b), mode
