"""The benchmark configurations at their real sizes on the GPU.

  * the 64-simulation uneven batch (128..512 x 128..512 x {64, 128}) against
    the FP64 C oracle, every simulation, and batched == alone;
  * the 512x512x128 workload's stack, source, 16+16 CPML and 32-cell absorber
    at a 128x128 cross-section for its full 9998 steps and 8 snapshots
    against the oracle (float32 drift over a whole run);
  * the batched inverse-design iteration at 384x384x128 (64 forward + 64
    adjoint simulations), whose checkpoints do not fit in device memory at
    once: run in waves under a memory budget, each simulation's gradient
    equal to the same simulation differentiated alone, and a directional
    derivative against central differences.
"""

import math

import numpy as np
import pytest
import torch

from oracle import c_oracle as C
from pjz_amd import fdtdz_torch as F
from pjz_amd import workloads

pytestmark = pytest.mark.gpu


def _rel(a, b):
  return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _host(kw):
  return {k: (v.cpu().numpy() if isinstance(v, torch.Tensor) else v) for k, v in kw.items()}


def test_batch64_full_size_matches_oracle_and_single_runs():
  """All 64 simulations of the uneven batch at their sizes, over 40 steps
  (a ramped sine from step 0, snapshots at 10, 20, 30) against the oracle;
  then 300 steps batched against each simulation alone."""
  steps = 40
  t = np.arange(steps)
  wave = np.zeros((steps, 2), np.float32)
  wave[:, 0] = np.sin(2 * math.pi / 20 * 0.5 * t) * (1 - np.exp(-t / 8.0))
  calls = workloads.batch64(steps=steps)
  for kw in calls:
    kw["source_waveform"] = torch.as_tensor(wave, device="cuda")
    kw["output_steps"] = (10, 40, 10)
  cells = sum(workloads.cells(kw) for kw in calls)
  print(f"batch64: {len(calls)} simulations, {cells / 1e6:.0f}M cells", flush=True)
  outs = F.fdtdz_batch(calls)
  torch.cuda.synchronize()
  worst = 0.0
  for i, (kw, out) in enumerate(zip(calls, outs)):
    ref = C.fdtdz(**_host(kw))
    err = _rel(out.cpu().numpy(), ref)
    worst = max(worst, err)
    assert err <= 1e-4, (i, tuple(kw["epsilon"].shape), err)
  print(f"batch64 vs oracle: worst rel L2 {worst:.2e}", flush=True)

  calls = workloads.batch64(steps=300)
  for kw in calls:
    kw["output_steps"] = (260, 300, 10)
  outs = F.fdtdz_batch(calls)
  exact, worst = 0, 0.0
  for i, (kw, out) in enumerate(zip(calls, outs)):
    alone = F.fdtdz(**dict(kw, launch_params=dict(kernel="stream2")))
    exact += int(torch.equal(out, alone))
    err = float(torch.linalg.vector_norm(out - alone) / torch.linalg.vector_norm(alone))
    worst = max(worst, err)
    assert err <= 1e-6, (i, err)
  print(f"batch64 300 steps: batched vs alone worst rel L2 {worst:.1e}, {exact} of 64 "
        f"bit for bit", flush=True)


def test_native_stack_full_run_matches_oracle():
  """The 512x512x128 workload's construction at 128x128x128 (its 50x50
  design square) for the full 10000-step schedule: 9998 steps, 8 snapshots
  (9976..9997), float32 kernel against the FP64 oracle."""
  kw = workloads.native_512(steps=10000, xx=128, yy=128, zz=128, design=50)
  print(f"native stack 128^3: {workloads.time_steps(kw)} steps", flush=True)
  got = F.fdtdz(**kw).cpu().numpy()
  ref = C.fdtdz(**kw)
  err = _rel(got, ref)
  per = [_rel(got[s], ref[s]) for s in range(got.shape[0])]
  print(f"rel L2 over all snapshots {err:.2e}; per snapshot "
        f"{', '.join(f'{e:.1e}' for e in per)}", flush=True)
  assert got.shape[0] == 8
  assert err <= 1e-4


def _design_objective(out, call, omega, monitor_x, mode):
  return workloads.overlap_objective(out, mode, [omega], call["output_steps"],
                                     monitor_x=monitor_x)


def test_design_batch_full_size_in_waves():
  """64 forward + 64 adjoint simulations at 384x384x128 (16 variants x 2
  ports x 2 frequencies, 160x160 design squares), 512 steps.  Each call gets
  its own epsilon tensor.  A 64 GiB memory budget splits the gradient into
  waves (one extra forward pass per wave but the last); the gradient of one
  simulation per variant must equal the same simulation differentiated
  alone, and one directional derivative must match central differences."""
  steps = 512
  calls, variant_of, monitors = workloads.design_batch(steps=steps, xx=384, yy=384, zz=128,
                                                       design=160)
  assert len(calls) == 64
  # Monitors inside the design square, which the 512-step fields reach.
  mon = [(w, 200 if c["source_position"] < 192 else 184, m) for c, (w, _, m)
         in zip(calls, monitors)]
  eps = [c["epsilon"].clone().requires_grad_(True) for c in calls]
  calls = [dict(c, epsilon=e, launch_params=dict(ckpt_every=32)) for c, e in zip(calls, eps)]
  probs = F._batch_problems(calls)[0]
  for p in probs:
    p.ckpt_every = 32
  budget = 64 * 2**30
  waves = F.plan_waves(probs, budget)
  print(f"design batch 384: {len(waves)} waves of {[len(w) for w in waves]}", flush=True)
  assert len(waves) >= 3
  outs = F.fdtdz_batch(calls, memory_budget=budget)
  loss = sum(_design_objective(o, c, *m) for o, c, m in zip(outs, calls, mon))
  loss.backward()
  torch.cuda.synchronize()
  print(f"design batch loss {float(loss):.4e}", flush=True)
  del outs, loss

  for v in range(16):
    i = 4 * v + (v % 4)  # one simulation per variant, every port/frequency
    e1 = calls[i]["epsilon"].detach().clone().requires_grad_(True)
    o1 = F.fdtdz(**dict(calls[i], epsilon=e1, launch_params=dict(kernel="stream2",
                                                                   ckpt_every=32)))
    _design_objective(o1, calls[i], *mon[i]).backward()
    got, alone = eps[i].grad, e1.grad
    err = float(torch.linalg.vector_norm(got - alone) / torch.linalg.vector_norm(alone))
    assert float(torch.linalg.vector_norm(alone)) > 0
    assert err <= 1e-5, (v, i, err)
  print("design batch: per-simulation gradients equal single calls", flush=True)

  i = 0
  base = calls[i]["epsilon"].detach()
  gen = torch.Generator(device="cuda").manual_seed(3)
  d = torch.zeros_like(base)
  reg = (slice(None), slice(112, 200), slice(112, 272), slice(60, 68))
  d[reg] = 0.5 + torch.rand(d[reg].shape, generator=gen, device="cuda")
  analytic = float((eps[i].grad * d).sum())

  def L(e):
    with torch.no_grad():
      return float(_design_objective(F.fdtdz(**dict(calls[i], epsilon=e)), calls[i], *mon[i]))

  h = 1e-2
  fd1 = (L(base + h * d) - L(base - h * d)) / (2 * h)
  fd2 = (L(base + h / 2 * d) - L(base - h / 2 * d)) / h
  fd = (4 * fd2 - fd1) / 3
  print(f"directional: adjoint {analytic:.6e}, differences {fd1:.6e} {fd2:.6e} -> {fd:.6e}",
        flush=True)
  assert abs(fd) > 0 and abs(analytic - fd) <= 5e-3 * abs(fd)
