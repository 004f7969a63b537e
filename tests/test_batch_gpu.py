"""Batched forward (pjz_fdtd_batch_forward) on the GPU.

Each simulation of an uneven batch must match the FP64 oracle (relative L2
<= 1e-4) and, bit for bit, the same simulation run alone: every cell sees the
same arithmetic whatever tile or launch it is computed in.
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


def test_uneven_batch_matches_oracle_and_single_runs():
  calls = [sim(*c) for c in CASES]
  outs = F.fdtdz_batch(calls)
  torch.cuda.synchronize()
  for case, kw, out in zip(CASES, calls, outs):
    print(f"case {case}", flush=True)
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


def test_batch64_builder_small():
  calls = workloads.batch64(steps=40, n=6, sizes=(32, 48), heights=(64, 128))
  outs = F.fdtdz_batch(calls)
  for kw, out in zip(calls, outs):
    single = F.fdtdz(**dict(kw, launch_params=dict(kernel="stream2")))
    assert torch.equal(out, single)


def test_batch_rejects_unsupported_height():
  kw = sim(9, 20, 20, 129, 30, 0)
  with pytest.raises(ValueError, match="zz"):
    F.fdtdz_batch([kw])
