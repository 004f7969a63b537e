"""pjz_fdtd_forward_noalloc: the forward without a workspace argument.

Called through the C ABI on two problems in a row (the second smaller, so it
reuses the library's scratch array after a larger run), each checked bit for
bit against pjz_fdtd_forward with a caller workspace (same kernels) and
against the FP64 oracle; problems beyond the array's size and runs with
checkpoints are rejected before any launch.
"""

import ctypes

import numpy as np
import pytest
import torch

from oracle import c_oracle
from pjz_amd import _lib
from pjz_amd import fdtdz_torch as F
from pjz_amd import workloads

pytestmark = pytest.mark.gpu


def _inputs(X, Y, Z, steps, kernel="auto"):
  kw = workloads.native_512(steps=steps, xx=X, yy=Y, zz=Z, design=min(X, Y) // 3)
  kw["launch_params"] = dict(kernel=kernel)
  kw["output_steps"] = (steps - 9, steps, 4)
  shapes = {k: np.shape(v) for k, v in kw.items() if isinstance(v, np.ndarray)}
  prob = F.make_problem(shapes["epsilon"], kw["dt"], shapes["source_field"],
                        kw["source_position"], shapes["absorption_mask"], shapes["pml_kappa"],
                        kw["pml_widths"], kw["output_steps"], shapes["source_waveform"][0],
                        kw["use_reduced_precision"], kw["launch_params"], kw["offset"])
  dev = [torch.as_tensor(kw[k], device="cuda").contiguous() for k in
         ("epsilon", "absorption_mask", "pml_kappa", "pml_sigma", "pml_alpha", "source_field",
          "source_waveform")]
  return kw, prob, dev


def _ptr(t):
  return ctypes.c_void_p(t.data_ptr())


def _noalloc(prob, dev):
  out = torch.full((_lib.num_outputs(prob), 3, prob.ex, prob.ey, prob.ez), float("nan"),
                   device="cuda")
  stream = torch.cuda.current_stream().cuda_stream
  rc = _lib.load().pjz_fdtd_forward_noalloc(ctypes.byref(prob), *[_ptr(t) for t in dev],
                                            _ptr(out), ctypes.c_void_p(stream))
  _lib.check(rc, "pjz_fdtd_forward_noalloc")
  return out


def _with_workspace(prob, dev):
  out = torch.empty((_lib.num_outputs(prob), 3, prob.ex, prob.ey, prob.ez), device="cuda")
  nbytes = _lib.workspace_bytes(prob)
  ws = torch.empty(nbytes + 256, dtype=torch.uint8, device="cuda")
  wp = ctypes.c_void_p((ws.data_ptr() + 255) & ~255)
  stream = torch.cuda.current_stream().cuda_stream
  rc = _lib.load().pjz_fdtd_forward(ctypes.byref(prob), *[_ptr(t) for t in dev], _ptr(out), wp,
                                    ctypes.c_size_t(nbytes), ctypes.c_void_p(stream))
  _lib.check(rc, "pjz_fdtd_forward")
  torch.cuda.synchronize()
  return out


@pytest.mark.parametrize("kernel", ["auto", "stream", "reference"])
def test_noalloc_equals_forward_two_calls(kernel):
  for (X, Y, Z, steps) in ((96, 80, 128, 120), (48, 40, 64, 90)):
    kw, prob, dev = _inputs(X, Y, Z, steps, kernel)
    print(f"noalloc {kernel}: {X}x{Y}x{Z}, {steps} steps, "
          f"{_lib.workspace_bytes(prob) / 2**20:.1f} MiB", flush=True)
    got = _noalloc(prob, dev)
    torch.cuda.synchronize()
    ref = _with_workspace(prob, dev)
    assert torch.equal(got, ref), (kernel, X, Y, Z)
    oracle = c_oracle.fdtdz(**kw)
    err = np.linalg.norm(got.cpu().numpy() - oracle) / np.linalg.norm(oracle)
    assert err <= 1e-4, (kernel, X, Y, Z, err)


def test_noalloc_rejects_what_does_not_fit():
  _, prob, dev = _inputs(48, 40, 64, 20)
  prob.ckpt_every = 4
  with pytest.raises(ValueError, match="ckpt_every"):
    _noalloc(prob, dev)
  big = F.make_problem((3, 512, 512, 128), 0.5, (2, 1, 512, 128), 40, (3, 512, 512), (128, 2),
                       (16, 16), (10, 20, 5), 20, True, None, (0, 0, 0))
  assert _lib.workspace_bytes(big) > (256 << 20)
  with pytest.raises(RuntimeError, match="scratch array"):
    _noalloc(big, dev)
