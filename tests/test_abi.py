"""CPU checks of the C ABI: the library loads, exports what the header
declares, and rejects bad problems with ValueError -- no GPU calls."""

import ctypes
import os
import re

import numpy as np
import pytest
import torch

from pjz_amd import _lib
from pjz_amd import fdtdz_torch as F
from pjz_amd import workloads

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
  names = []
  for fn in sorted(os.listdir(os.path.join(ROOT, "include"))):
    if fn.endswith(".h"):
      text = open(os.path.join(ROOT, "include", fn)).read()
      names += re.findall(r"PJZ_FDTD_API\s+[\w\s\*]*?\b(pjz_fdtd_\w+)\s*\(", text)
  return names


def test_header_declares_entry_points():
  names = declared_symbols()
  assert set(names) == set(_lib.EXPORTS)


def test_library_exports_every_declared_symbol():
  lib = ctypes.CDLL(_lib.LIB_PATH)
  for name in declared_symbols():
    assert hasattr(lib, name), name
  assert _lib.load().pjz_fdtd_abi_version() == _lib.ABI_VERSION


def _problem(**over):
  kw = workloads.layered_64(steps=40)
  kw.update(over)
  shapes = {k: np.shape(v) for k, v in kw.items() if isinstance(v, np.ndarray)}
  return F.make_problem(shapes["epsilon"], kw["dt"], shapes["source_field"],
                        kw["source_position"], shapes["absorption_mask"],
                        shapes["pml_kappa"], kw["pml_widths"], kw["output_steps"],
                        shapes["source_waveform"][0], kw["use_reduced_precision"],
                        kw["launch_params"], kw["offset"])


def test_workspace_and_outputs():
  p = _problem()
  assert _lib.num_outputs(p) == 4
  n = _lib.workspace_bytes(p)
  cells = 64 ** 3
  assert n >= 4 * 5 * 3 * cells  # E, H ping-pong + cb


@pytest.mark.parametrize("over, msg", [
    (dict(source_position=11), "even"),
    (dict(output_steps=(0, 100, 1)), "waveform rows"),
    (dict(output_steps=(0, 10, 0)), "step"),
    (dict(pml_widths=(40, 40)), "pml_widths"),
    (dict(offset=(1, 0, 0)), "does not fit"),
    (dict(dt=0.0), "dt"),
])
def test_bad_problem_raises_value_error(over, msg):
  with pytest.raises(ValueError, match=msg):
    _lib.workspace_bytes(_problem(**over))


def test_deep_kernel_ids_checked():
  """Three and four steps per launch: accepted at zz = 64, refused at
  zz = 96 (one cell per lane only)."""
  for kernel in ("stream3", "stream4"):
    assert _lib.workspace_bytes(_problem(launch_params=dict(kernel=kernel))) > 0
  z96 = dict(epsilon=np.zeros((3, 64, 64, 96), np.float32),
             source_field=np.zeros((2, 1, 64, 96), np.float32),
             pml_kappa=np.ones((96, 2), np.float32))
  with pytest.raises(ValueError, match="three-step"):
    _lib.workspace_bytes(_problem(launch_params=dict(kernel="stream3"), **z96))
  with pytest.raises(ValueError, match="four-step"):
    _lib.workspace_bytes(_problem(launch_params=dict(kernel="stream4"), **z96))


def test_bad_shapes_raise_value_error():
  with pytest.raises(ValueError, match="source_field"):
    F.source_geometry((2, 3, 64, 64), (64, 64, 64))
  with pytest.raises(ValueError, match="ambiguous"):
    F.source_geometry((2, 1, 1, 8), (1, 1, 8))
  assert F.source_geometry((2, 2, 8, 9, 1), (8, 9, 10)) == (2, 2)
  assert F.source_geometry((2, 8, 9, 1), (8, 9, 10)) == (2, 1)
  assert F.source_geometry((2, 8, 1, 10), (8, 9, 10)) == (1, 1)
  with pytest.raises(ValueError, match="launch_params"):
    F._launch_options(dict(kernel="nope"))


@pytest.mark.skipif(torch.cuda.is_available(), reason="checks the no-GPU failure")
def test_fdtdz_fails_loudly_without_gpu():
  with pytest.raises(RuntimeError, match="GPU"):
    F.fdtdz(**workloads.layered_64(steps=20))


def test_copy_rejects_bad_arguments_before_any_launch():
  lib = _lib.load()
  buf = ctypes.create_string_buffer(64)
  base = ctypes.addressof(buf)
  assert lib.pjz_fdtd_copy(None, ctypes.c_void_p(base), 16, None) == _lib.EINVAL
  assert lib.pjz_fdtd_copy(ctypes.c_void_p(base), ctypes.c_void_p(base), 15, None) == _lib.EINVAL
  misaligned = base + (16 - base % 16) + 4
  assert lib.pjz_fdtd_copy(ctypes.c_void_p(misaligned), ctypes.c_void_p(base), 16,
                           None) == _lib.EINVAL
  assert "16" in lib.pjz_fdtd_last_error().decode()
  aligned = ctypes.c_void_p(base + (16 - base % 16) % 16)
  assert lib.pjz_fdtd_copy(aligned, aligned, 0, None) == _lib.OK


def test_density_and_render_reject_bad_arguments_before_any_launch():
  """Argument checks of the density and render entry points run on the host
  (no GPU): sizes, radius, layer count, NULL pointers, workspace size."""
  lib = _lib.load()
  sz = ctypes.c_size_t(0)
  assert lib.pjz_fdtd_density_workspace_bytes(1, 16, 12, 2.5, ctypes.byref(sz)) == _lib.OK
  assert sz.value > 0
  assert lib.pjz_fdtd_density_workspace_bytes(0, 16, 12, 2.5, ctypes.byref(sz)) == _lib.EINVAL
  assert lib.pjz_fdtd_density_workspace_bytes(1, 16, 12, 0.0, ctypes.byref(sz)) == _lib.EINVAL
  assert lib.pjz_fdtd_density_workspace_bytes(1, 16, 12, 41.0, ctypes.byref(sz)) == _lib.EINVAL
  assert b"radius" in lib.pjz_fdtd_last_error()
  args = (1, 16, 12, 2.5, 0.5, 1.0, 0.5, 0.25, 0.75)
  assert lib.pjz_fdtd_density_forward(*args, None, None, None, None, None) == _lib.EINVAL
  assert lib.pjz_fdtd_density_forward(1, 16, 12, 2.5, 0.5, 1.0, 0.5, 0.8, 0.2,
                                      None, None, None, None, None) == _lib.EINVAL
  assert b"eta_lo" in lib.pjz_fdtd_last_error()
  buf = ctypes.create_string_buffer(1024)
  p = ctypes.c_void_p(ctypes.addressof(buf))
  assert lib.pjz_fdtd_density_backward(*args, p, p, p, p, p, 16, None) == _lib.ESPACE

  assert lib.pjz_fdtd_render_workspace_bytes(3, 8, 8, 16, ctypes.byref(sz)) == _lib.OK
  assert lib.pjz_fdtd_render_workspace_bytes(17, 8, 8, 16, ctypes.byref(sz)) == _lib.EINVAL
  assert b"layers" in lib.pjz_fdtd_last_error()
  assert lib.pjz_fdtd_render_forward(2, 0, 8, 8, 16, 0, p, p, p, p, p, p, None) == _lib.EINVAL
  assert lib.pjz_fdtd_render_forward(2, 1, 8, 8, 16, 0, p, None, p, p, p, p, None) == _lib.EINVAL
  assert lib.pjz_fdtd_render_backward(2, 1, 8, 8, 16, 0, p, p, p, p, p, p, p, p, p, 16,
                                      None) == _lib.ESPACE
