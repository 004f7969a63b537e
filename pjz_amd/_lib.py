"""ctypes binding of libpjzfdtd.so (C ABI declared in include/pjz_fdtd.h).

The library is loaded from this package directory.  There is no fallback: if
the library is missing, or no GPU is visible, the calls raise.
"""

import ctypes
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpjzfdtd.so")
ABI_VERSION = 1

OK, EINVAL, EHIP, ESPACE = 0, 1, 2, 3
KERNEL_AUTO, KERNEL_REFERENCE, KERNEL_STREAM, KERNEL_STREAM2, KERNEL_AUTOTUNE = 0, 1, 2, 3, 4
KERNEL_STREAM3, KERNEL_STREAM4 = 5, 6

# Every entry point include/pjz_fdtd.h declares.
EXPORTS = (
    "pjz_fdtd_abi_version",
    "pjz_fdtd_last_error",
    "pjz_fdtd_num_outputs",
    "pjz_fdtd_workspace_bytes",
    "pjz_fdtd_forward",
    "pjz_fdtd_forward_noalloc",
    "pjz_fdtd_backward",
    "pjz_fdtd_batch_workspace_bytes",
    "pjz_fdtd_batch_forward",
    "pjz_fdtd_batch_backward",
    "pjz_fdtd_copy",
    "pjz_fdtd_stepper",
    "pjz_fdtd_density_workspace_bytes",
    "pjz_fdtd_density_forward",
    "pjz_fdtd_density_backward",
    "pjz_fdtd_render_workspace_bytes",
    "pjz_fdtd_render_forward",
    "pjz_fdtd_render_backward",
    "pjz_fdtd_mode_workspace_bytes",
    "pjz_fdtd_mode_solve",
)


class Problem(ctypes.Structure):
  """Mirror of ``pjz_fdtd_problem``."""
  _fields_ = [
      ("xx", ctypes.c_int32), ("yy", ctypes.c_int32), ("zz", ctypes.c_int32),
      ("ex", ctypes.c_int32), ("ey", ctypes.c_int32), ("ez", ctypes.c_int32),
      ("x0", ctypes.c_int32), ("y0", ctypes.c_int32), ("z0", ctypes.c_int32),
      ("pml_lo", ctypes.c_int32), ("pml_hi", ctypes.c_int32),
      ("src_axis", ctypes.c_int32), ("src_pos", ctypes.c_int32),
      ("src_quad", ctypes.c_int32), ("tt", ctypes.c_int32),
      ("out_start", ctypes.c_int32), ("out_stop", ctypes.c_int32),
      ("out_step", ctypes.c_int32),
      ("dt", ctypes.c_float),
      ("flags", ctypes.c_int32), ("kernel", ctypes.c_int32),
      ("tile_x", ctypes.c_int32), ("tile_y", ctypes.c_int32),
      ("waves_per_simd", ctypes.c_int32), ("ckpt_every", ctypes.c_int32),
      ("graph_steps", ctypes.c_int32), ("ckpt_inner", ctypes.c_int32),
  ]


_lock = threading.Lock()
_lib = None


def load():
  """Loads (once) and returns the library handle with typed entry points."""
  global _lib
  with _lock:
    if _lib is not None:
      return _lib
    # PJZ_FDTD_LIB selects another build of the same library (A/B timing).
    path = os.environ.get("PJZ_FDTD_LIB") or LIB_PATH
    if not os.path.exists(path):
      raise RuntimeError(
          f"{path} is missing: build it with `python -m pjz_amd.build` "
          "(the FDTD kernels have no CPU fallback)")
    lib = ctypes.CDLL(path)
    P = ctypes.POINTER(Problem)
    vp, sz = ctypes.c_void_p, ctypes.c_size_t
    lib.pjz_fdtd_abi_version.restype = ctypes.c_int
    lib.pjz_fdtd_abi_version.argtypes = []
    lib.pjz_fdtd_last_error.restype = ctypes.c_char_p
    lib.pjz_fdtd_last_error.argtypes = []
    lib.pjz_fdtd_num_outputs.restype = ctypes.c_int
    lib.pjz_fdtd_num_outputs.argtypes = [P, ctypes.POINTER(ctypes.c_int32)]
    lib.pjz_fdtd_workspace_bytes.restype = ctypes.c_int
    lib.pjz_fdtd_workspace_bytes.argtypes = [P, ctypes.POINTER(sz)]
    lib.pjz_fdtd_forward.restype = ctypes.c_int
    lib.pjz_fdtd_forward.argtypes = [P] + [vp] * 8 + [vp, sz, vp]
    lib.pjz_fdtd_forward_noalloc.restype = ctypes.c_int
    lib.pjz_fdtd_forward_noalloc.argtypes = [P] + [vp] * 8 + [vp]
    lib.pjz_fdtd_backward.restype = ctypes.c_int
    lib.pjz_fdtd_backward.argtypes = [P] + [vp] * 9 + [vp, sz, vp]
    lib.pjz_fdtd_batch_workspace_bytes.restype = ctypes.c_int
    lib.pjz_fdtd_batch_workspace_bytes.argtypes = [P, ctypes.c_int, ctypes.POINTER(sz)]
    lib.pjz_fdtd_batch_forward.restype = ctypes.c_int
    lib.pjz_fdtd_batch_forward.argtypes = [P, ctypes.c_int] + [ctypes.POINTER(vp)] * 8 + [vp, sz, vp]
    lib.pjz_fdtd_batch_backward.restype = ctypes.c_int
    lib.pjz_fdtd_batch_backward.argtypes = ([P, ctypes.c_int] + [ctypes.POINTER(vp)] * 9 +
                                            [vp, sz, vp])
    lib.pjz_fdtd_copy.restype = ctypes.c_int
    lib.pjz_fdtd_copy.argtypes = [vp, vp, sz, vp]
    lib.pjz_fdtd_stepper.restype = ctypes.c_int
    lib.pjz_fdtd_stepper.argtypes = [P, ctypes.POINTER(ctypes.c_int32)]
    i32 = ctypes.c_int32
    lib.pjz_fdtd_mode_workspace_bytes.restype = ctypes.c_int
    lib.pjz_fdtd_mode_workspace_bytes.argtypes = [i32] * 4 + [ctypes.POINTER(sz)]
    lib.pjz_fdtd_mode_solve.restype = ctypes.c_int
    lib.pjz_fdtd_mode_solve.argtypes = [i32] * 4 + [vp] * 7 + [i32, ctypes.c_float, vp, sz, vp]
    f32 = ctypes.c_float
    lib.pjz_fdtd_density_workspace_bytes.restype = ctypes.c_int
    lib.pjz_fdtd_density_workspace_bytes.argtypes = [i32] * 3 + [f32, ctypes.POINTER(sz)]
    lib.pjz_fdtd_density_forward.restype = ctypes.c_int
    lib.pjz_fdtd_density_forward.argtypes = [i32] * 3 + [f32] * 6 + [vp] * 4 + [vp]
    lib.pjz_fdtd_density_backward.restype = ctypes.c_int
    lib.pjz_fdtd_density_backward.argtypes = [i32] * 3 + [f32] * 6 + [vp] * 4 + [vp, sz, vp]
    lib.pjz_fdtd_render_workspace_bytes.restype = ctypes.c_int
    lib.pjz_fdtd_render_workspace_bytes.argtypes = [i32] * 4 + [ctypes.POINTER(sz)]
    lib.pjz_fdtd_render_forward.restype = ctypes.c_int
    lib.pjz_fdtd_render_forward.argtypes = [i32] * 6 + [vp] * 6 + [vp]
    lib.pjz_fdtd_render_backward.restype = ctypes.c_int
    lib.pjz_fdtd_render_backward.argtypes = [i32] * 6 + [vp] * 8 + [vp, sz, vp]
    if lib.pjz_fdtd_abi_version() != ABI_VERSION:
      raise RuntimeError("libpjzfdtd.so ABI version mismatch; rebuild it")
    _lib = lib
    return lib


def check(rc, what):
  """Raises ValueError / RuntimeError for a non-zero return code."""
  if rc == OK:
    return
  msg = load().pjz_fdtd_last_error().decode(errors="replace")
  if rc == EINVAL:
    raise ValueError(f"{what}: {msg}")
  raise RuntimeError(f"{what} failed (code {rc}): {msg}")


def workspace_bytes(prob):
  n = ctypes.c_size_t(0)
  check(load().pjz_fdtd_workspace_bytes(ctypes.byref(prob), ctypes.byref(n)),
        "pjz_fdtd_workspace_bytes")
  return n.value


def stepper(prob):
  """The stepper kernel id a forward call with ``prob`` runs on the current
  device (AUTO / AUTOTUNE resolved)."""
  k = ctypes.c_int32(0)
  check(load().pjz_fdtd_stepper(ctypes.byref(prob), ctypes.byref(k)), "pjz_fdtd_stepper")
  return k.value


def num_outputs(prob):
  n = ctypes.c_int32(0)
  check(load().pjz_fdtd_num_outputs(ctypes.byref(prob), ctypes.byref(n)),
        "pjz_fdtd_num_outputs")
  return n.value
