"""ctypes driver of oracle/fdtd_ref.c -- TEST INFRASTRUCTURE ONLY.

``build()`` compiles fdtd_ref.c with gcc -O2 -fopenmp into
oracle/liboracle_fdtd.so.  ``fdtdz(...)`` has the signature of
``fdtd_spec.fdtdz`` plus ``threads`` and returns float64 snapshots.
"""

import ctypes
import os
import subprocess

import numpy as np

from . import fdtd_spec as S

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "fdtd_ref.c")
LIB = os.path.join(HERE, "liboracle_fdtd.so")
_lib = None


def build():
  tmp = LIB + ".tmp"
  cmd = ["gcc", "-O2", "-fopenmp", "-fPIC", "-shared", "-fno-fast-math", "-ffp-contract=off",
         "-o", tmp, SRC, "-lm"]
  subprocess.run(cmd, check=True, capture_output=True, text=True)
  os.replace(tmp, LIB)
  return LIB


def load():
  global _lib
  if _lib is None:
    if not os.path.exists(LIB):
      build()
    lib = ctypes.CDLL(LIB)
    d, i = ctypes.c_double, ctypes.c_int
    pd = ctypes.POINTER(ctypes.c_double)
    lib.oracle_fdtd_run.restype = i
    lib.oracle_fdtd_run.argtypes = ([i] * 5 + [d] + [pd] * 5 + [i] * 3 + [pd, pd] + [i] * 4 +
                                    [i] * 6 + [pd, i, i, pd, pd])
    _lib = lib
  return _lib


def _p(a):
  return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def fdtdz(epsilon, dt, source_field, source_waveform, source_position,
          absorption_mask, pml_kappa, pml_sigma, pml_alpha, pml_widths,
          output_steps, use_reduced_precision=False, launch_params=None,
          offset=(0, 0, 0), threads=0, nsteps=-1, final_state=False):
  """FP64 snapshots (nn, 3, xx, yy, ze), same contract as fdtd_spec.fdtdz.

  ``nsteps`` >= 0 runs exactly that many steps (timing samples);
  ``final_state`` also returns the final (E, H) arrays.
  """
  lib = load()
  f64 = lambda a: np.ascontiguousarray(np.asarray(a, dtype=np.float64))
  eps = f64(epsilon)
  mask = f64(absorption_mask)
  X, Y = mask.shape[1:]
  Z = np.shape(pml_sigma)[0]
  eps_full = np.ascontiguousarray(S.full_epsilon(np, eps, (X, Y, Z), offset))
  sf = f64(source_field)
  axis = S.source_axis(sf, (X, Y, Z))
  quad = sf.shape[0] if sf.ndim == 5 else 1
  wave = f64(source_waveform)
  start, stop, step = (int(v) for v in output_steps)
  nn = len(range(start, stop, step))
  ex, ey, ez = eps.shape[1:]
  out = np.zeros((nn, 3, ex, ey, ez))
  Ef = np.zeros((3, X, Y, Z)) if final_state else None
  Hf = np.zeros((3, X, Y, Z)) if final_state else None
  null = ctypes.POINTER(ctypes.c_double)()
  rc = lib.oracle_fdtd_run(
      X, Y, Z, int(pml_widths[0]), int(pml_widths[1]), float(dt), _p(eps_full), _p(mask),
      _p(f64(pml_kappa)), _p(f64(pml_sigma)), _p(f64(pml_alpha)), axis,
      int(source_position), quad, _p(sf), _p(wave), wave.shape[0], start, stop, step,
      int(offset[0]), int(offset[1]), int(offset[2]), ex, ey, ez, _p(out), int(nsteps),
      int(threads), _p(Ef) if final_state else null, _p(Hf) if final_state else null)
  if rc != 0:
    raise MemoryError("oracle_fdtd_run could not allocate its arrays")
  if final_state:
    return out, Ef, Hf
  return out
