"""Pins the field() glue of the oracle (oracle/field_ref.py) by mathematics.

The reference has no field() test, and field_ref restates
src/pjz/_field.py:171-279, so comparing the product glue with it alone would
pass a misreading of pjz's code.  These tests check field_ref's output
against what the mathematics fixes instead: a continuous-wave run reaches a
time-harmonic steady state, so its complex fields must satisfy the
time-harmonic form of the discrete update, with a residual that vanishes
everywhere except on the source plane, where it equals the source field()
builds from the mode: mode / eps on the requested plane, a -i phase (sine
waveform) for x/y planes, the [imag, real] quadratures for a z plane, and
1/ww per frequency.  The derivation is in tests/cw_checks.py.

Each planted-error variant alters one step of the glue in the oracle's
inputs (a source moved one cell, the /eps scaling dropped, the z quadratures
swapped, the odd-position waveform flip left out, the 1/ww averaging lost)
and must fail the same check.
"""

import numpy as np
import pytest

import cw_checks as CW
from oracle import field_ref
from pjz_amd._field import SimParams

P = SimParams(omega_range=(2 * np.pi / 22, 2 * np.pi / 18), tt=2500, absorption_padding=8,
              absorption_coeff=4e-3, pml_widths=(6, 6), pml_alpha_coeff=0.05,
              use_reduced_precision=False)
OMEGA = np.array([2 * np.pi / 21, 2 * np.pi / 19])
SHAPE = (16, 14, 12)
TOL = 2e-5


def structure(seed=0):
  rng = np.random.default_rng(seed)
  return 1.5 + 2.0 * rng.random((3,) + SHAPE)


CASES = [
    # (axis, position, complex source, z as batch, reduced precision, frequencies)
    ("x", 7, False, False, False, 2),   # odd x: shifted + flipped waveform
    ("x", 6, False, False, False, 2),   # even x
    ("y", 5, False, False, False, 2),   # odd y
    ("y", 8, False, False, False, 1),   # even y, one frequency
    ("z", 5, True, False, False, 2),    # z plane, complex mode: two quadratures
    ("z", 6, True, False, False, 1),
    ("x", 9, False, True, False, 2),    # z as batch (kappa = inf)
    ("x", 7, False, False, True, 2),    # pjz's default 128-cell z budget
]


@pytest.mark.parametrize("axis,pos,cplx,zbatch,reduced,nw", CASES)
def test_field_ref_is_the_cw_steady_state(axis, pos, cplx, zbatch, reduced, nw):
  eps = structure()
  # z slices that do not couple lose energy only through the xy absorber, so
  # their transient dies more slowly (residual 3.4e-5 after 4000 steps)
  p = P._replace(use_z_as_batch=zbatch, use_reduced_precision=reduced,
                 tt=4000 if zbatch else P.tt)
  omega = OMEGA[:nw] if nw == 2 else np.array([2 * np.pi / 20])
  mode, J = CW.divergence_free_source(axis, SHAPE, eps, pos, complex_=cplx, only_normal=zbatch)
  fields = field_ref.field(eps, mode, omega, pos, p)
  err = CW.cw_error(fields, eps, mode, J, omega, pos, axis, p)
  print(f"{axis}@{pos} cplx={cplx} zbatch={zbatch} reduced={reduced}: residual {err:.2e}")
  assert err <= (5 * TOL if zbatch else TOL)


def _planted(name, kw, eps, pos):
  kw = dict(kw)
  if name == "source two cells over":
    kw["source_position"] = kw["source_position"] + 2
  elif name == "source one cell over":  # z planes take any position
    kw["source_position"] = kw["source_position"] + 1
  elif name == "no waveform flip":  # the odd x position lands one cell over
    kw["source_waveform"] = np.ascontiguousarray(kw["source_waveform"][:, ::-1])
  elif name == "no 1/ww":
    kw["source_waveform"] = kw["source_waveform"] * len(OMEGA)
  elif name == "quadratures swapped":
    kw["source_field"] = np.ascontiguousarray(kw["source_field"][::-1])
  return kw


@pytest.mark.parametrize("name", ["source two cells over", "no eps scaling", "no waveform flip",
                                  "no 1/ww"])
def test_planted_glue_errors_fail_the_x_plane_check(name):
  eps = structure()
  pos = 7
  mode, J = CW.divergence_free_source("x", SHAPE, eps, pos)
  kw = field_ref.inputs(eps, mode, OMEGA, pos, P)
  if name == "no eps scaling":  # the source field is the mode itself
    bad = field_ref.inputs(eps, mode * eps[1:3, pos:pos + 1], OMEGA, pos, P)
  else:
    bad = _planted(name, kw, eps, pos)
  err = CW.cw_error(CW.oracle_fields(bad, OMEGA, P.dt), eps, mode, J, OMEGA, pos, "x", P, kw)
  print(f"planted '{name}': residual {err:.2e}")
  assert err > 100 * TOL


@pytest.mark.parametrize("name", ["quadratures swapped", "source one cell over"])
def test_planted_glue_errors_fail_the_z_plane_check(name):
  eps = structure()
  pos = 5
  mode, J = CW.divergence_free_source("z", SHAPE, eps, pos, complex_=True)
  kw = field_ref.inputs(eps, mode, OMEGA, pos, P)
  bad = _planted(name, kw, eps, pos)
  err = CW.cw_error(CW.oracle_fields(bad, OMEGA, P.dt), eps, mode, J, OMEGA, pos, "z", P, kw)
  print(f"planted '{name}': residual {err:.2e}")
  assert err > 100 * TOL
