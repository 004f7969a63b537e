"""Pins the C oracle (oracle/fdtd_ref.c) against the pinned numpy oracle."""

import numpy as np
import pytest

from oracle import c_oracle as C, fdtd_spec as S, helpers as H


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_c_oracle_matches_numpy_oracle(axis):
  rng = np.random.default_rng(20 + axis)
  X, Y, Z = 12, 10, 14
  eps = 1 + rng.random((3, 8, 7, 9))
  sf = {0: rng.standard_normal((2, 1, Y, Z)), 1: rng.standard_normal((2, X, 1, Z)),
        2: rng.standard_normal((2, 2, X, Y, 1))}[axis]
  pos = {0: 4, 1: 6, 2: 5}[axis]
  args = (eps, 0.5, sf, rng.standard_normal((40, 2)), pos, 0.1 * rng.random((3, X, Y)),
          1 + rng.random((Z, 2)), H.pml_sigma((3, 4), Z, 0.5, 3.0),
          0.05 * rng.random((Z, 2)), (3, 4), (20, 40, 3))
  a = S.fdtdz(*args, offset=(2, 1, 3))
  b = C.fdtdz(*args, offset=(2, 1, 3), threads=2)
  np.testing.assert_allclose(b, a, rtol=0, atol=1e-12 * np.abs(a).max())


def test_c_oracle_kappa_inf_and_one_quadrature():
  rng = np.random.default_rng(30)
  X, Y, Z = 9, 8, 6
  eps = 1 + rng.random((3, X, Y, Z))
  kap = np.full((Z, 2), np.inf)
  args = (eps, 0.4, rng.standard_normal((2, X, Y, 1)), rng.standard_normal((25, 2)), 3,
          0.1 * rng.random((3, X, Y)), kap, np.zeros((Z, 2)), np.zeros((Z, 2)), (1, 1),
          (5, 25, 5))
  np.testing.assert_allclose(C.fdtdz(*args, threads=1), S.fdtdz(*args), rtol=0, atol=1e-12)
