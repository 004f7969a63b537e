"""Pins the helper restatements with the reference's own vectors.

Expected values are the literal vectors and properties of the reference tests
tests/test_boundaries.py:8-41, tests/test_waveforms.py:8-14 and
tests/test_frequencies.py:8-38, re-targeted at both the oracle's helpers
(oracle/helpers.py) and the product's (pjz_amd/boundaries.py, which returns
float32 / complex64 like pjz's jax code); the vectors are read from
tests/golden/reference_vectors.json.
"""

import json
import pathlib

import numpy as np
import pytest

from oracle import helpers as OH
from pjz_amd import _field as Fd
from pjz_amd import boundaries as PB

GOLDEN = json.loads((pathlib.Path(__file__).parent / "golden" / "reference_vectors.json").read_text())


@pytest.fixture(params=["oracle", "product"])
def H(request):
  return OH if request.param == "oracle" else PB


def test_absorption_mask_rows(H):  # tests/test_boundaries.py:8-24
  g = GOLDEN["absorption_mask"]
  m = H.absorption_mask(**g["inputs"])
  for row in g["rows"]:
    got = m[row["component"], :, row["y"]] if row["x"] is None else m[row["component"], row["x"], :]
    np.testing.assert_array_equal(got, row["values"])


PML_TABLE = GOLDEN["pml_sigma"]["table"]


def test_pml_sigma_table(H):  # tests/test_boundaries.py:27-41
  i = GOLDEN["pml_sigma"]["inputs"]
  np.testing.assert_array_almost_equal(
      H.pml_sigma(tuple(i["pml_widths"]), i["zz"], i["ln_R"], i["m"]),
      PML_TABLE)


def test_pml_sigma_zero_width(H):
  s = H.pml_sigma((0, 3), 8, 0.5, 2.0)
  assert np.all(s[:5] == 0) and np.all(s[5:] > 0)


def test_ramped_sin_asymptote(H):  # tests/test_waveforms.py:8-14
  omega, dt, tt = 2 * np.pi / 40, 0.5, 100000
  w = H.ramped_sin(omega, width=3, delay=4, dt=dt, tt=tt)
  np.testing.assert_array_almost_equal(
      np.imag(w)[-10:], np.sin(omega * dt * np.arange(tt))[-10:])


def test_sampling_interval_orthogonal(H):  # tests/test_frequencies.py:8-18
  wmin, wmax, n, dt = 2 * np.pi / 40, 2 * np.pi / 36, 10, 0.5
  interval = H.sampling_interval(wmin, wmax, n, dt)
  ws = np.linspace(wmin, wmax, n)
  theta = ws * dt * interval * np.arange(2 * n)[:, None]
  A = np.hstack([np.sin(theta), np.cos(theta)])
  A /= np.linalg.norm(A, ord=2, axis=0)
  assert np.all(A.T @ A - np.eye(2 * n) < 1e-1)


def test_sampling_interval_single(H):
  # one frequency: a quarter period, _field.py:126-128
  assert H.sampling_interval(2 * np.pi / 20, 2 * np.pi / 20, 1, 0.5) == 10


def test_frequency_components_phase_convention(H):  # tests/test_frequencies.py:21-29
  omegas = np.array([2 * np.pi / 40, 2 * np.pi / 39])
  steps = np.array([10, 20, 30, 40])
  dt = 0.5
  signal = np.sin(omegas[0] * dt * steps) + np.cos(omegas[1] * dt * steps)
  out = H.frequency_components(signal[:, None, None, None, None], steps, omegas, dt)
  assert out[0] == pytest.approx(-1j, rel=1e-3)
  assert out[1] == pytest.approx(1, rel=1e-3)


def test_frequency_components_exact_multitone(H):
  rng = np.random.default_rng(3)
  omegas = np.array([2 * np.pi / 22, 2 * np.pi / 20, 2 * np.pi / 18])
  dt, interval = 0.5, H.sampling_interval(omegas.min(), omegas.max(), 3, 0.5)
  steps = np.arange(1000, 1000 + 7 * interval, interval)
  amp = rng.standard_normal(3) + 1j * rng.standard_normal(3)
  t = omegas[:, None] * dt * steps[None, :]
  sig = np.real(amp[:, None] * np.exp(1j * t)).sum(0)
  tol = 1e-9 if H is OH else 1e-6  # the product returns complex64
  np.testing.assert_allclose(H.frequency_components(sig, steps, omegas, dt), amp,
                             rtol=tol, atol=tol)


def test_source_amplitude_half_step(H):  # tests/test_frequencies.py:32-38
  omega, dt, tt = 2 * np.pi / 40, 0.5, 10000
  out = H.source_amplitude(np.sin(omega * dt * (np.arange(tt) - 0.5))[:, None], omega, dt)
  assert out[0] == pytest.approx(-1j, rel=1e-6 if H is OH else 1e-5)


def test_pad_zz_identities(H):
  for ez in (1, 7, 24, 50, 96):
    for pml in ((16, 16), (10, 6), (0, 0)):
      for red in (True, False):
        total = H.zz(pml, red)
        if ez > total:
          continue
        bot, top = H.pad_zz(ez, pml, red)
        assert bot + top + ez == total and 0 <= top - bot <= 1
  assert H.zz((16, 16), True) == 96 and H.zz((16, 16), False) == 32


def test_prop_axis(H):
  prop_axis = H.prop_axis if H is OH else Fd.prop_axis
  assert prop_axis((3, 2, 1, 5, 6)) == "x"
  assert prop_axis((2, 5, 1, 6)) == "y"
  assert prop_axis((2, 5, 6, 1)) == "z"
  with pytest.raises(ValueError):
    prop_axis((2, 1, 1, 6))
