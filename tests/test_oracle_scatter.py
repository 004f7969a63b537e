"""Pins the scatter() oracle (oracle/scatter_ref.py) by mathematics.

The reference has no scatter() test, so scatter_ref (a restatement of
src/pjz/_field.py:293-398) is checked against what a straight, uniform
waveguide fixes:

  * transmission: a guided mode launched at port 1 reaches port 2 unchanged
    in magnitude, |S21| = 1, with the phase exp(-i beta L) of the distance
    L between the ports (beta measured along the guide from the same run);
  * reflection: nothing scatters the mode back, so S11 is only the near
    field of the source sheet next to the port (a few per cent);
  * reciprocity: S12 = S21 (both ports launch the same mode);
  * passivity: sum_j |S_ij|^2 <= 1 (up to that near field);
  * pjz's reciprocity gradient f_i f_j / a_i (_field.py:380-382): the
    derivative of S_ij along any permittivity change d is
    K <f_i f_j / a_i, d>, with one constant K per frequency for every d (the
    discrete Green's function is symmetric); K is close to real and positive
    under this repository's source sign (the sign under which pjz's
    optimisation loop, examples/_utils.py:27-29, descends).

Planted errors (an overlap sampled one cell early, the backward port's beta
left unflipped, f_i conj(f_j) in the gradient) must fail these checks.
"""

import numpy as np
import pytest
import torch

from oracle import field_ref, scatter_ref
from pjz_amd import _mode
from pjz_amd._field import SimParams

XX, YY, ZE = 40, 22, 20
P1, P2 = 8, XX - 9
OMEGA = np.array([2 * np.pi / 21, 2 * np.pi / 19])
P = SimParams(omega_range=(2 * np.pi / 22, 2 * np.pi / 18), tt=4000, absorption_padding=30,
              absorption_coeff=1.1e-3, pml_widths=(6, 6), pml_alpha_coeff=0.05,
              use_reduced_precision=False)


@pytest.fixture(scope="module")
def guide():
  eps = np.ones((3, XX, YY, ZE))
  eps[:, :, 7:15, 8:13] = 4.0  # 8 x 5 core: no degenerate polarisations
  beta, exc, _, _ = _mode.mode(torch.as_tensor(eps[:, :1]), torch.as_tensor(OMEGA), 1)
  m = exc[..., 0].double().numpy()
  b = beta[:, 0].double().numpy()
  svals, grads, fields = scatter_ref.scatter(eps, OMEGA, [m, m], [b, b], (P1, P2),
                                             (True, False), P)
  return dict(eps=eps, mode=m, beta=b, S=np.array(svals), grads=grads, fields=fields)


def _beta_along_guide(g):
  """Propagation constant of the simulated guide: the phase advance per cell
  of the mode overlap between the ports."""
  f1, m = g["fields"][0], g["mode"]
  ov = np.array([[np.sum(m[k] * f1[k][[1, 2], x:x + 1]) for x in range(P1 + 3, P2 - 2)]
                 for k in range(len(OMEGA))])
  return -np.mean(np.angle(ov[:, 1:] / ov[:, :-1]), axis=1)


def _s21(g, overlap=scatter_ref.overlap):
  f1, m, b = g["fields"][0], g["mode"], g["beta"]
  return overlap(m, b, P2, False, f1)[:, 1] / overlap(m, b, P1, True, f1)[:, 0]


def test_straight_guide_transmits_reciprocally(guide):
  S = guide["S"]  # S[i, j] = port i -> port j, per frequency
  print("|S|", np.abs(S).round(4).tolist())
  np.testing.assert_allclose(np.abs(S[0, 1]), 1.0, atol=0.03)
  np.testing.assert_allclose(np.abs(S[1, 0]), 1.0, atol=0.03)
  assert np.max(np.abs(S[1, 0] - S[0, 1])) <= 2e-3
  assert np.max(np.abs(S[0, 0])) <= 0.1 and np.max(np.abs(S[1, 1])) <= 0.1
  assert np.max(np.sum(np.abs(S) ** 2, axis=1)) <= 1.03
  # phase of the transmission: the propagation over the ports' distance (the
  # two-plane fits use the mode solver's beta, ~1% off the simulated guide's)
  phase = np.angle(S[0, 1] * np.exp(1j * _beta_along_guide(guide) * (P2 - P1)))
  print("S21 phase error (rad)", phase)
  assert np.max(np.abs(phase)) <= 0.1


def _overlap_one_cell_early(mode, beta, pos, is_fwd, output):
  if is_fwd:
    return scatter_ref.overlap(mode, beta, pos - 1, is_fwd, output)
  return scatter_ref.overlap(mode, beta, pos, is_fwd, output)


def _overlap_beta_unflipped(mode, beta, pos, is_fwd, output):
  if is_fwd is False:
    return scatter_ref.overlap(mode, -np.asarray(beta), pos, is_fwd, output)
  return scatter_ref.overlap(mode, beta, pos, is_fwd, output)


def test_planted_overlap_errors_fail(guide):
  good = _s21(guide)
  beta = _beta_along_guide(guide)
  early = _s21(guide, _overlap_one_cell_early)
  assert np.min(np.abs(np.angle(early * np.exp(1j * beta * (P2 - P1))))) > 0.2
  unflipped = _s21(guide, _overlap_beta_unflipped)
  assert np.min(np.abs(np.abs(unflipped) - 1.0)) > 0.5
  assert np.max(np.abs(np.abs(good) - 1.0)) <= 0.03


def _directions():
  """Two permittivity changes inside the guide core, at different x."""
  out = []
  for x0 in (14, 21):
    d = np.zeros((3, XX, YY, ZE))
    d[:, x0:x0 + 2, 9:13, 9:12] = 1.0
    out.append(d)
  return out


def _s_of(eps, g):
  f1 = field_ref.field(eps, g["mode"].mean(0), OMEGA, P1, P)
  m, b = g["mode"], g["beta"]
  a1 = scatter_ref.overlap(m, b, P1, True, f1)[:, 0]
  return (scatter_ref.overlap(m, b, P1, True, f1)[:, 1] / a1,
          scatter_ref.overlap(m, b, P2, False, f1)[:, 1] / a1)


def test_reciprocity_gradient_is_proportional_to_the_derivative(guide):
  eps, grads, (f1, f2) = guide["eps"], guide["grads"], guide["fields"]
  a1 = f1 * f1 / grads[0][0]  # a_1 broadcast
  conj_term = f1 * np.conj(f2) / a1  # planted: a conjugated field
  h = 1e-2
  ratios = {"S11": [], "S21": [], "S21 planted": []}
  for d in _directions():
    (s11p, s21p), (s11m, s21m) = _s_of(eps + h * d, guide), _s_of(eps - h * d, guide)
    fd11, fd21 = (s11p - s11m) / (2 * h), (s21p - s21m) / (2 * h)
    dot = lambda G: np.sum(G * d, axis=(1, 2, 3, 4))
    ratios["S11"].append(fd11 / dot(grads[0][0]))
    ratios["S21"].append(fd21 / dot(grads[0][1]))
    ratios["S21 planted"].append(fd21 / dot(conj_term))
  for name, (r1, r2) in ratios.items():
    print(name, "K at two positions:", np.round(r1, 3), np.round(r2, 3))
  for name in ("S11", "S21"):
    r1, r2 = ratios[name]
    assert np.max(np.abs(r1 - r2) / np.abs(r1)) <= 0.15, name
    for r in (r1, r2):
      assert np.all(np.real(r) > 0) and np.max(np.abs(np.angle(r))) <= 0.3, name
  r1, r2 = ratios["S21 planted"]
  assert np.max(np.abs(r1 - r2) / np.abs(r1)) > 0.5
