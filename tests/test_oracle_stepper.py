"""Pins the FP64 stepper oracle (oracle/fdtd_spec.py) against mathematics.

No reference vectors exist for the kernel update (its source is not vendored),
so the oracle is pinned by checks a plausible mistake would fail:
  - per-cell brute force (indexing, difference directions, CPML columns,
    source planes / columns, edge fill);
  - closed-form evolution of a 3-D discrete cavity mode (curl operators, dt
    scaling, boundary conditions);
  - exact discrete energy conservation (curl_b is the transpose of curl_f);
  - per-step mode damping det(M) = exp(-sigma dt / eps) to third order
    (absorber coefficients, sign and factor 2);
  - kappa = inf decouples z slices (2-D cavity closed form per slice);
  - first-step closed form E^1 = cb J(0) (source placement and sign);
  - CPML reflection below a bound (sign of the psi recursion);
  - edge-replicated epsilon equals explicit padding;  linearity;
  - reciprocity (symmetric discrete Green's function: source placement on
    the E samples, curl directions, per-cell epsilon and absorber);
  - time-shift invariance (no step-dependent coefficient).
"""

import math

import numpy as np
import pytest

from oracle import brute, fdtd_spec as S, helpers as H


def _vacuum(X, Y, Z, tt=200):
  return dict(eps=np.ones((3, X, Y, Z)), mask=np.zeros((3, X, Y)),
              kap=np.ones((Z, 2)), sig=np.zeros((Z, 2)), alp=np.zeros((Z, 2)),
              sf=np.zeros((2, 2, X, Y, 1)), wave=np.zeros((tt, 2)))


def _sim(d, dt=0.5, pos=0, widths=(0, 0), **kw):
  return S.Simulation(d["eps"], dt, d["sf"], d["wave"], pos, d["mask"], d["kap"],
                      d["sig"], d["alp"], widths, **kw)


def _mode_1d(N, m):
  """Integer-position and half-position profiles of a discrete cavity mode."""
  k = math.pi * (m + 0.5) / (N + 0.5)
  i = np.arange(N)
  return np.cos(k * (i + 0.5)), np.sin(k * (i + 1)), 2 * math.sin(k / 2)


def _cavity_mode(X, Y, Z, mx=1, my=2, mz=1):
  cx, sx, kx = _mode_1d(X, mx)
  cy, sy, ky = _mode_1d(Y, my)
  cz, sz, kz = _mode_1d(Z, mz)
  kv = np.array([kx, ky, kz])
  A = np.cross(kv, [0.3, 1.0, -0.2])  # transverse => discrete divergence free
  E = [A[0] * sx[:, None, None] * cy[None, :, None] * cz[None, None, :],
       A[1] * cx[:, None, None] * sy[None, :, None] * cz[None, None, :],
       A[2] * cx[:, None, None] * cy[None, :, None] * sz[None, None, :]]
  return E, float((kv ** 2).sum())


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_brute_force_matches_vectorised(axis):
  rng = np.random.default_rng(axis)
  X, Y, Z = 5, 6, 9
  eps = 1 + 3 * rng.random((3, 4, 4, 7))
  mask = 0.3 * rng.random((3, X, Y))
  widths = (3, 2)
  sig = H.pml_sigma(widths, Z, 0.5, 2.0)
  kap = 1 + rng.random((Z, 2))
  alp = 0.1 * rng.random((Z, 2))
  tt = 6
  wave = rng.standard_normal((tt, 2))
  sf, pos = {0: (rng.standard_normal((2, 1, Y, Z)), 2),
             1: (rng.standard_normal((2, X, 1, Z)), 4),
             2: (rng.standard_normal((2, 2, X, Y, 1)), 4)}[axis]
  sim = S.Simulation(eps, 0.5, sf, wave, pos, mask, kap, sig, alp, widths, offset=(1, 1, 1))
  for _ in range(tt):
    sim.step()
  Eb, Hb = brute.run(sim.eps_full, mask, 0.5, kap, sig, alp, widths, axis, pos, sf, wave, tt)
  np.testing.assert_allclose(np.stack(sim.E), Eb, rtol=0, atol=1e-13)
  np.testing.assert_allclose(np.stack(sim.H), Hb, rtol=0, atol=1e-13)


def test_cavity_mode_closed_form():
  X, Y, Z, dt = 9, 7, 8, 0.5
  E0, lam = _cavity_mode(X, Y, Z)
  sim = _sim(_vacuum(X, Y, Z), dt=dt, pos=3, E0=E0)
  theta = math.acos(1 - dt * dt * lam / 2)
  for n in range(1, 60):
    sim.step()
    a = math.cos(n * theta + theta / 2) / math.cos(theta / 2)
    np.testing.assert_allclose(np.stack(sim.E), a * np.stack(E0), rtol=0, atol=1e-12)


def test_energy_invariant():
  X, Y, Z = 7, 6, 8
  rng = np.random.default_rng(1)
  d = _vacuum(X, Y, Z)
  d["eps"] = 1 + 5 * rng.random((3, X, Y, Z))
  d["kap"] = np.full((Z, 2), 1.7)  # uniform kappa keeps curl_b = curl_f^T
  E0 = [rng.standard_normal((X, Y, Z)) for _ in range(3)]
  H0 = [rng.standard_normal((X, Y, Z)) for _ in range(3)]
  sim = _sim(d, pos=3, E0=E0, H0=H0)
  prev = [e.copy() for e in sim.E]
  sim.step()
  W = []
  for _ in range(100):
    W.append(S.energy(sim, prev))
    prev = [e.copy() for e in sim.E]
    sim.step()
  W = np.array(W)
  assert np.abs(W - W[0]).max() <= 1e-12 * abs(W[0])


def test_energy_invariant_fails_for_planted_sign_error(monkeypatch):
  """The energy check has teeth: a flipped d/dz in the H update breaks it."""
  X, Y, Z = 6, 5, 7
  rng = np.random.default_rng(2)
  d = _vacuum(X, Y, Z)
  E0 = [rng.standard_normal((X, Y, Z)) for _ in range(3)]
  real = S.dfwd

  def bad(xp, f, axis):
    out = real(xp, f, axis)
    return -out if axis == 2 else out
  monkeypatch.setattr(S, "dfwd", bad)
  sim = _sim(d, pos=3, E0=E0)
  prev = [e.copy() for e in sim.E]
  sim.step()
  W = []
  for _ in range(40):
    W.append(S.energy(sim, prev))
    prev = [e.copy() for e in sim.E]
    sim.step()
  assert np.abs(np.array(W) - W[0]).max() > 1e-3 * abs(W[0])


def _mode_matrix(sim_factory, E0, X, Y, Z):
  """2x2 map (a, h) -> (a', h') of one step restricted to a cavity mode."""
  e = np.stack(E0)
  Ce = None
  cols = []
  for a, h in ((1.0, 0.0), (0.0, 1.0)):
    sim = sim_factory()
    if Ce is None:
      probe = sim_factory(E0=E0)
      probe.step()
      # H after one step from (E0, 0) is -dt * C e
      Ce = -np.stack(probe.H) / probe.dt
    sim.E = [a * x for x in E0]
    sim.H = [h * Ce[c] for c in range(3)]
    sim.step()
    an = float((np.stack(sim.E) * e).sum() / (e * e).sum())
    hn = float((np.stack(sim.H) * Ce).sum() / (Ce * Ce).sum())
    cols.append([an, hn])
  return np.array(cols).T


def test_absorber_damping_law():
  X, Y, Z, dt, sigma, eps_r = 9, 7, 8, 0.5, 0.02, 2.0
  E0, lam = _cavity_mode(X, Y, Z)
  d = _vacuum(X, Y, Z)
  d["eps"][:] = eps_r
  d["mask"][:] = sigma
  M = _mode_matrix(lambda **kw: _sim(d, dt=dt, pos=3, **kw), E0, X, Y, Z)
  # energy of the mode decays by exp(-sigma dt / eps) per step
  assert np.linalg.det(M) == pytest.approx(math.exp(-sigma * dt / eps_r), rel=1e-6)
  d["mask"][:] = 0
  M0 = _mode_matrix(lambda **kw: _sim(d, dt=dt, pos=3, **kw), E0, X, Y, Z)
  assert np.linalg.det(M0) == pytest.approx(1.0, abs=1e-12)
  # lossless trace: 2 cos(w dt) with the discrete dispersion of eps_r
  assert np.trace(M0) == pytest.approx(2 - dt * dt * lam / eps_r, abs=1e-12)


def test_kappa_inf_decouples_z_slices():
  X, Y, Z, dt = 8, 7, 5, 0.5
  cx, _, kx = _mode_1d(X, 1)
  cy, _, ky = _mode_1d(Y, 0)
  lam = kx ** 2 + ky ** 2
  amp = np.array([0.3, -1.0, 2.0, 0.5, 1.5])
  Ez = cx[:, None, None] * cy[None, :, None] * amp[None, None, :]
  d = _vacuum(X, Y, Z)
  d["kap"][:] = np.inf
  zeros = np.zeros((X, Y, Z))
  sim = _sim(d, dt=dt, pos=2, widths=(2, 1), E0=[zeros, zeros, Ez])
  theta = math.acos(1 - dt * dt * lam / 2)
  for n in range(1, 40):
    sim.step()
    a = math.cos(n * theta + theta / 2) / math.cos(theta / 2)
    np.testing.assert_allclose(sim.E[2], a * Ez, rtol=0, atol=1e-12)
    assert np.abs(sim.E[0]).max() < 1e-12 and np.abs(sim.E[1]).max() < 1e-12


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_first_step_is_cb_J(axis):
  rng = np.random.default_rng(10 + axis)
  X, Y, Z, dt = 6, 7, 8, 0.5
  eps = 1 + rng.random((3, X, Y, Z))
  mask = 0.2 * rng.random((3, X, Y))
  wave = rng.standard_normal((3, 2))
  sf, pos = {0: (rng.standard_normal((2, 1, Y, Z)), 4),
             1: (rng.standard_normal((2, X, 1, Z)), 2),
             2: (rng.standard_normal((2, 2, X, Y, 1)), 5)}[axis]
  out = S.fdtdz(eps, dt, sf, wave, pos, mask, np.ones((Z, 2)), np.zeros((Z, 2)),
                np.zeros((Z, 2)), (0, 0), (0, 1, 1))
  cb = (dt / eps) / (1 + mask[..., None] * dt / (2 * eps))
  J = np.zeros((3, X, Y, Z))
  if axis == 0:
    J[1, pos], J[2, pos] = sf[0, 0] * wave[0, 0], sf[1, 0] * wave[0, 0]
    J[1, pos - 1], J[2, pos - 1] = sf[0, 0] * wave[0, 1], sf[1, 0] * wave[0, 1]
  elif axis == 1:
    J[0, :, pos], J[2, :, pos] = sf[0, :, 0] * wave[0, 0], sf[1, :, 0] * wave[0, 0]
    J[0, :, pos - 1], J[2, :, pos - 1] = sf[0, :, 0] * wave[0, 1], sf[1, :, 0] * wave[0, 1]
  else:
    for c in range(2):
      J[c, :, :, pos] = sf[0, c, :, :, 0] * wave[0, 0] + sf[1, c, :, :, 0] * wave[0, 1]
  np.testing.assert_allclose(out[0], cb * J, rtol=0, atol=1e-14)


def test_edge_fill_equals_explicit_padding():
  rng = np.random.default_rng(5)
  X, Y, Z = 10, 9, 12
  eps = 1 + 4 * rng.random((3, 4, 3, 6))
  off = (3, 2, 4)
  pads = [(0, 0)] + [(off[d], (X, Y, Z)[d] - off[d] - eps.shape[d + 1]) for d in range(3)]
  eps_full = np.pad(eps, pads, mode="edge")
  sf = rng.standard_normal((2, 2, X, Y, 1))
  wave = rng.standard_normal((12, 2))
  args = (0.5, sf, wave, 6, 0.1 * rng.random((3, X, Y)), np.ones((Z, 2)),
          H.pml_sigma((2, 2), Z, 0.5, 3.0), np.full((Z, 2), 0.05), (2, 2))
  a = S.fdtdz(eps, *args, (5, 12, 3), offset=off)
  b = S.fdtdz(eps_full, *args, (5, 12, 3))[:, :, off[0]:off[0] + 4, off[1]:off[1] + 3,
                                           off[2]:off[2] + 6]
  np.testing.assert_allclose(a, b, rtol=0, atol=1e-14)


def test_linearity_and_zero_source():
  rng = np.random.default_rng(6)
  X, Y, Z = 8, 8, 10
  eps = 1 + rng.random((3, X, Y, Z))
  common = dict(dt=0.5, source_position=4, absorption_mask=0.05 * rng.random((3, X, Y)),
                pml_kappa=np.ones((Z, 2)), pml_sigma=H.pml_sigma((3, 3), Z, 0.5, 3.0),
                pml_alpha=np.full((Z, 2), 0.05), pml_widths=(3, 3), output_steps=(10, 20, 4))
  s1, s2 = rng.standard_normal((2, 1, Y, Z)), rng.standard_normal((2, 1, Y, Z))
  w1, w2 = rng.standard_normal((20, 2)), rng.standard_normal((20, 2))
  f = lambda s, w: S.fdtdz(eps, source_field=s, source_waveform=w, **common)
  np.testing.assert_allclose(f(2 * s1 - 3 * s2, w1), 2 * f(s1, w1) - 3 * f(s2, w1),
                             atol=1e-12)
  np.testing.assert_allclose(f(s1, w1 + w2), f(s1, w1) + f(s1, w2), atol=1e-12)
  assert np.abs(f(s1, 0 * w1)).max() == 0.0


def _zpulse(Z, zsrc, widths, steps, sigma_max=0.5):
  X = Y = 20
  xs = np.arange(X) - X / 2
  prof = np.exp(-(xs[:, None] ** 2 + xs[None, :] ** 2) / 8.0)
  sf = np.zeros((2, 2, X, Y, 1))
  sf[0, 0, :, :, 0] = prof
  t = np.arange(steps)
  wave = np.zeros((steps, 2))
  wave[:, 0] = np.sin(2 * np.pi / 12 * 0.5 * t) * np.exp(-((t - 40) / 12.0) ** 2)
  lo, hi = widths
  sig = H.pml_sigma(widths, Z, 1.0, 3.0) * (sigma_max / 2.0) if lo + hi else np.zeros((Z, 2))
  eps = np.ones((3, X, Y, 16))
  off = (0, 0, zsrc - 8)
  return S.fdtdz(eps, 0.5, sf, wave, zsrc, H.absorption_mask(X, Y, 4, 2e-2),
                 np.ones((Z, 2)), sig, np.full((Z, 2), 0.05), widths,
                 (steps - 1, steps, 1), offset=off)


def test_cpml_reflection_bound():
  steps = 150
  ref = _zpulse(300, 150, (0, 0), steps)       # walls too far to reach the window
  pml = _zpulse(40, 20, (10, 10), steps)       # PML 12 cells from the window
  bare = _zpulse(40, 20, (0, 0), steps)        # same box, no PML
  err = np.linalg.norm(pml - ref) / np.linalg.norm(ref)
  err_bare = np.linalg.norm(bare - ref) / np.linalg.norm(ref)
  assert err_bare > 0.3          # the walls do reflect within the window/time
  assert err < 0.02 * err_bare   # the CPML removes nearly all of it


def _point_response(eps, mask, src, comp_src, rec, comp_rec, wave, dbwd=None, monkeypatch=None):
  """E_comp_rec at `rec` over every step, driven by a unit current in
  component comp_src (x or y) at `src`, through a z-plane source."""
  X, Y, Z = eps.shape[1:]
  sf = np.zeros((2, 2, X, Y, 1))
  sf[0, comp_src, src[0], src[1], 0] = 1.0
  tt = wave.shape[0]
  out = S.fdtdz(eps, 0.5, sf, wave, src[2], mask, np.ones((Z, 2)), np.zeros((Z, 2)),
                np.zeros((Z, 2)), (0, 0), (0, tt, 1))
  return out[:, comp_rec, rec[0], rec[1], rec[2]]


def _reciprocity_pair():
  rng = np.random.default_rng(11)
  X, Y, Z = 9, 8, 10
  eps = 1 + 4 * rng.random((3, X, Y, Z))      # anisotropic, diagonal: reciprocal
  mask = 0.08 * rng.random((3, X, Y))        # conductivity: reciprocal loss
  t = np.arange(60)
  wave = np.zeros((60, 2))
  wave[:, 0] = np.sin(0.4 * t) * np.exp(-((t - 15) / 6.0) ** 2)
  A, B = (2, 3, 4), (6, 5, 7)
  return eps, mask, wave, A, B


def test_reciprocity():
  """Yee's scheme in E is (eps d_t^2 + dt^2 C^T C) E = -dt d_t J with a
  symmetric spatial operator, and the absorber adds a diagonal term: the
  discrete Green's function is symmetric.  So Ey at B driven by Jx at A
  equals Ex at A driven by Jy at B, step by step (no CPML: kappa = 1,
  sigma = 0).  A wrong index or direction in one curl term, a source placed
  off its E sample, or epsilon / absorber taken from a neighbouring cell on
  one side breaks it."""
  eps, mask, wave, A, B = _reciprocity_pair()
  ab = _point_response(eps, mask, A, 0, B, 1, wave)
  ba = _point_response(eps, mask, B, 1, A, 0, wave)
  assert np.abs(ab).max() > 1e-4            # the signal has arrived
  np.testing.assert_allclose(ab, ba, rtol=0, atol=1e-12 * np.abs(ab).max())


def test_reciprocity_fails_for_planted_shift(monkeypatch):
  """The reciprocity check has teeth: a backward x difference taken one
  cell off (no longer the transpose of the forward one) breaks it."""
  eps, mask, wave, A, B = _reciprocity_pair()
  real = S.dbwd

  def bad(xp, f, axis):
    out = real(xp, f, axis)
    return np.roll(out, 1, axis=0) if axis == 0 else out
  monkeypatch.setattr(S, "dbwd", bad)
  ab = _point_response(eps, mask, A, 0, B, 1, wave)
  ba = _point_response(eps, mask, B, 1, A, 0, wave)
  assert np.abs(ab - ba).max() > 1e-3 * np.abs(ab).max()


def test_time_shift_invariance():
  """Delaying the waveform by k rows delays every snapshot by k steps, bit
  for bit, with CPML, absorber and an x-plane source: no coefficient depends
  on the step index."""
  rng = np.random.default_rng(12)
  X, Y, Z, tt, k = 10, 8, 12, 40, 7
  eps = 1 + 3 * rng.random((3, X, Y, Z))
  common = dict(dt=0.5, source_field=rng.standard_normal((2, 1, Y, Z)), source_position=4,
                absorption_mask=H.absorption_mask(X, Y, 3, 2e-2),
                pml_kappa=1 + rng.random((Z, 2)), pml_sigma=H.pml_sigma((3, 3), Z, 0.5, 3.0),
                pml_alpha=np.full((Z, 2), 0.05), pml_widths=(3, 3))
  wave = rng.standard_normal((tt, 2))
  late = np.concatenate([np.zeros((k, 2)), wave])
  a = S.fdtdz(eps, source_waveform=wave, output_steps=(10, tt, 3), **common)
  b = S.fdtdz(eps, source_waveform=late, output_steps=(10 + k, tt + k, 3), **common)
  assert np.abs(a).max() > 0
  np.testing.assert_array_equal(a, b)
