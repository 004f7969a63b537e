"""Continuous-wave checks of field() outputs (helper module of the tests).

A field() result is the complex amplitude of a time-harmonic steady state, so
it must satisfy the time-harmonic form of the discrete update it came from.
For the stepper spec (oracle/fdtd_spec.py) with z = exp(i omega dt) and the
convention that snapshot s is E after the E half step of step s:

  H of step n, H^(n+1/2) = H^(n-1/2) - dt curl_f(E^n), gives
      H_hat (1 - 1/z) = -dt curl_f^s(E_hat)
  psi recursions  psi^n = b psi^(n-1) + c d_z(...)  give stretched z
      derivatives  s(k) d_z  with  s = 1/kappa + c / (1 - b/z)  in PML cells
      (1/kappa elsewhere), column 1 for the H update, column 0 for E
  E^(n+1) = ca E^n + cb (curl_b^s(H^(n+1/2)) + J(n)) gives
      r(E_hat) = (z - ca) E_hat - cb curl_b^s(H_hat) = z cb J_hat
  where J(n) = Re(J_hat z^n), and the factor z on the right is the one step
  between J(n) and the snapshot label s = n that E^(n+1) carries.

field() (src/pjz/_field.py:171-279) drives the kernel with the waveform
ramped_sin -> mean over omega of exp(i omega dt t) (steady part): an x/y
plane takes its imaginary part (J_hat = -i/ww * source), a z plane the two
quadratures [imag, real] against [imag(src), real(src)] (J_hat = conj(src) /
ww), with source = mode / eps on the source plane (_field.py:166-167) at the
requested position.  So away from the source plane r must vanish, and on it
r must equal z cb J_hat.  The residual is evaluated on the eps window
interior (one cell in from every face) where every stencil point is inside
the window.

Sources are built divergence-free on their plane (J = curl of a bump), so no
static charge builds up: a static field would otherwise leak into the
projection (pjz's few snapshots cannot separate it from the harmonic part).
"""

import numpy as np

from oracle import c_oracle, field_ref, fdtd_spec
from oracle import helpers as H

TRANSVERSE = {"x": (1, 2), "y": (0, 2), "z": (0, 1)}


def _dfwd(f, ax):
  return np.concatenate([np.diff(f, axis=ax), -np.take(f, [-1], axis=ax)], axis=ax)


def _dbwd(f, ax):
  return np.concatenate([np.take(f, [0], axis=ax), np.diff(f, axis=ax)], axis=ax)


def stretch(p, ZZ, z):
  """Per-z stretched-derivative factors (s_E, s_H) of the CPML at z = e^(i w dt)."""
  if p.use_z_as_batch:
    kappa = np.full((ZZ, 2), np.inf)
    sigma = np.zeros((ZZ, 2))
    alpha = np.zeros((ZZ, 2))
  else:
    sigma = H.pml_sigma(p.pml_widths, ZZ, p.pml_sigma_lnr, p.pml_sigma_m)
    kappa = np.ones_like(sigma)
    alpha = p.pml_alpha_coeff * np.ones_like(sigma)
  c = fdtd_spec.pml_coefficients(kappa, sigma, alpha, p.dt)
  lo, hi = p.pml_widths
  inpml = np.zeros(ZZ, bool)
  inpml[:lo] = True
  inpml[ZZ - hi:] = True
  s_e = c["ike"] + np.where(inpml, c["ce"] / (1 - c["be"] / z), 0)
  s_h = c["ikh"] + np.where(inpml, c["ch"] / (1 - c["bh"] / z), 0)
  return s_e, s_h


def residual(E, eps, kw, p, omega):
  """(r, cb) of one frequency's (3, xx, yy, ze) complex window field E."""
  dt = p.dt
  z = np.exp(1j * omega * dt)
  ZZ = np.shape(kw["pml_sigma"])[0]
  x0, y0, z0 = kw["offset"]
  xx, yy, ze = E.shape[1:]
  s_e, s_h = stretch(p, ZZ, z)
  s_e, s_h = s_e[z0:z0 + ze], s_h[z0:z0 + ze]
  Ex, Ey, Ez = E
  cf = [_dfwd(Ez, 1) - s_h * _dfwd(Ey, 2), s_h * _dfwd(Ex, 2) - _dfwd(Ez, 0),
        _dfwd(Ey, 0) - _dfwd(Ex, 1)]
  Hx, Hy, Hz = [-dt * c / (1 - 1 / z) for c in cf]
  cb_ = np.array([_dbwd(Hz, 1) - s_e * _dbwd(Hy, 2), s_e * _dbwd(Hx, 2) - _dbwd(Hz, 0),
                  _dbwd(Hy, 0) - _dbwd(Hx, 1)])
  mask = np.asarray(kw["absorption_mask"], np.float64)[:, x0:x0 + xx, y0:y0 + yy, None]
  q = mask * dt / (2 * eps)
  cb = (dt / eps) / (1 + q)
  ca = (1 - q) / (1 + q)
  return (z - ca) * E - cb * cb_, cb


def divergence_free_source(axis, shape, eps, pos, seed=0, complex_=False, only_normal=False):
  """A mode argument for field() whose current mode/eps is divergence-free.

  ``shape`` is the (xx, yy, ze) window; the source plane is normal to
  ``axis`` at ``pos``.  J's two transverse components are (d_b v psi,
  -d_b u psi) for a Gaussian bump psi on the plane (u, v the plane's axes), so
  d_b u J_u + d_b v J_v = 0.  ``only_normal`` keeps only the z component
  (a z-as-batch run has no z coupling, so only Ez is divergence-free there).
  Returns (mode, J) with mode of field()'s source shape and J (2, plane).
  """
  rng = np.random.default_rng(seed)
  a = "xyz".index(axis)
  plane_axes = [d for d in range(3) if d != a]
  n_u, n_v = shape[plane_axes[0]], shape[plane_axes[1]]

  def bump():
    cu, cv = rng.uniform(0.35, 0.65) * n_u, rng.uniform(0.35, 0.65) * n_v
    uu, vv = np.meshgrid(np.arange(n_u), np.arange(n_v), indexing="ij")
    b = np.exp(-((uu - cu) ** 2 / (0.04 * n_u ** 2) + (vv - cv) ** 2 / (0.04 * n_v ** 2)))
    b[0, :] = b[-1, :] = 0
    b[:, 0] = b[:, -1] = 0
    return b

  def curl(b):
    ju = b - np.roll(b, 1, axis=1)     # d_b v psi
    jv = -(b - np.roll(b, 1, axis=0))  # -d_b u psi
    if only_normal:
      ju = np.zeros_like(ju)
    return np.stack([ju, jv])

  J = curl(bump())
  if complex_:
    J = J + 1j * curl(bump())
  comps = list(TRANSVERSE[axis])
  sl = [slice(None)] * 3
  sl[a] = slice(pos, pos + 1)
  eps_plane = eps[tuple([comps] + sl)]
  Jx = np.expand_dims(J, a + 1)
  return Jx * eps_plane, J


def expected_residual(axis, J, cb, omega, pos, dt, ww):
  """z cb J_hat on the source plane, zero elsewhere ((3, xx, yy, ze))."""
  z = np.exp(1j * omega * dt)
  a = "xyz".index(axis)
  J_hat = np.conj(J) / ww if axis == "z" else -1j * J / ww
  out = np.zeros(cb.shape, np.complex128)
  for i, comp in enumerate(TRANSVERSE[axis]):
    sl = [comp] + [slice(None)] * 3
    sl[1 + a] = pos
    out[tuple(sl)] = z * cb[tuple(sl)] * J_hat[i]
  return out


def interior(a):
  return a[:, 1:-1, 1:-1, 1:-1]


def cw_error(fields, eps, mode, J, omega, pos, axis, p, kw=None):
  """max over frequencies of |r - expected| / |expected| on the interior."""
  kw = kw if kw is not None else field_ref.inputs(eps, mode, omega, pos, p)
  worst = 0.0
  for k, w in enumerate(np.atleast_1d(omega)):
    r, cb = residual(fields[k], eps, kw, p, w)
    e = expected_residual(axis, J, cb, w, pos, p.dt, len(np.atleast_1d(omega)))
    worst = max(worst, np.linalg.norm(interior(r - e)) / np.linalg.norm(interior(e)))
  return worst


def oracle_fields(kw, omega, dt):
  """field_ref.field's last two lines, for inputs that a test has altered."""
  out = c_oracle.fdtdz(**kw)
  return H.frequency_components(out, np.arange(*kw["output_steps"]), omega, dt)
