"""field() / SimParams: host glue on CPU, end-to-end parity on the GPU."""

import numpy as np
import pytest
import torch

import cw_checks as CW
from oracle import field_ref
from pjz_amd import _field as Fd

P = Fd.SimParams(omega_range=(2 * np.pi / 22, 2 * np.pi / 18), tt=700, absorption_padding=6,
                 absorption_coeff=4e-3, pml_widths=(6, 6), pml_alpha_coeff=0.05,
                 use_reduced_precision=False)


def problem(axis, seed=0, xx=14, yy=12, zz=20):
  rng = np.random.default_rng(seed)
  eps = (1 + 3 * rng.random((3, xx, yy, zz))).astype(np.float32)
  shape = {"x": (2, 1, yy, zz), "y": (2, xx, 1, zz), "z": (2, xx, yy, 1)}[axis]
  src = rng.standard_normal(shape).astype(np.float32)
  pos = {"x": 5, "y": 4, "z": 9}[axis]
  return eps, src, pos


def test_simparams_defaults_match_reference():
  p = Fd.SimParams(omega_range=(1.0, 2.0), tt=100)
  assert (p.dt, p.source_ramp, p.source_delay, p.absorption_padding, p.absorption_coeff,
          p.pml_widths, p.pml_alpha_coeff, p.pml_sigma_lnr, p.pml_sigma_m, p.use_z_as_batch,
          p.use_reduced_precision, p.launch_params) == (
              0.5, 4.0, 4.0, 50, 1e-4, (16, 16), 0.0, 0.5, 1.3, False, True, None)


@pytest.mark.parametrize("axis", ["x", "y", "z"])
@pytest.mark.parametrize("pos_shift", [0, 1])
@pytest.mark.parametrize("z_batch", [False, True])
def test_simulation_inputs_match_oracle_glue(axis, pos_shift, z_batch):
  eps, src, pos = problem(axis)
  pos += pos_shift
  p = P._replace(use_z_as_batch=z_batch)
  omega = np.array([2 * np.pi / 20, 2 * np.pi / 19])
  kw, steps = Fd.simulation_inputs(torch.as_tensor(eps), torch.as_tensor(src), omega, pos, p)
  ref = field_ref.inputs(eps.astype(np.float64), src, omega, pos, p)
  for k in ("source_position", "pml_widths", "output_steps", "offset", "dt"):
    assert kw[k] == ref[k], k
  for k in ("source_field", "source_waveform", "absorption_mask", "pml_kappa", "pml_sigma",
            "pml_alpha"):
    np.testing.assert_allclose(np.asarray(kw[k]), ref[k], rtol=2e-6, atol=1e-7, err_msg=k)


def test_prop_axis_error():
  with pytest.raises(ValueError):
    Fd.prop_axis((2, 1, 1, 5))


@pytest.mark.gpu
@pytest.mark.parametrize("axis", ["x", "y", "z"])
def test_field_matches_oracle(axis, gpu):
  eps, src, pos = problem(axis, seed=1)
  pos += 1  # odd: exercises the position shift / waveform flip for x, y
  omega = np.array([2 * np.pi / 21, 2 * np.pi / 19])
  got = Fd.field(torch.as_tensor(eps, device=gpu), torch.as_tensor(src, device=gpu), omega,
                 pos, P).cpu().numpy()
  ref = field_ref.field(eps, src, omega, pos, P)
  assert got.shape == ref.shape == (2, 3) + eps.shape[1:]
  err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
  print(f"field {axis}: rel L2 {err:.2e}")
  assert err <= 1e-4


@pytest.mark.gpu
@pytest.mark.parametrize("z_batch", [False, True])
def test_field_reduced_precision_default(z_batch, gpu):
  """pjz's default use_reduced_precision=True gives the larger z budget
  (128 - PML cells, _field.py:56-58); z as batch switches the CPML off."""
  eps, src, pos = problem("x", seed=2)
  p = P._replace(use_reduced_precision=True, use_z_as_batch=z_batch)
  omega = np.array([2 * np.pi / 20])
  got = Fd.field(torch.as_tensor(eps, device=gpu), torch.as_tensor(src, device=gpu), omega,
                 pos, p).cpu().numpy()
  ref = field_ref.field(eps, src, omega, pos, p)
  err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
  print(f"field reduced z_batch={z_batch}: rel L2 {err:.2e}", flush=True)
  assert err <= 1e-4


@pytest.mark.gpu
def test_fields_batch_equals_field_per_source(gpu):
  """fields() (scatter's per-port loop as one batch) gives field() per source."""
  eps, _, _ = problem("x", seed=3)
  rng = np.random.default_rng(4)
  srcs = [torch.as_tensor(rng.standard_normal((2, 1, 12, 20)).astype(np.float32), device=gpu)
          for _ in range(3)]
  positions = [4, 7, 9]
  omega = np.array([2 * np.pi / 21, 2 * np.pi / 19])
  e = torch.as_tensor(eps, device=gpu)
  got = Fd.fields(e, srcs, omega, positions, P)
  for g, s, p in zip(got, srcs, positions):
    ref = Fd.field(e, s, omega, p, P)
    err = float(torch.linalg.vector_norm(g - ref) / torch.linalg.vector_norm(ref))
    print(f"port at {p}: rel {err:.2e}", flush=True)
    assert err <= 1e-6


@pytest.mark.gpu
@pytest.mark.parametrize("axis,pos,cplx", [("x", 7, False), ("y", 6, False), ("z", 5, True)])
def test_field_is_the_cw_steady_state(axis, pos, cplx, gpu):
  """field() on the GPU (float32) satisfies the time-harmonic form of the
  discrete update: zero residual off the source plane, mode / eps with pjz's
  phase and 1/ww on it (derivation in tests/cw_checks.py)."""
  shape = (16, 14, 12)
  eps = 1.5 + 2.0 * np.random.default_rng(0).random((3,) + shape)
  p = P._replace(tt=2500, absorption_padding=8)
  omega = np.array([2 * np.pi / 21, 2 * np.pi / 19])
  mode, J = CW.divergence_free_source(axis, shape, eps, pos, complex_=cplx)
  got = Fd.field(torch.as_tensor(eps, dtype=torch.float32, device=gpu),
                 torch.as_tensor(mode, device=gpu), omega, pos, p).cpu().numpy()
  err = CW.cw_error(got.astype(np.complex128), eps.astype(np.float32).astype(np.float64), mode,
                    J, omega, pos, axis, p)
  print(f"field {axis}@{pos}: CW residual {err:.2e}", flush=True)
  assert err <= 2e-3
