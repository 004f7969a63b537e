"""Permittivity on the Yee cell from a layered stack: pjz's ``epsilon()``.

Mirrors ``src/pjz/_epsilon.py``: each layer is a 2-D profile magnified
``2 m`` times in x and y; a component's samples are shifted by half a cell
(``m`` sub-pixels, edge-replicated) along the in-plane axes it is not
directed along (``:12-16``), so Ex sits at x + 1/2, Ey at y + 1/2 and Ez at
the cell corner in xy and at z + 1/2 (the second grid column, ``:19-24``).
Within each 2m x 2m tile the layer's mean, mean of the inverse and first
moments (gradient estimates, ``:30-42``) are taken; along z each layer's
share of a cell comes from clipping its interface interval to the cell
(``:55-61``).  The result is Farjadpour et al.'s diagonal sub-pixel
average: ``1 / (p aoi + (1 - p) / avg)`` with ``p`` the squared share of the
component's axis in the gradient (``:72-99``), or the plain volume average.

A float32 stack on the GPU runs the HIP kernels of
``csrc/kernels_render.hip`` (forward and exact reverse pass for the layers
and the interface positions); a CPU stack runs the torch restatement below
on the host.  Both are differentiable with respect to the layers and the
interface positions.
"""

import torch


def _as_tensor(x, ref):
  return torch.as_tensor(x, dtype=ref.dtype, device=ref.device)


def _shift_in_plane(layers, m, dim):
  """Drop the last ``m`` sub-pixels along ``dim`` and edge-pad ``m`` at the
  front: the half-cell offset of ``_epsilon.py:12-16``."""
  body = torch.narrow(layers, dim, 0, layers.shape[dim] - m)
  front = torch.narrow(layers, dim, 0, 1).expand(
      *[m if d == dim else -1 for d in range(layers.ndim)])
  return torch.cat([front, body], dim=dim)


def _render_single(layers, layer_pos, grid_start, grid_end, m, axis, use_simple_averaging):
  """One component (``axis`` in "xyz"); ``_epsilon.py:10-99``."""
  if axis != "x":
    layers = _shift_in_plane(layers, m, 1)
  if axis != "y":
    layers = _shift_in_plane(layers, m, 2)
  col = 1 if axis == "z" else 0
  gs, ge = grid_start[:, col], grid_end[:, col]

  ll, sx, sy = layers.shape
  tiles = layers.reshape(ll, sx // (2 * m), 2 * m, sy // (2 * m), 2 * m)

  # First moment of each sub-pixel row / column about the tile centre; the
  # 12 (2m) factor turns it into a slope for a unit cell.
  w = (torch.arange(2 * m, dtype=layers.dtype, device=layers.device) - (m - 0.5)) / (2 * m) ** 2
  gx = torch.mean(12 * (2 * m) * tiles * w[:, None, None], dim=(2, 4))
  gy = torch.mean(12 * (2 * m) * tiles * w, dim=(2, 4))
  avg = torch.mean(tiles, dim=(2, 4))
  aoi = torch.mean(1 / tiles, dim=(2, 4))

  inf = torch.full((1,), float("inf"), dtype=layers.dtype, device=layers.device)
  lo = torch.cat([-inf, layer_pos])[:, None]
  hi = torch.cat([layer_pos, inf])[:, None]
  p0 = torch.minimum(torch.maximum(lo, gs), ge)
  p1 = torch.minimum(torch.maximum(hi, gs), ge)
  u = (p1 - p0) / (ge - gs)  # (ll, zz): share of each cell held by each layer

  def cross(a, b):
    return torch.einsum("lxy,lz->xyz", a, b)

  if use_simple_averaging:
    return cross(avg, u)

  zc = (p0 + p1) / 2 - (gs + ge) / 2
  aoi = cross(aoi, u)
  ioa = 1 / cross(avg, u)
  grads = [cross(gx, u), cross(gy, u), cross(avg, u * zc) / ((ge - gs) ** 2 / 12)]
  total = sum(g ** 2 for g in grads)
  pii = grads["xyz".index(axis)] ** 2 / torch.where(total == 0, torch.ones_like(total), total)
  return 1 / (pii * aoi + (1 - pii) * ioa)


class _RenderKernel(torch.autograd.Function):
  """render() through the C ABI (``pjz_fdtd_render_forward`` / ``_backward``)
  for a float32 stack on the GPU."""

  @staticmethod
  def forward(ctx, layers, pos, gs, ge, m, simple):
    from . import _lib
    layers, pos = layers.detach().contiguous(), pos.detach().contiguous()
    gs, ge = gs.detach().contiguous(), ge.detach().contiguous()
    ll, sx, sy = layers.shape
    xx, yy, zz = sx // (2 * m), sy // (2 * m), gs.shape[0]
    eps = torch.empty((3, xx, yy, zz), dtype=layers.dtype, device=layers.device)
    stats = torch.empty(12 * ll * xx * yy + 4 * ll * zz, dtype=layers.dtype, device=layers.device)
    with torch.cuda.device(layers.device):
      stream = torch.cuda.current_stream(layers.device).cuda_stream
      _lib.check(_lib.load().pjz_fdtd_render_forward(
          ll, m, xx, yy, zz, int(simple), layers.data_ptr(), pos.data_ptr() if ll > 1 else None,
          gs.data_ptr(), ge.data_ptr(), eps.data_ptr(), stats.data_ptr(), stream),
          "pjz_fdtd_render_forward")
    ctx.save_for_backward(layers, pos, gs, ge, stats)
    ctx.dims = (ll, m, xx, yy, zz, int(simple))
    return eps

  @staticmethod
  def backward(ctx, g_eps):
    from . import _lib
    import ctypes
    layers, pos, gs, ge, stats = ctx.saved_tensors
    ll, m, xx, yy, zz, simple = ctx.dims
    lib = _lib.load()
    nbytes = ctypes.c_size_t(0)
    _lib.check(lib.pjz_fdtd_render_workspace_bytes(ll, xx, yy, zz, ctypes.byref(nbytes)),
               "pjz_fdtd_render_workspace_bytes")
    # torch's caching allocator hands out 512-byte aligned blocks.
    ws = torch.empty(max(nbytes.value, 1), dtype=torch.uint8, device=layers.device)
    g_eps = g_eps.to(layers.dtype).contiguous()
    g_layers = torch.empty_like(layers)
    g_pos = torch.zeros_like(pos)
    with torch.cuda.device(layers.device):
      stream = torch.cuda.current_stream(layers.device).cuda_stream
      _lib.check(lib.pjz_fdtd_render_backward(
          ll, m, xx, yy, zz, simple, layers.data_ptr(), pos.data_ptr() if ll > 1 else None,
          gs.data_ptr(), ge.data_ptr(), stats.data_ptr(), g_eps.data_ptr(), g_layers.data_ptr(),
          g_pos.data_ptr() if ll > 1 else None, ws.data_ptr(), nbytes.value, stream),
          "pjz_fdtd_render_backward")
    return g_layers, g_pos, None, None, None, None


def render(layers, interface_positions, grid_start, grid_end, m, use_simple_averaging=False):
  """(3, xx, yy, zz) permittivity for explicit z grids (``_epsilon.py:102-108``).

  ``layers`` is (ll, 2 m xx, 2 m yy); ``interface_positions`` (ll - 1,);
  ``grid_start`` / ``grid_end`` are (zz, 2): the lower and upper z bounds of
  each cell for the integer (column 0) and half-integer (column 1) grids.
  """
  layers = torch.as_tensor(layers)
  if not layers.is_floating_point():
    layers = layers.to(torch.float32)
  pos = _as_tensor(interface_positions, layers).reshape(-1)
  gs, ge = _as_tensor(grid_start, layers), _as_tensor(grid_end, layers)
  if layers.ndim != 3 or layers.shape[1] % (2 * m) or layers.shape[2] % (2 * m):
    raise ValueError(f"layers must be (ll, 2*m*xx, 2*m*yy) with m={m}, got {tuple(layers.shape)}")
  if pos.shape[0] != layers.shape[0] - 1:
    raise ValueError(f"{layers.shape[0]} layers need {layers.shape[0] - 1} interface positions, "
                     f"got {pos.shape[0]}")
  if layers.is_cuda:
    if layers.dtype != torch.float32:
      raise ValueError(f"render() on the GPU computes in float32, got {layers.dtype}")
    if gs.requires_grad or ge.requires_grad:
      raise ValueError("render() on the GPU differentiates the layers and interface positions, "
                       "not the z grids")
    return _RenderKernel.apply(layers, pos, gs, ge, int(m), bool(use_simple_averaging))
  return torch.stack([_render_single(layers, pos, gs, ge, m, a, use_simple_averaging)
                      for a in "xyz"])


def epsilon(layers, interface_positions, magnification: int, zz: int,
            use_simple_averaging: bool = False):
  """pjz's ``epsilon()`` (``_epsilon.py:111-155``): cells of unit height at
  z = 0 .. zz - 1, the integer grid spanning [k - 1/2, k + 1/2] and the
  half-integer grid [k, k + 1]."""
  layers = torch.as_tensor(layers)
  if not layers.is_floating_point():
    layers = layers.to(torch.float32)
  k = torch.arange(zz, dtype=layers.dtype, device=layers.device)[:, None]
  gs = k + torch.tensor([[-0.5, 0.0]], dtype=layers.dtype, device=layers.device)
  ge = k + torch.tensor([[0.5, 1.0]], dtype=layers.dtype, device=layers.device)
  return render(layers, interface_positions, gs, ge, magnification, use_simple_averaging)
