// Shared definitions of the gfx950 Yee FDTD library: problem geometry,
// workspace layout, the per-z CPML coefficient table and launch plumbing.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

#include "../../include/pjz_fdtd.h"

namespace pjz {

// Rows of the per-z coefficient table (zt), each `zz` floats long.
//   IKE/BE/CE: 1/kappa, b, c of pml column 0 (E update, z = k)
//   IKH/BH/CH: 1/kappa, b, c of pml column 1 (H update, z = k + 1/2)
enum ZRow { IKE = 0, BE = 1, CE = 2, IKH = 3, BH = 4, CH = 5, ZROWS = 8 };

// psi components, each (xx, yy, np) with np = pml_lo + pml_hi:
//   PSI_HX: d/dz Ey in the Hx update   PSI_HY: d/dz Ex in the Hy update
//   PSI_EX: d/dz Hy in the Ex update   PSI_EY: d/dz Hx in the Ey update
//   PSI_HX2, PSI_HY2: second buffer of the H pair.  The streaming kernel
//   recomputes H on ghost rows/planes owned by other workgroups, so it reads
//   psi_H from one buffer and writes the other (like E and H themselves).
enum PsiComp { PSI_HX = 0, PSI_HY = 1, PSI_EX = 2, PSI_EY = 3, PSI_HX2 = 4, PSI_HY2 = 5,
               PSI_N = 6 };

struct Geom {
  int X, Y, Z;      // domain
  int lo, hi, NP;   // pml widths, NP = lo + hi
  long long N;      // X*Y*Z
};

// Source description handed to the E update.
struct Src {
  int axis, pos, quad;
  const float* field;
  const float* wave;  // (tt, 2)
};

// Output window of the snapshots.
struct OutWin {
  int x0, y0, z0, ex, ey, ez;
  float* out;  // (nn, 3, ex, ey, ez)
};

// Byte offsets of the workspace segments (each 256-byte aligned).
struct Layout {
  size_t E0, H0, E1, H1, cb, psi, zt, total;
};

inline size_t align256(size_t v) { return (v + 255) & ~size_t(255); }

inline Layout make_layout(const Geom& g) {
  Layout L;
  size_t f3 = align256(sizeof(float) * 3 * (size_t)g.N);
  size_t off = 0;
  L.E0 = off; off += f3;
  L.H0 = off; off += f3;
  L.E1 = off; off += f3;
  L.H1 = off; off += f3;
  L.cb = off; off += f3;
  L.psi = off;
  off += align256(sizeof(float) * PSI_N * (size_t)g.X * g.Y * (size_t)(g.NP > 0 ? g.NP : 1));
  L.zt = off; off += align256(sizeof(float) * ZROWS * (size_t)g.Z);
  L.total = off;
  return L;
}

// Launchers implemented in the kernel translation units.
hipError_t launch_prepare(const Geom& g, const pjz_fdtd_problem& p, const float* eps,
                          const float* mask, const float* kappa, const float* sigma,
                          const float* alpha, float* cb, float* zt, hipStream_t s);

// Clears n floats with a fill kernel (no hipMemsetAsync, so the call can sit
// inside a captured graph).
hipError_t launch_zero(float* ptr, size_t n, hipStream_t s);

hipError_t run_reference(const Geom& g, const pjz_fdtd_problem& p, const float* mask,
                         const Src& src, const OutWin& win, char* ws, const Layout& L,
                         int nsteps, hipStream_t s);

hipError_t run_stream(const Geom& g, const pjz_fdtd_problem& p, const float* mask,
                      const Src& src, const OutWin& win, char* ws, const Layout& L,
                      int nsteps, hipStream_t s);

}  // namespace pjz
