/*
 * fdtd_ref.c -- FP64 C + OpenMP restatement of the stepper in fdtd_spec.py.
 * TEST INFRASTRUCTURE ONLY: used by tests/ as a fast checker and by bench.py
 * as the CPU baseline.  Never linked into the product library.
 *
 * Same update, order and conventions as oracle/fdtd_spec.py (which states the
 * equations and is pinned by brute force and closed forms); this file is
 * pinned against fdtd_spec.py in tests/test_oracle_c.py.
 *
 * Arrays are C order, z fastest: fields (3, X, Y, Z); eps_full (3, X, Y, Z),
 * already edge-replicated; mask (3, X, Y); kappa/sigma/alpha (Z, 2);
 * source_field as in fdtdz(); wave (tt, 2); out (nn, 3, ex, ey, ez).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define IDX(i, j, k) (((long)(i) * Y + (j)) * Z + (k))

int oracle_fdtd_run(int X, int Y, int Z, int lo, int hi, double dt, const double* eps_full,
                    const double* mask, const double* kappa, const double* sigma,
                    const double* alpha, int src_axis, int src_pos, int src_quad,
                    const double* src_field, const double* wave, int tt, int out_start,
                    int out_stop, int out_step, int x0, int y0, int z0, int ex, int ey, int ez,
                    double* out, int nsteps_limit, int nthreads, double* E_final,
                    double* H_final) {
  /* Thread count of this call only (num_threads clauses below): setting the
     process-wide default would leave later calls with threads = 0 on the
     count an earlier call asked for. */
  int nt = 1;
#ifdef _OPENMP
  nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#endif
  (void)nt;
  const long N = (long)X * Y * Z;
  const int NP = lo + hi;
  double* E = calloc(3 * N, sizeof(double));
  double* H = calloc(3 * N, sizeof(double));
  double* ca = malloc(3 * N * sizeof(double));
  double* cb = malloc(3 * N * sizeof(double));
  double* psi = calloc(4 * (long)X * Y * (NP > 0 ? NP : 1), sizeof(double));
  double* zt = malloc(6 * Z * sizeof(double));
  if (!E || !H || !ca || !cb || !psi || !zt) return 1;
  for (long c = 0; c < 3; ++c)
    for (long i = 0; i < X; ++i)
      for (long j = 0; j < Y; ++j)
        for (long k = 0; k < Z; ++k) {
          long n = c * N + IDX(i, j, k);
          double e = eps_full[n];
          double q = mask[(c * X + i) * Y + j] * dt / (2.0 * e);
          cb[n] = (dt / e) / (1.0 + q);
          ca[n] = (1.0 - q) / (1.0 + q);
        }
  for (int k = 0; k < Z; ++k)
    for (int col = 0; col < 2; ++col) {
      double kap = kappa[2 * k + col], sig = sigma[2 * k + col], alp = alpha[2 * k + col];
      double ik = isinf(kap) ? 0.0 : 1.0 / kap;
      double b = exp(-(sig * ik + alp) * dt);
      double den = sig * ik + alp;
      double c = (sig == 0.0 || ik == 0.0 || den == 0.0) ? 0.0 : sig * ik * ik * (b - 1.0) / den;
      zt[(3 * col + 0) * Z + k] = ik;
      zt[(3 * col + 1) * Z + k] = b;
      zt[(3 * col + 2) * Z + k] = c;
    }
  const double *ikE = zt, *bE = zt + Z, *cE = zt + 2 * Z;
  const double *ikH = zt + 3 * Z, *bH = zt + 4 * Z, *cH = zt + 5 * Z;
  const long PS = (long)X * Y * (NP > 0 ? NP : 1);
  int nn = out_stop > out_start ? (out_stop - out_start + out_step - 1) / out_step : 0;
  int nsteps = nn > 0 ? out_start + (nn - 1) * out_step + 1 : 0;
  if (nsteps_limit >= 0) nsteps = nsteps_limit;
  (void)tt;
  double *Ex = E, *Ey = E + N, *Ez = E + 2 * N, *Hx = H, *Hy = H + N, *Hz = H + 2 * N;
  for (int n = 0; n < nsteps; ++n) {
#pragma omp parallel for num_threads(nt) schedule(static)
    for (int i = 0; i < X; ++i)
      for (int j = 0; j < Y; ++j)
        for (int k = 0; k < Z; ++k) {
          long c0 = IDX(i, j, k);
          double ey_zp = k + 1 < Z ? Ey[c0 + 1] : 0.0, ex_zp = k + 1 < Z ? Ex[c0 + 1] : 0.0;
          double ez_yp = j + 1 < Y ? Ez[c0 + Z] : 0.0, ex_yp = j + 1 < Y ? Ex[c0 + Z] : 0.0;
          double ez_xp = i + 1 < X ? Ez[c0 + (long)Y * Z] : 0.0;
          double ey_xp = i + 1 < X ? Ey[c0 + (long)Y * Z] : 0.0;
          double dzEy = ey_zp - Ey[c0], dzEx = ex_zp - Ex[c0];
          double tEy = ikH[k] * dzEy, tEx = ikH[k] * dzEx;
          if (k < lo || k >= Z - hi) {
            long p = ((long)i * Y + j) * NP + (k < lo ? k : k - (Z - hi) + lo);
            psi[p] = bH[k] * psi[p] + cH[k] * dzEy;
            psi[PS + p] = bH[k] * psi[PS + p] + cH[k] * dzEx;
            tEy += psi[p];
            tEx += psi[PS + p];
          }
          Hx[c0] = Hx[c0] - dt * ((ez_yp - Ez[c0]) - tEy);
          Hy[c0] = Hy[c0] - dt * (tEx - (ez_xp - Ez[c0]));
          Hz[c0] = Hz[c0] - dt * ((ey_xp - Ey[c0]) - (ex_yp - Ex[c0]));
        }
    int out_idx = -1;
    if (n >= out_start && n < out_stop && (n - out_start) % out_step == 0)
      out_idx = (n - out_start) / out_step;
#pragma omp parallel for num_threads(nt) schedule(static)
    for (int i = 0; i < X; ++i)
      for (int j = 0; j < Y; ++j)
        for (int k = 0; k < Z; ++k) {
          long c0 = IDX(i, j, k);
          double hy_zm = k > 0 ? Hy[c0 - 1] : 0.0, hx_zm = k > 0 ? Hx[c0 - 1] : 0.0;
          double hz_ym = j > 0 ? Hz[c0 - Z] : 0.0, hx_ym = j > 0 ? Hx[c0 - Z] : 0.0;
          double hz_xm = i > 0 ? Hz[c0 - (long)Y * Z] : 0.0;
          double hy_xm = i > 0 ? Hy[c0 - (long)Y * Z] : 0.0;
          double dzHy = Hy[c0] - hy_zm, dzHx = Hx[c0] - hx_zm;
          double tHy = ikE[k] * dzHy, tHx = ikE[k] * dzHx;
          if (k < lo || k >= Z - hi) {
            long p = ((long)i * Y + j) * NP + (k < lo ? k : k - (Z - hi) + lo);
            psi[2 * PS + p] = bE[k] * psi[2 * PS + p] + cE[k] * dzHy;
            psi[3 * PS + p] = bE[k] * psi[3 * PS + p] + cE[k] * dzHx;
            tHy += psi[2 * PS + p];
            tHx += psi[3 * PS + p];
          }
          double cur[3];
          cur[0] = (Hz[c0] - hz_ym) - tHy;
          cur[1] = tHx - (Hz[c0] - hz_xm);
          cur[2] = (Hy[c0] - hy_xm) - (Hx[c0] - hx_ym);
          if (src_axis == 0 && (i == src_pos || i == src_pos - 1)) {
            double w = wave[2 * n + (i == src_pos ? 0 : 1)];
            cur[1] += src_field[(long)j * Z + k] * w;
            cur[2] += src_field[(long)Y * Z + (long)j * Z + k] * w;
          } else if (src_axis == 1 && (j == src_pos || j == src_pos - 1)) {
            double w = wave[2 * n + (j == src_pos ? 0 : 1)];
            cur[0] += src_field[(long)i * Z + k] * w;
            cur[2] += src_field[(long)X * Z + (long)i * Z + k] * w;
          } else if (src_axis == 2 && k == src_pos) {
            long xy = (long)i * Y + j, XY = (long)X * Y;
            for (int q = 0; q < src_quad; ++q) {
              cur[0] += src_field[(2 * q + 0) * XY + xy] * wave[2 * n + q];
              cur[1] += src_field[(2 * q + 1) * XY + xy] * wave[2 * n + q];
            }
          }
          for (int c = 0; c < 3; ++c) {
            long m = c * N + c0;
            E[m] = ca[m] * E[m] + cb[m] * cur[c];
          }
          if (out_idx >= 0 && i >= x0 && i < x0 + ex && j >= y0 && j < y0 + ey && k >= z0 &&
              k < z0 + ez)
            for (int c = 0; c < 3; ++c)
              out[((((long)out_idx * 3 + c) * ex + (i - x0)) * ey + (j - y0)) * ez + (k - z0)] =
                  E[c * N + c0];
        }
  }
  if (E_final) memcpy(E_final, E, 3 * N * sizeof(double));
  if (H_final) memcpy(H_final, H, 3 * N * sizeof(double));
  free(E);
  free(H);
  free(ca);
  free(cb);
  free(psi);
  free(zt);
  return 0;
}
