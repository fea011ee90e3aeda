/* C/OpenMP restatement of src/autojax/numba.py:197-207 (test oracle and CPU baseline only).
 *
 *   W_ij += cos(dg_ij1 * u_k0 + dg_ij0 * u_k1) / n_k^2,   dg = 2 pi (g_i - g_j)
 *
 * Like the reference, the K loop is the parallel loop and every thread keeps a
 * private M x M accumulator (numba's prange array reduction), reduced at the end.
 * Build: see oracle/Makefile.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

int oracle_w_tilde(const double* noise_map_real, const double* uv, const double* grid, int64_t K, int64_t M,
                   int nthreads, double* w) {
  const double two_pi = 2.0 * M_PI;
  if (nthreads <= 0) nthreads = omp_get_max_threads();
  memset(w, 0, sizeof(double) * M * M);
  double* priv = (double*)calloc((size_t)nthreads * M * M, sizeof(double));
  if (!priv) return 1;
#pragma omp parallel num_threads(nthreads)
  {
    double* acc = priv + (size_t)omp_get_thread_num() * M * M;
#pragma omp for schedule(static)
    for (int64_t k = 0; k < K; ++k) {
      const double u0 = uv[2 * k], u1 = uv[2 * k + 1];
      const double inv = 1.0 / (noise_map_real[k] * noise_map_real[k]);
      for (int64_t i = 0; i < M; ++i) {
        const double gi0 = two_pi * grid[2 * i], gi1 = two_pi * grid[2 * i + 1];
        double* row = acc + i * M;
        for (int64_t j = 0; j < M; ++j) {
          const double d0 = gi0 - two_pi * grid[2 * j];
          const double d1 = gi1 - two_pi * grid[2 * j + 1];
          row[j] += cos(d1 * u0 + d0 * u1) * inv;
        }
      }
    }
  }
  for (int t = 0; t < nthreads; ++t)
    for (int64_t e = 0; e < M * M; ++e) w[e] += priv[(size_t)t * M * M + e];
  free(priv);
  return 0;
}

/* Rows rows[0 .. R-1] of the same matrix (out[r * M + j]), for sizes whose full M x M matrix is out of
 * reach on the CPU.  Each element sums its K terms in visibility order, as a row of the loop above
 * would with one thread; the work is split over (row, block of 256 columns) pairs. */
int oracle_w_tilde_rows(const double* noise_map_real, const double* uv, const double* grid, int64_t K, int64_t M,
                        const int64_t* rows, int64_t R, int nthreads, double* out) {
  const double two_pi = 2.0 * M_PI;
  const int64_t JB = 256;
  const int64_t nb = (M + JB - 1) / JB;
  if (nthreads <= 0) nthreads = omp_get_max_threads();
  for (int64_t r = 0; r < R; ++r)
    if (rows[r] < 0 || rows[r] >= M) return 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
  for (int64_t t = 0; t < R * nb; ++t) {
    const int64_t r = t / nb, j0 = (t % nb) * JB, j1 = j0 + JB < M ? j0 + JB : M;
    const int64_t i = rows[r];
    const double gi0 = two_pi * grid[2 * i], gi1 = two_pi * grid[2 * i + 1];
    double acc[256], d0[256], d1[256];
    for (int64_t j = j0; j < j1; ++j) {
      acc[j - j0] = 0.0;
      d0[j - j0] = gi0 - two_pi * grid[2 * j];
      d1[j - j0] = gi1 - two_pi * grid[2 * j + 1];
    }
    for (int64_t k = 0; k < K; ++k) {
      const double u0 = uv[2 * k], u1 = uv[2 * k + 1];
      const double inv = 1.0 / (noise_map_real[k] * noise_map_real[k]);
      for (int64_t j = 0; j < j1 - j0; ++j) acc[j] += cos(d1[j] * u0 + d0[j] * u1) * inv;
    }
    for (int64_t j = j0; j < j1; ++j) out[r * M + j] = acc[j - j0];
  }
  return 0;
}
