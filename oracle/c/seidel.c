/* C restatement of seidel_2d_numpy.kernel
 * (npbench/benchmarks/polybench/seidel_2d/seidel_2d_numpy.py:4-13) for grids
 * too large for the NumPy oracle's per-element Python loop.
 *
 * Test infrastructure only: the checker, never the product.  Built by
 * oracle/c/Makefile with -ffp-contract=off and no fast-math, so every
 * operation is one correctly rounded IEEE binary64 operation in the
 * reference's order:
 *   row i:  s_j = ((((((up[j-1] + up[j]) + up[j+1]) + row[j+1]) + dn[j-1])
 *                + dn[j]) + dn[j+1]), all read before row i changes
 *           (the NumPy statement :8-10 evaluates its right side first);
 *           row[j] = row[j] + s_j                          (:8, +=)
 *           row[j] = (row[j] + row[j-1]) / 9.0  for j = 1 .. N-2 in order
 *                                                      (:11-13)
 * Pinned bit for bit against oracle.stencils.seidel_2d_kernel in
 * tests/test_oracle.py. */
#include <stdint.h>
#include <stdlib.h>

int oracle_seidel_2d(int64_t tsteps, int64_t n, double *A) {
  if (n < 3 || tsteps <= 1) return 0;
  double *s = (double *)malloc((size_t)n * sizeof(double));
  if (!s) return -1;
  for (int64_t t = 0; t < tsteps - 1; ++t) {
    for (int64_t i = 1; i < n - 1; ++i) {
      const double *up = A + (i - 1) * n;
      double *row = A + i * n;
      const double *dn = A + (i + 1) * n;
      for (int64_t j = 1; j < n - 1; ++j) {
        double v = up[j - 1] + up[j];
        v = v + up[j + 1];
        v = v + row[j + 1];
        v = v + dn[j - 1];
        v = v + dn[j];
        v = v + dn[j + 1];
        s[j] = v;
      }
      for (int64_t j = 1; j < n - 1; ++j) row[j] = row[j] + s[j];
      for (int64_t j = 1; j < n - 1; ++j) row[j] = (row[j] + row[j - 1]) / 9.0;
    }
  }
  free(s);
  return 0;
}
