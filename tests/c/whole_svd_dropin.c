/* whole_svd_dropin.c -- a C caller of the drop-in svd_gpu, built the way the
 * reference builds its own driver (gcc -std=gnu99, Makefile:14-17) but linked
 * against libddcsvd instead of the OpenCL objects.  It follows the driver
 * test-whole-svd.c:29-99: the same input draw, the same call (:75), and it
 * runs the reconstruction check that the reference driver has compiled out
 * (:81-96, "#if 0"), as assertions.
 *
 *   whole_svd_dropin M N rand OUT      A = U[1,4) from glibc rand(), never
 *                                      seeded (test-whole-svd.c:18-24, :69-73)
 *   whole_svd_dropin M N file:PATH OUT A read from PATH (M*N doubles,
 *                                      column-major)
 *
 * Checks (exit status 1 on failure): sigma finite and ascending; residual
 * ||A - U S V^T||_F / ||A||_F <= 1e-12; ||U^T U - I||_F and ||V^T V - I||_F
 * <= 2e-13 * min(M, N); for M >= N, A is left holding unit-norm reflectors
 * (columns A[i:M, i], rows A[i, i+1:N] for i < N-2) as bidiag_par leaves it
 * (bidiag_par.c:415-418); a zero column or row tail is stored as a zero
 * vector ("no reflection", where the reference divides 0 by 0).  OUT
 * receives sigma then the overwritten A, so the calling test can compare
 * them with the golden fixtures. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "svd_gpu.h"

static double rand_d(double a, double b) {
  double c = rand() / (1.0 * RAND_MAX + 1); /* in [a, b) */
  return a + c * (b - a);
}

static int fail(const char* what, double v, double bound) {
  fprintf(stderr, "FAIL %s = %.3e (bound %.3e)\n", what, v, bound);
  return 1;
}

int main(int argc, char** argv) {
  if (argc != 5) {
    fprintf(stderr, "usage: %s M N rand|file:PATH OUT\n", argv[0]);
    return 2;
  }
  const int m = atoi(argv[1]), n = atoi(argv[2]);
  if (m < 1 || n < 1) return 2;
  const int mn = m < n ? m : n;
  double* A = malloc(sizeof(double) * (size_t)m * n);
  double* A0 = malloc(sizeof(double) * (size_t)m * n);
  double* sigma = malloc(sizeof(double) * mn);
  double* U = calloc((size_t)m * m, sizeof(double));
  double* V = calloc((size_t)n * n, sizeof(double));
  if (!A || !A0 || !sigma || !U || !V) return 2;
  if (strcmp(argv[3], "rand") == 0) {
    for (long i = 0; i < (long)m * n; ++i) A[i] = rand_d(1., 4.);
  } else if (strncmp(argv[3], "file:", 5) == 0) {
    FILE* f = fopen(argv[3] + 5, "rb");
    if (!f || fread(A, sizeof(double), (size_t)m * n, f) != (size_t)m * n) return 2;
    fclose(f);
  } else {
    return 2;
  }
  memcpy(A0, A, sizeof(double) * (size_t)m * n);

  svd_gpu(m, n, A, sigma, U, V);

  int bad = 0;
  for (int i = 0; i < mn; ++i) {
    if (!isfinite(sigma[i])) bad |= fail("non-finite sigma", i, 0);
    if (i > 0 && sigma[i] < sigma[i - 1]) bad |= fail("sigma not ascending at", i, 0);
  }
  /* residual: R = A0 - sum_k sigma_k U[:,k] V[:,k]^T */
  double rn = 0.0, an = 0.0;
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < m; ++i) {
      double r = A0[i + (long)j * m];
      for (int k = 0; k < mn; ++k) r -= U[i + (long)k * m] * sigma[k] * V[j + (long)k * n];
      rn += r * r;
      an += A0[i + (long)j * m] * A0[i + (long)j * m];
    }
  const double res = sqrt(rn) / sqrt(an);
  if (!(res <= 1e-12)) bad |= fail("residual", res, 1e-12);
  /* orthogonality of the first mn columns */
  double ou = 0.0, ov = 0.0;
  for (int p = 0; p < mn; ++p)
    for (int q = 0; q < mn; ++q) {
      double du = 0.0, dv = 0.0;
      for (int i = 0; i < m; ++i) du += U[i + (long)p * m] * U[i + (long)q * m];
      for (int i = 0; i < n; ++i) dv += V[i + (long)p * n] * V[i + (long)q * n];
      du -= p == q;
      dv -= p == q;
      ou += du * du;
      ov += dv * dv;
    }
  ou = sqrt(ou);
  ov = sqrt(ov);
  const double ob = 2e-13 * (mn > 8 ? mn : 8);
  if (!(ou <= ob)) bad |= fail("||U^T U - I||", ou, ob);
  if (!(ov <= ob)) bad |= fail("||V^T V - I||", ov, ob);
  /* the overwritten A holds unit-norm reflectors */
  double worst = 0.0;
  if (m >= n) {
    for (int i = 0; i < n; ++i) {
      double c = 0.0;
      for (int r = i; r < m; ++r) c += A[r + (long)i * m] * A[r + (long)i * m];
      worst = fmax(worst, c == 0.0 ? 0.0 : fabs(c - 1.0));
      if (i + 2 < n) {
        double w = 0.0;
        for (int k = i + 1; k < n; ++k) w += A[i + (long)k * m] * A[i + (long)k * m];
        worst = fmax(worst, w == 0.0 ? 0.0 : fabs(w - 1.0));
      }
    }
    if (!(worst <= 1e-12)) bad |= fail("reflector norm deviation", worst, 1e-12);
  }
  FILE* o = fopen(argv[4], "wb");
  if (!o) return 2;
  fwrite(sigma, sizeof(double), mn, o);
  fwrite(A, sizeof(double), (size_t)m * n, o);
  fclose(o);
  printf("svd_gpu %dx%d: residual %.3e, orth U %.3e, orth V %.3e, reflectors %.1e, %s\n", m, n,
         res, ou, ov, worst, bad ? "FAILED" : "ok");
  free(A);
  free(A0);
  free(sigma);
  free(U);
  free(V);
  return bad;
}
