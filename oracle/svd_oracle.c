/*
 * svd_oracle.c -- CPU restatement of the ddc-svd host path.
 *
 * TEST INFRASTRUCTURE ONLY (see svd_oracle.h).  It restates the reference's
 * algorithms operation by operation so that it reproduces the reference's
 * results, including its known inaccuracies; the comparison against the
 * compiled reference (oracle/_ref) pins it bit for bit.  The one deliberate
 * deviation is opt-in: fix_multv=1 repairs the non-square indexing of multV.
 *
 * Built with the reference's flags (gcc -std=gnu99 -O3 -fopenmp, no -march),
 * so no fused multiply-adds are formed and the rounding matches.
 */
#include "svd_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

static double g_stage_s[4];

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

void orc_last_timings(double out[4]) {
  for (int i = 0; i < 4; ++i) out[i] = g_stage_s[i];
}

/* ------------------------------------------------------------------------ */
/* Stage 1: serial bidiagonalization                                        */
/* ------------------------------------------------------------------------ */

/* sum of squares then sqrt, in index order: matrix_helper.c:54-61 (column)
 * and :82-89 (row, stride ld). */
static double strided_norm(int len, const double *x, int stride) {
  double acc = 0.0;
  for (int t = 0; t < len; ++t) acc += x[(long)t * stride] * x[(long)t * stride];
  return sqrt(acc);
}

/* Turn the partial vector x (len entries, stride) into the unit-norm
 * reflector in place and return the new diagonal value -sign(x0)*||x||.
 * bidiag.c:76-94 (column) and :101-120 (row).  The reflector length is
 * sqrt(2)*sqrt(nrm^2 + |nrm*x0|); the vector is scaled by its reciprocal. */
static double make_reflector(int len, double *x, int stride) {
  double x0 = x[0];
  int sgn = (x0 < 0) ? -1 : 1;
  double nrm = strided_norm(len, x, stride);
  double diag = -sgn * nrm;
  x[0] += sgn * nrm;
  double rl = sqrt(2.0) * sqrt(nrm * nrm + fabs(nrm * x0));
  double inv = 1.0 / rl;
  for (int t = 0; t < len; ++t) x[(long)t * stride] *= inv;
  return diag;
}

/* Apply I - 2 v v^T from the left to the l x k block at C (leading dim ld).
 * bidiag.c:188-218. */
static void reflect_left(int ld, int l, int k, const double *v, double *C) {
  for (int j = 0; j < k; ++j) {
    double *col = C + (long)j * ld;
    double dp = 0.0;
    for (int i = 0; i < l; ++i) dp += v[i] * col[i];
    for (int i = 0; i < l; ++i) col[i] -= 2 * v[i] * dp;
  }
}

/* Apply I - 2 w w^T from the right to the l x k block at C, where w is a
 * matrix row (stride ld).  bidiag.c:220-250. */
static void reflect_right(int ld, int l, int k, const double *w, double *C) {
  for (int i = 0; i < l; ++i) {
    double dp = 0.0;
    for (int j = 0; j < k; ++j) dp += w[(long)j * ld] * C[i + (long)j * ld];
    for (int j = 0; j < k; ++j) C[i + (long)j * ld] -= 2 * w[(long)j * ld] * dp;
  }
}

/* The first `steps` column/row steps of the main loop only (a bounded
 * sample of the serial bidiagonalization for timing); steps < 0 = all. */
static void bidiag_seq_steps(int m, int n, double *A, double *alpha,
                             double *beta, int steps);

void orc_bidiag_seq(int m, int n, double *A, double *alpha, double *beta) {
  bidiag_seq_steps(m, n, A, alpha, beta, -1);
}

void orc_bidiag_sample(int m, int n, double *A, double *alpha, double *beta,
                       int steps) {
  bidiag_seq_steps(m, n, A, alpha, beta, steps);
}

static void bidiag_seq_steps(int m, int n, double *A, double *alpha,
                             double *beta, int steps) {
  int mn = (m < n) ? m : n;
  long ld = m;
  int last = (steps >= 0 && steps < mn - 1) ? steps : mn - 1;
  for (int i = 0; i < last; ++i) {
    double *aii = A + i + i * ld;
    alpha[i] = make_reflector(m - i, aii, 1);
    reflect_left((int)ld, m - i, n - i - 1, aii, aii + ld);
    double *row = A + i + (i + 1) * ld;
    if (i < n - 2) {
      beta[i] = make_reflector(n - i - 1, row, (int)ld);
      reflect_right((int)ld, m - i - 1, n - i - 1, row, row + 1);
    } else {
      /* one entry left on the right: taken as beta, slot zeroed
       * (bidiag.c:125-129) */
      beta[i] = row[0];
      row[0] = 0;
    }
  }
  if (last < mn - 1) return;
  if (n >= m + 1) {
    /* wide: final row reflector, diagonal taken as is (bidiag.c:133-160) */
    double *row = A + (mn - 1) + mn * ld;
    beta[mn - 1] = make_reflector(n - mn, row, (int)ld);
    double *d = A + (mn - 1) + (mn - 1) * ld;
    alpha[mn - 1] = d[0];
    d[0] = 0;
  } else {
    /* last column reflected, the 1 x 1 case included (bidiag.c:161-183) */
    double *d = A + (mn - 1) + (mn - 1) * ld;
    alpha[mn - 1] = make_reflector(m - mn + 1, d, 1);
  }
}

/* ------------------------------------------------------------------------ */
/* Stage 2: double divide and conquer for the singular values               */
/* ------------------------------------------------------------------------ */

/* Merge the pseudo-sorted pole list (d[0] = 0, then d[1..K] ascending, then
 * d[K+1..N-1] ascending) into ascending order; perm[t] receives the source
 * index of sorted entry t.  Ties take the first list
 * (Calculations-Parallel.c:71-112, :372-457). */
static void merge_order(int K, int N, const double *d, int *perm) {
  int a = 1, b = K + 1, t = 1;
  perm[0] = 0;
  while (a <= K && b < N) perm[t++] = (d[a] <= d[b]) ? a++ : b++;
  while (a <= K) perm[t++] = a++;
  while (b < N) perm[t++] = b++;
}

/* Root i of the secular equation in the reciprocal variable gamma = 1/t,
 * t = sigma^2 - pole_i.  Calculations-Parallel.c:127-343 (Newton with the
 * reference's safeguards, bisection if 30 steps do not settle). */
static double secular_root(int N, int i, const double *w, const double *c,
                           double *sh /* scratch, N */) {
  const double tol = 1e-12;
  const double maxit = 30;
  for (int j = 0; j < N; ++j) sh[j] = w[j] - w[i];
  double prev = 9999.0, last = 0.0, it = 0;
  double gam, g, gp, step;

  if (c[i] < 1e-20) return sqrt(w[i]) + 1e-14;

  if (i < N - 1) {
    double gap = sh[i + 1];
    double qa = -c[i];
    double qb0 = 1.0 + (c[i] + c[i + 1]) / gap;
    double qc = -1.0 / gap;
    double qb = qb0, dqb;
    gam = (-qb - sqrt(qb * qb - 4.0 * qa * qc)) / (2.0 * qa);
    g = 0.0;
    gp = 1.0;
    while (((gp > 0.0) || (fabs(g) > 1e-5) ||
            (fabs(prev - last) / (fabs(prev) + fabs(last)) > tol)) &&
           (it < maxit)) {
      prev = last;
      qb = qb0;
      dqb = 0.0;
      for (int k = 0; k < N; ++k) {
        if (k == i || k == i + 1) continue;
        double den = sh[k] * gam - 1.0;
        qb += c[k] * (gam + qc) / den;
        dqb += -c[k] * (sh[k] * qc + 1.0) / (den * den);
      }
      g = qc + qb * gam + qa * gam * gam;
      gp = 2.0 * qa * gam + qb + gam * dqb;
      step = gam - (g / gp);
      if (step < 1.0 / gap) {
        gam = gam / 2.0 + (1.0 / (2.0 * gap));
      } else if ((g > 0.0) && (gp > 0)) {
        if (it < 5)
          gam = ((gam > 1.0) ? gam : 1.0) * 10.0;
        else
          gam = (-qb - sqrt(qb * qb - 4.0 * qa * qc)) / (2.0 * qa);
      } else if ((g > 0.0) && (gam > 10e24)) {
        prev = last;
        gp = 0.0;
        g = 0.0;
      } else {
        gam = step;
      }
      it++;
      last = gam;
    }
    if (it == maxit) {
      double lo = 1 / gap, hi = -1.0;
      gam = lo;
      while (hi < 0) {
        gam *= 10.0;
        qb = qb0;
        for (int k = 0; k < N; ++k) {
          if (k == i || k == i + 1) continue;
          qb += c[k] * (gam + qc) / (sh[k] * gam - 1.0);
        }
        g = qc + qb * gam + qa * gam * gam;
        if (g < 0.0) hi = gam; else lo = gam;
      }
      gam = (lo + hi) / 2.0;
      while (fabs(lo - hi) / (fabs(lo) + fabs(hi)) > tol) {
        qb = qb0;
        for (int k = 0; k < N; ++k) {
          if (k == i || k == i + 1) continue;
          qb += c[k] * (gam + qc) / (sh[k] * gam - 1.0);
        }
        g = qc + qb * gam + qa * gam * gam;
        if (g > 0.0) lo = gam; else hi = gam;
        gam = (lo + hi) / 2.0;
      }
    }
    return sqrt(w[i] + 1.0 / gam);
  }

  /* largest root: g(gamma) = 1 - c_i gamma + sum_k c_k gamma/(sh_k gamma - 1) */
  gam = (c[i] > 1e-14 ? 1.0 / c[i] : 1);
  g = 0.0;
  gp = 1.0;
  while (((gp > 0) || (fabs(g) > 1e-5) ||
          (fabs(prev - last) / (fabs(prev) + fabs(last)) > tol)) &&
         (it < maxit)) {
    prev = last;
    g = 1 - c[i] * gam;
    gp = -c[i];
    for (int k = 0; k < N - 1; ++k) {
      double den = sh[k] * gam - 1.0;
      g += c[k] * gam / den;
      gp -= c[k] / (den * den);
    }
    step = gam - g / gp;
    if (step < 0.0) {
      gam = gam / 2.0;
    } else if ((g > 0.0) && (gam > 10e24)) {
      prev = last;
      gp = 0.0;
      g = 0.0;
    } else {
      gam = step;
    }
    last = gam;
    it++;
  }
  if (it == maxit) {
    double lo = 0.0, hi = -1.0;
    gam = 0.1;
    while (hi < 0.0) {
      gam *= 10.0;
      g = 1 - c[i] * gam;
      for (int k = 0; k < N - 1; ++k) g += c[k] * gam / (sh[k] * gam - 1);
      if (g < 0.0) hi = gam; else lo = gam;
    }
    gam = (lo + hi) / 2.0;
    while (fabs(lo - hi) / (fabs(lo) + fabs(hi)) > tol) {
      g = 1 - c[i] * gam;
      for (int k = 0; k < N - 1; ++k) g += c[k] * gam / (sh[k] * gam - 1);
      if (g > 0.0) lo = gam; else hi = gam;
      gam = (lo + hi) / 2.0;
    }
  }
  return sqrt(w[i] + 1.0 / gam);
}

/* All N roots of the merged problem, ascending.  Calculations-Parallel.c:48-348. */
static void secular_all(int K, int N, const double *d, const double *z,
                        double *sigma) {
  int *perm = (int *)malloc(sizeof(int) * N);
  double *w = (double *)malloc(sizeof(double) * N);
  double *c = (double *)malloc(sizeof(double) * N);
  merge_order(K, N, d, perm);
  w[0] = 0.0;
  c[0] = z[0] * z[0];
  for (int t = 1; t < N; ++t) {
    w[t] = d[perm[t]] * d[perm[t]];
    c[t] = z[perm[t]] * z[perm[t]];
  }
#pragma omp parallel
  {
    double *sh = (double *)malloc(sizeof(double) * N);
#pragma omp for
    for (int i = 0; i < N; ++i) sigma[i] = secular_root(N, i, w, c, sh);
    free(sh);
  }
  free(perm);
  free(w);
  free(c);
}

/* First and last rows of the merged right singular vectors, from the
 * log-domain Loewner z-hat.  Calculations-Parallel.c:351-588. */
static void edge_rows(int K, int N, const double *d, const double *sigma,
                      const double *z, double *first, double *last, int needs) {
  int need_f = (needs == 1 || needs == 3);
  int need_l = (needs == 2 || needs == 3);
  int *perm = (int *)malloc(sizeof(int) * N);
  double *w = (double *)malloc(sizeof(double) * N);
  double *zo = (double *)malloc(sizeof(double) * N);
  double *hf = (double *)malloc(sizeof(double) * N);
  double *hl = (double *)malloc(sizeof(double) * N);
  double *s2 = (double *)malloc(sizeof(double) * N);
  double *lz = (double *)malloc(sizeof(double) * N);
  double *nv = (double *)malloc(sizeof(double) * N);
  double *nu = (double *)malloc(sizeof(double) * N);
  double *zs = (double *)malloc(sizeof(double) * N);

  merge_order(K, N, d, perm);
  for (int t = 0; t < N; ++t) {
    int p = perm[t];
    w[t] = (t == 0) ? 0.0 : d[p] * d[p];
    zo[t] = z[p];
    if (need_f) hf[t] = first[p];
    if (need_l) hl[t] = last[p];
  }
  for (int t = 0; t < N; ++t) s2[t] = sigma[t] * sigma[t];

#pragma omp parallel for
  for (int i = 0; i < N; ++i) {
    if ((fabs(w[i] - s2[i]) < 1e-14) || (i > 0 && (fabs(w[i] - s2[i - 1]) < 1e-14))) {
      lz[i] = 0.0; /* marks a root that sits on its pole */
    } else {
      double acc = log(s2[N - 1] - w[i]);
      for (int j = 0; j < i; ++j) acc += log(w[i] - s2[j]) - log(w[i] - w[j]);
      for (int j = i; j < N - 1; ++j) acc += log(s2[j] - w[i]) - log(w[j + 1] - w[i]);
      lz[i] = acc / 2;
    }
  }

#pragma omp parallel for
  for (int i = 0; i < N; ++i) {
    if (lz[i] == 0.0) {
      nv[i] = 1.0;
      nu[i] = 1.0;
    } else {
      double sv = 0.0, su = 1.0;
      for (int j = 0; j < N; ++j) {
        double t = exp(lz[j] - log(fabs(w[j] - s2[i])));
        t = t * t;
        sv += t;
        su += t * w[j];
      }
      nv[i] = sqrt(sv);
      nu[i] = sqrt(su);
    }
  }

#pragma omp parallel for
  for (int i = 0; i < N; ++i) {
    if (lz[i] == 0.0) {
      if (fabs(w[i] - s2[i]) < 1e-14)
        zs[i] = (zo[i] > 0 ? 1 : -1);
      else
        zs[i] = (zo[i] > 0 ? -1 : 1);
    } else {
      double acc = 0.0;
      for (int k = 0; k < N; ++k) acc += sigma[k] / nv[k] / nu[k] / (w[i] - s2[k]);
      zs[i] = -1 * (zo[i] > 0 ? 1 : -1) * (acc > 0 ? 1 : -1);
    }
  }

#pragma omp parallel
  {
    double *v = (double *)malloc(sizeof(double) * N);
#pragma omp for
    for (int i = 0; i < N; ++i) {
      if (need_f) first[i] = 0.0;
      if (need_l) last[i] = 0.0;
      for (int j = 0; j < N; ++j) {
        if (lz[i] == 0.0) {
          v[j] = (j == i ? 1.0 : 0.0);
        } else if (lz[j] == 0.0) {
          v[j] = 0.0;
        } else {
          v[j] = exp(lz[j] - log((j > i ? w[j] - s2[i] : s2[i] - w[j])));
          v[j] *= (j > i ? 1 : -1);
        }
        v[j] *= zs[j];
      }
      for (int j = 0; j < N; ++j) {
        v[j] /= nv[i];
        if (need_f) first[i] += v[j] * hf[j];
        if (need_l) last[i] += v[j] * hl[j];
      }
    }
    free(v);
  }
  free(perm); free(w); free(zo); free(hf); free(hl);
  free(s2); free(lz); free(nv); free(nu); free(zs);
}

/* Closed forms for the 1 x 2 and 2 x 3 leaves.  Calculations-Parallel.c:590-703. */
static void dc_leaf(int N, const double *b1, const double *b2, double *sigma,
                    double *first, double *last, double *phi, double *psi,
                    int needs) {
  int need_f = (needs == 1 || needs == 3);
  int need_l = (needs == 2 || needs == 3);
  if (N == 1) {
    sigma[0] = sqrt(b1[0] * b1[0] + b2[0] * b2[0]);
    if (need_f) first[0] = b1[0] / sigma[0];
    if (need_l) last[0] = b2[0] / sigma[0];
    phi[0] = b2[0] / sigma[0];
    psi[0] = -b1[0] / sigma[0];
    return;
  }
  if (N != 2) {
    fprintf(stderr, "oracle dc_leaf: size %d > 2\n", N);
    abort();
  }
  double a = b1[0] * b1[0] + b2[0] * b2[0];
  double dd = b1[1] * b1[1] + b2[1] * b2[1];
  double h0 = (a + dd) / 2.0;
  double h1 = 0.0;
  double h2 = (a - dd) / 2.0;
  double h3 = b2[0] * b1[1];
  double e = h1 * h1 + h2 * h2 + h3 * h3;
  double n1 = h0 * h0 + e;
  double n2 = h0 * h0 - e;
  sigma[0] = sqrt(sqrt(n1 - sqrt(n1 * n1 - n2 * n2)));
  sigma[1] = sqrt(sqrt(n1 + sqrt(n1 * n1 - n2 * n2)));
  /* null vector (first/last components) */
  phi[0] = (b2[1] == 0.0 ? 0.0 : (b1[0] == 0.0 ? 1.0 : -b2[0] / b1[0]));
  psi[0] = (b2[1] == 0.0 ? 1.0 : (b1[0] == 0.0 ? 0.0 : -b1[1] / b2[1]));
  double nrm = (((b2[1] == 0.0) || (b1[0] == 0.0)) ? 1.0
                                                    : sqrt(1 + phi[0] * phi[0] + psi[0] * psi[0]));
  phi[0] /= nrm;
  psi[0] /= nrm;
  for (int r = 0; r < 2; ++r) {
    double v1 = -b2[0] * b1[0] / (b1[0] * b1[0] - sigma[r] * sigma[r]);
    double v3 = -b2[1] * b1[1] / (b2[1] * b2[1] - sigma[r] * sigma[r]);
    nrm = sqrt(1 + v1 * v1 + v3 * v3);
    v1 /= nrm;
    v3 /= nrm;
    if (need_f) first[r] = v1;
    if (need_l) last[r] = v3;
  }
}

/* Recursive merge, splitting row K = N/2.  Calculations-Parallel.c:706-850.
 * The child's singular values land in this level's scratch and the child
 * uses this level's sigma as its scratch. */
static void dc_rec(int N, const double *b1, const double *b2, double *sigma,
                   double *scratch, double *first, double *last, double *phi,
                   double *psi, int needs) {
  if (N <= 2) {
    dc_leaf(N, b1, b2, sigma, first, last, phi, psi, needs);
    return;
  }
  int need_f = (needs == 1 || needs == 3);
  int need_l = (needs == 2 || needs == 3);
  int K = N / 2;
  double phic[2], psic[2];
  double *l1 = (double *)malloc(sizeof(double) * (K > 0 ? K : 1));
  double *f2 = (double *)malloc(sizeof(double) * (N - K - 1 > 0 ? N - K - 1 : 1));
  scratch[0] = 0;
  dc_rec(K, b1, b2, scratch + 1, sigma + 1, need_f ? first + 1 : NULL, l1,
         &phic[0], &psic[0], need_f ? 3 : 2);
  dc_rec(N - K - 1, b1 + K + 1, b2 + K + 1, scratch + K + 1, sigma + K + 1,
         f2, need_l ? last + K + 1 : NULL, &phic[1], &psic[1], need_l ? 3 : 1);

  double r0 = sqrt((b1[K] * psic[0]) * (b1[K] * psic[0]) + (b2[K] * phic[1]) * (b2[K] * phic[1]));
  double c0 = b1[K] * psic[0] / r0;
  double s0 = b2[K] * phic[1] / r0;
  double *z = (double *)malloc(sizeof(double) * N);
  z[0] = r0;
  for (int l = 1; l <= K; ++l) z[l] = b1[K] * l1[l - 1];
  for (int l = K + 1; l < N; ++l) z[l] = b2[K] * f2[l - K - 1];

  secular_all(K, N, scratch, z, sigma);

  if (need_f) {
    first[0] = c0 * phic[0];
    for (int l = K + 1; l < N; ++l) first[l] = 0.0;
  }
  if (need_l) {
    last[0] = s0 * psic[1];
    for (int l = 1; l <= K; ++l) last[l] = 0.0;
  }
  if (needs != 0) {
    edge_rows(K, N, scratch, sigma, z, first, last, needs);
    phi[0] = -s0 * phic[0];
    psi[0] = c0 * psic[1];
  }
  if (need_f) {
    double acc = 0.0;
    for (int l = 0; l < N; ++l) acc += first[l] * first[l];
    acc += phi[0] * phi[0];
    acc = sqrt(acc);
    for (int l = 0; l < N; ++l) first[l] /= acc;
    phi[0] /= acc;
  }
  if (need_l) {
    double acc = 0.0;
    for (int l = 0; l < N; ++l) acc += last[l] * last[l];
    acc += psi[0] * psi[0];
    acc = sqrt(acc);
    for (int l = 0; l < N; ++l) last[l] /= acc;
    psi[0] /= acc;
  }
  free(z);
  free(l1);
  free(f2);
}

void orc_dc_values(int N, const double *b1, const double *b2, double *sigma) {
  double *scratch = (double *)malloc(sizeof(double) * N);
  double phi, psi;
  dc_rec(N, b1, b2, sigma, scratch, NULL, NULL, &phi, &psi, 0);
  free(scratch);
}

/* ------------------------------------------------------------------------ */
/* Stage 3: twisted factorization of B^T B - sigma^2 I                       */
/* ------------------------------------------------------------------------ */

void orc_right_vectors(int n, int m, const double *a, const double *b,
                       const double *sigma, double *X) {
  /* tridiagonal T = B^T B: diagonal ta (m), off-diagonal tb (m-1);
   * parallel-twisted.c:290-316 */
  double *ta = (double *)malloc(sizeof(double) * m);
  double *tb = (double *)malloc(sizeof(double) * m);
  long sz = (long)n * m;
  double *L = (double *)malloc(sizeof(double) * sz);  /* P: forward multipliers */
  double *D1 = (double *)malloc(sizeof(double) * sz);
  double *U = (double *)malloc(sizeof(double) * sz);  /* Q: backward multipliers */
  double *D2 = (double *)malloc(sizeof(double) * sz);
  double *G = (double *)malloc(sizeof(double) * sz);  /* twist values gamma */
  double *xt = (double *)malloc(sizeof(double) * sz);

  ta[0] = a[0] * a[0];
  for (int i = 0; i < n - 1; ++i) {
    tb[i] = a[i] * b[i];
    ta[i + 1] = a[i + 1] * a[i + 1] + b[i] * b[i];
  }
  if (m > n) {
    tb[n - 1] = a[n - 1] * b[n - 1];
    ta[n] = b[n - 1] * b[n - 1];
  }

  /* forward L D1 L^T and backward U D2 U^T of T - sigma_k^2 I
   * (parallel-twisted.c:320-362) */
#pragma omp parallel for
  for (int k = 0; k < n; ++k) {
    long o = (long)k * m;
    double s2 = sigma[k] * sigma[k];
    D1[o] = ta[0] - s2;
    D2[o + m - 1] = ta[m - 1] - s2;
    for (int i = 0; i < m - 1; ++i) {
      L[o + i] = tb[i] / D1[o + i];
      D1[o + i + 1] = (ta[i + 1] - s2 - L[o + i] * L[o + i] * D1[o + i]);
      int r = m - 2 - i;
      U[o + r] = tb[r] / D2[o + r + 1];
      D2[o + r] = ta[r] - s2 - D2[o + r + 1] * U[o + r] * U[o + r];
    }
  }

  /* twist values, twist index, and the solve N_k x = e_k
   * (parallel-twisted.c:431-528) */
#pragma omp parallel for
  for (int k = 0; k < n; ++k) {
    long o = (long)k * m;
    double s2 = sigma[k] * sigma[k];
    G[o] = D1[o] + D2[o] - (a[0] * a[0] - s2);
    for (int j = 1; j < n; ++j)
      G[o + j] = D1[o + j] + D2[o + j] - (a[j] * a[j] + b[j - 1] * b[j - 1] - s2);
    if (m > n) G[o + m - 1] = D1[o + m - 1] + D2[o + m - 1] - (b[n - 1] * b[n - 1] - s2);
  }

#pragma omp parallel for
  for (int k = 0; k < n; ++k) {
    long o = (long)k * m;
    /* smallest |gamma|, the last one on ties (which_min_gamma,
     * parallel-twisted.c:240-288, keeps the right half on ties) */
    int tw = 0;
    double best = G[o];
    for (int j = 1; j < m; ++j)
      if (!(fabs(G[o + tw]) < fabs(G[o + j]))) { tw = j; best = G[o + j]; }
    (void)best;
    double *x = X + o;
    x[tw] = 1.0;
    for (int j = tw + 1; j < m; ++j) x[j] = -1.0 * U[o + j - 1] * x[j - 1];
    for (int j = tw - 1; j >= 0; --j) x[j] = -1.0 * L[o + j] * x[j + 1];
  }

  /* one refinement solve twisted at m/2 (parallel-twisted.c:365-427); the
   * reference's index pattern needs m >= 3 */
  if (m >= 3) {
#pragma omp parallel for
    for (int k = 0; k < n; ++k) {
      long o = (long)k * m;
      double *x = X + o, *t = xt + o;
      const double *Lk = L + o, *Uk = U + o, *D1k = D1 + o, *D2k = D2 + o;
      int h = m / 2;
      t[0] = x[0];
      t[m - 1] = x[m - 1];
      for (int j = 1; j < h; ++j) t[j] = x[j] - Lk[j - 1] * t[j - 1];
      for (int j = m - 2; j > h; --j) t[j] = x[j] - Uk[j] * t[j + 1];
      t[h] = x[h] - Uk[h] * t[h + 1] - Lk[h - 1] * t[h - 1];
      x[h] = t[h] / G[o + h];
      for (int j = h + 1; j < m; ++j) x[j] = (t[j] - D2k[j] * Uk[j - 1] * x[j - 1]) / D2k[j];
      for (int j = h - 1; j >= 0; --j) x[j] = (t[j] - D1k[j] * Lk[j] * x[j + 1]) / D1k[j];
    }
  }

  /* normalise each vector (parallel-twisted.c:88-123) */
#pragma omp parallel for
  for (int k = 0; k < n; ++k) {
    double *x = X + (long)k * m;
    double nrm = 0.0;
    for (int j = 0; j < m; ++j) nrm = nrm + x[j] * x[j];
    nrm = sqrt(nrm);
    for (int j = 0; j < m; ++j) x[j] = x[j] / nrm;
  }

  free(ta); free(tb); free(L); free(D1); free(U); free(D2); free(G); free(xt);
}

void orc_left_vectors(int n, int m, const double *a, const double *b,
                      const double *sigma, const double *X, double *Y) {
  for (int k = 0; k < n; ++k) {
    const double *x = X + (long)k * m;
    double *y = Y + (long)k * n;
#pragma omp parallel for
    for (int i = 0; i < n - 1; ++i) y[i] = a[i] * x[i] + b[i] * x[i + 1];
    if (n == m)
      y[n - 1] = a[n - 1] * x[n - 1];
    else
      y[n - 1] = a[n - 1] * x[n - 1] + b[n - 1] * x[n];
    for (int j = 0; j < n; ++j) y[j] = y[j] / sigma[k];
  }
}

/* ------------------------------------------------------------------------ */
/* Stage 4: back-transformation                                             */
/* ------------------------------------------------------------------------ */

void orc_mult_u(int m, int n, int vec, const double *A_mod, const double *Y,
                double *U) {
  int mn = (m < n) ? m : n;
  int last = (m < n) ? mn : mn - 1;
  for (int i = 0; i < mn; ++i) U[i] = Y[(long)vec * mn + i];
  for (int i = mn; i < m; ++i) U[i] = 0.0;
  for (int j = last; j >= 0; --j) {
    const double *h = A_mod + j + (long)j * m;
    double dp = 0.0;
    for (int t = 0; t < m - j; ++t) dp += U[j + t] * h[t];
    for (int k = j; k < m; ++k) U[k] -= 2 * A_mod[k + (long)j * m] * dp;
  }
}

void orc_mult_v(int m, int n, int vec, const double *AT, const double *X,
                double *V, int fix_multv) {
  int mn = (m < n) ? m : n;
  int nref = (m < n) ? mn : mn - 1;
  int len_beta = (m >= n) ? n - 1 : m;
  long xs = fix_multv ? (long)(len_beta + 1) : (long)mn; /* X row stride */
  long ld = fix_multv ? (long)n : (long)m;               /* AT leading dim */
  for (int i = 0; i < nref + 1; ++i) V[i] = X[(long)vec * xs + i];
  for (int i = nref + 1; i < n; ++i) V[i] = 0.0;
  /* The reference starts at reflector nref+1 (bidiag_par.c:1031); for
   * m >= n the two top indices are empty, for m < n they name columns of AT
   * that do not exist (the reference reads past the array there), so the
   * wide case starts at the last real reflector m-1. */
  int jstart = (m < n) ? mn - 1 : nref + 1;
  for (int j = jstart; j >= 0; --j) {
    const double *h = AT + j + 1 + (long)j * ld;
    double dp = 0.0;
    for (int t = 0; t < n - j - 1; ++t) dp += h[t] * V[j + 1 + t];
    for (int k = j + 1; k < n; ++k) V[k] -= 2 * AT[(long)j * ld + k] * dp;
  }
}

/* ------------------------------------------------------------------------ */
/* Orchestration: svd_gpu.c:53-131                                           */
/* ------------------------------------------------------------------------ */

int orc_svd(int m, int n, double *A, double *sigma, double *U, double *V,
            double *alpha_out, double *beta_out, int fix_multv) {
  int mn = (m >= n) ? n : m;
  int len_beta = (m >= n) ? n - 1 : m;
  double *AT = (double *)malloc(sizeof(double) * (long)m * n);
  double *alpha = (double *)malloc(sizeof(double) * mn);
  /* one spare slot: the value solver reads beta[mn-1], which is 0 for m >= n */
  double *beta = (double *)calloc((size_t)len_beta + 1, sizeof(double));
  double *X = (double *)malloc(sizeof(double) * (long)mn * (len_beta + 1));
  double *Y = (double *)malloc(sizeof(double) * (long)mn * mn);
  if (!AT || !alpha || !beta || !X || !Y) {
    free(AT); free(alpha); free(beta); free(X); free(Y);
    return -1;
  }
  double t0 = now_s();
  orc_bidiag_seq(m, n, A, alpha, beta);
  double t1 = now_s();
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < m; ++j) AT[i + (long)j * n] = A[j + (long)i * m];
  orc_dc_values(mn, alpha, beta, sigma);
  double t2 = now_s();
  orc_right_vectors(mn, len_beta + 1, alpha, beta, sigma, X);
  orc_left_vectors(mn, len_beta + 1, alpha, beta, sigma, X, Y);
  double t3 = now_s();
#pragma omp parallel for
  for (int i = 0; i < mn; ++i) {
    orc_mult_u(m, n, i, A, Y, U + (long)i * m);
    orc_mult_v(m, n, i, AT, X, V + (long)i * n, fix_multv);
  }
  double t4 = now_s();
  g_stage_s[0] = t1 - t0;
  g_stage_s[1] = t2 - t1;
  g_stage_s[2] = t3 - t2;
  g_stage_s[3] = t4 - t3;
  if (alpha_out) memcpy(alpha_out, alpha, sizeof(double) * mn);
  if (beta_out) memcpy(beta_out, beta, sizeof(double) * len_beta);
  free(AT); free(alpha); free(beta); free(X); free(Y);
  return 0;
}

/* form_u, bidiag.c:252-302: column i of U is H_0 ... H_top e_i with
 * top = min(i, mn-1), the reflectors applied from the last one down. */
void orc_form_u(int m, int n, const double *A_mod, double *U) {
  const int mn = m < n ? m : n;
  for (int i = 0; i < m; ++i) {
    double *u = U + (size_t)i * m;
    for (int r = 0; r < m; ++r) u[r] = 0.0;
    u[i] = 1.0;
    const int top = i < mn - 1 ? i : mn - 1;
    for (int j = top; j >= 0; --j) {
      const double *v = A_mod + j + (size_t)j * m;
      double d = 0.0;
      for (int r = 0; r < m - j; ++r) d += u[j + r] * v[r];
      for (int r = 0; r < m - j; ++r) u[j + r] -= 2.0 * v[r] * d;
    }
  }
}

/* form_v, bidiag.c:305-359: column i of V is G_0 ... G_top e_i, the right
 * reflector j being row j of A_mod from column j+1 on, num_refs = mn-1 for
 * m >= n (mn for m < n), top = min(i-1, num_refs-1). */
void orc_form_v(int m, int n, const double *A_mod, double *V) {
  const int mn = m < n ? m : n;
  const int num_refs = m < n ? mn : mn - 1;
  for (int i = 0; i < n; ++i) {
    double *v = V + (size_t)i * n;
    for (int r = 0; r < n; ++r) v[r] = 0.0;
    v[i] = 1.0;
    const int top = i < num_refs ? i - 1 : num_refs - 1;
    for (int j = top; j >= 0; --j) {
      const double *w = A_mod + j + (size_t)(j + 1) * m;  /* row j, stride m */
      double d = 0.0;
      for (int r = 0; r < n - j - 1; ++r) d += w[(size_t)r * m] * v[j + 1 + r];
      for (int r = 0; r < n - j - 1; ++r) v[j + 1 + r] -= 2.0 * w[(size_t)r * m] * d;
    }
  }
}
