/*
 * Deterministic synthetic sparse matrices for the benchmark workloads.
 *
 * Stream layout: row r draws from its own counter-based stream
 *   key_r = mix64(seed * PHI + (r + 1) * ROWK),   u_k = mix64(key_r + k * PHI), k = 1, 2, ...
 * (mix64 = splitmix64 finaliser, PHI = 0x9E3779B97F4A7C15,
 * ROWK = 0xD1B54A32D192ED03). Within a row the draws are: the row length,
 * then column draws (with rejection of repeats), then one value per entry in
 * ascending column order. Rows are independent, so the output does not
 * depend on the thread count.
 *
 * Kinds:
 *   0 FIXED_UNIFORM:  p1 distinct uniform columns per row; values U(0,1)
 *   1 POISSON_ZIPF:   nnz ~ Poisson(p1) clipped to [1, n]; columns without
 *                     replacement, popularity of rank r proportional to r^-p2
 *                     over a seeded permutation of the columns;
 *                     values log1p(Poisson(2) + 1)
 *   2 PARETO_UNIFORM: nnz = min(floor(Pareto(alpha = p1, x_m = p2)), p3);
 *                     uniform columns without replacement; values N(0,1)
 *   3 POISSON_UNIFORM: nnz ~ Poisson(p1) clipped to [1, n]; uniform columns
 *                     without replacement; values log1p(Poisson(2) + 1)
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define PHI 0x9E3779B97F4A7C15ULL
#define ROWK 0xD1B54A32D192ED03ULL

static inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

typedef struct {
  uint64_t key, ctr;
} stream_t;

static inline uint64_t draw(stream_t *s) { return mix64(s->key + (++s->ctr) * PHI); }
static inline double unif(stream_t *s) { return (double)(draw(s) >> 11) * 0x1.0p-53; } /* [0,1) */

static stream_t row_stream(uint64_t seed, int64_t r) {
  stream_t s = {mix64(seed * PHI + (uint64_t)(r + 1) * ROWK), 0};
  return s;
}

static int64_t poisson(stream_t *s, double lam) {
  double u = unif(s), p = exp(-lam), c = p;
  int64_t k = 0;
  while (u > c && k < 100000) {
    ++k;
    p *= lam / (double)k;
    c += p;
  }
  return k;
}

static double normal(stream_t *s) {
  double u1 = 1.0 - unif(s), u2 = unif(s);
  return sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
}

static int64_t row_len(int kind, stream_t *s, int64_t n, double p1, double p2, double p3) {
  int64_t k;
  switch (kind) {
    case 0: k = (int64_t)p1; break;
    case 1:
    case 3: k = poisson(s, p1); if (k < 1) k = 1; break;
    case 2: {
      double x = p2 / pow(1.0 - unif(s), 1.0 / p1);
      k = (int64_t)floor(x);
      if (k > (int64_t)p3) k = (int64_t)p3;
      break;
    }
    default: k = 0;
  }
  if (k > n) k = n;
  return k;
}

void synth_row_lengths(int kind, int64_t m, int64_t n, uint64_t seed, double p1, double p2, double p3,
                       int64_t *lens) {
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < m; ++r) {
    stream_t s = row_stream(seed, r);
    lens[r] = row_len(kind, &s, n, p1, p2, p3);
  }
}

static int cmp_i64(const void *a, const void *b) {
  int64_t x = *(const int64_t *)a, y = *(const int64_t *)b;
  return (x > y) - (x < y);
}

/* Fill colidx (int64) and vals (f64) for rows with offsets rowptr. */
void synth_fill(int kind, int64_t m, int64_t n, uint64_t seed, double p1, double p2, double p3,
                const int64_t *rowptr, int64_t *colidx, double *vals) {
  /* Zipf table and column permutation (kind 1). */
  double *cdf = NULL;
  int64_t *perm = NULL;
  if (kind == 1) {
    cdf = (double *)malloc((size_t)n * sizeof(double));
    perm = (int64_t *)malloc((size_t)n * sizeof(int64_t));
    double acc = 0.0;
    for (int64_t i = 0; i < n; ++i) {
      acc += pow((double)(i + 1), -p2);
      cdf[i] = acc;
    }
    for (int64_t i = 0; i < n; ++i) cdf[i] /= acc;
    stream_t ps = {mix64(seed ^ 0x5045524D55544531ULL), 0};
    for (int64_t i = 0; i < n; ++i) perm[i] = i;
    for (int64_t i = n - 1; i > 0; --i) { /* Fisher-Yates */
      int64_t j = (int64_t)(draw(&ps) % (uint64_t)(i + 1));
      int64_t t = perm[i];
      perm[i] = perm[j];
      perm[j] = t;
    }
  }
#pragma omp parallel
  {
    uint8_t *seen = (uint8_t *)calloc((size_t)n, 1);
#pragma omp for schedule(dynamic, 256)
    for (int64_t r = 0; r < m; ++r) {
      stream_t s = row_stream(seed, r);
      const int64_t k = row_len(kind, &s, n, p1, p2, p3);
      int64_t *cols = colidx + rowptr[r];
      int64_t got = 0;
      if (k == n) {
        for (int64_t c = 0; c < n; ++c) cols[got++] = c;
      } else {
        while (got < k) {
          int64_t c;
          if (kind == 1) {
            const double u = unif(&s);
            int64_t lo = 0, hi = n - 1;
            while (lo < hi) {
              const int64_t mid = (lo + hi) / 2;
              if (cdf[mid] < u)
                lo = mid + 1;
              else
                hi = mid;
            }
            c = perm[lo];
          } else {
            c = (int64_t)(draw(&s) % (uint64_t)n);
          }
          if (!seen[c]) {
            seen[c] = 1;
            cols[got++] = c;
          }
        }
        for (int64_t i = 0; i < k; ++i) seen[cols[i]] = 0;
      }
      qsort(cols, (size_t)k, sizeof(int64_t), cmp_i64);
      double *v = vals + rowptr[r];
      for (int64_t i = 0; i < k; ++i) {
        switch (kind) {
          case 0: v[i] = unif(&s); break;
          case 2: v[i] = normal(&s); break;
          default: v[i] = log1p((double)poisson(&s, 2.0) + 1.0); break;
        }
      }
    }
    free(seen);
  }
  free(cdf);
  free(perm);
}
