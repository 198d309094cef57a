/*
 * CPU oracle for the Poisson bmnd() path -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference denoiser's default Poisson branch
 * (cduhme/bmnd v0.3.0, the modules under src/bmnd/).  Every function cites the reference
 * lines it follows.  It exists to check the HIP path (bit-for-bit where the
 * reference's arithmetic is discrete) and to be timed as the CPU baseline.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * it; the product package never does.
 *
 * Arithmetic rules (they mirror numba/numpy semantics of the reference):
 *  - built with -ffp-contract=off so a*b+c is never fused implicitly;
 *  - float32 dot products of the separable transforms are written as
 *    explicit fmaf chains in index order (the reference's BLAS sgemm order is
 *    implementation-defined, so this order is the one this project fixes and
 *    the GPU kernels reproduce exactly);
 *  - the Poisson deviance is evaluated in float64 (numba promotes the
 *    expression through the float64 `1.0/poisson_count_scale`,
 *    blockmatching.py:338-371) with the C library's log();
 *  - risks are float64, weights float32 (aggregationweights.py:213-259,
 *    616-630); scatter, division and variance sums are float32;
 *  - the SSD distance is a float32 sum; under the Poisson gamma correction
 *    the corrected distance is a float64 (the bias is a float64 sum), ordered
 *    against the float32 values the candidate buffers keep
 *    (blockmatching.py:432-443, 1099-1102, 1140-1145).
 *
 * Parity pins: every function here is checked in tests/test_oracle_*.py
 * against the reference's own known answers (blockmatching, shifts,
 * variance, weights, scatter tests) and against brute-force / closed-form
 * references (dense T*Sigma*T^T variance, exhaustive top-K, np.kaiser).
 * The within-segment order of the exact-variance reduction is the one
 * documented deviation: the reference sorts with numba's unstable argsort,
 * this oracle (and the GPU) use the stable order (increasing member index).
 */
#include "bmnd_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define F32_EPS 1.1920928955078125e-07 /* np.finfo(np.float32).eps */

/* ======================================================================== */
/* geometry helpers                                                          */
/* ======================================================================== */

static int64_t prod_int(const int *v, int n) {
  int64_t p = 1;
  for (int i = 0; i < n; ++i) p *= v[i];
  return p;
}

static void c_strides(int ndim, const int64_t *shape, int64_t *strides) {
  int64_t s = 1;
  for (int d = ndim - 1; d >= 0; --d) {
    strides[d] = s;
    s *= shape[d];
  }
}

/* local coordinates of the patch entries in C order (utils.py:15-31) */
static void local_coords(int ndim, const int *block, int p, int *coords) {
  for (int d = ndim - 1; d >= 0; --d) {
    coords[d] = p % block[d];
    p /= block[d];
  }
}

/* ======================================================================== */
/* Kaiser window: np.kaiser restated (utils.py:5-12)                          */
/* ======================================================================== */

static long double bessel_i0(long double x) {
  /* I0(x) = sum_k ((x/2)^2)^k / (k!)^2, summed until the terms vanish */
  long double q = (x * 0.5L) * (x * 0.5L);
  long double term = 1.0L, sum = 1.0L;
  for (int k = 1; k < 500; ++k) {
    term *= q / ((long double)k * (long double)k);
    sum += term;
    if (term < sum * 1e-24L) break;
  }
  return sum;
}

void oracle_kaiser(int m, double beta, double *out) {
  if (m == 1) {
    out[0] = 1.0;
    return;
  }
  double alpha = (m - 1) / 2.0;
  double denom = (double)bessel_i0((long double)beta);
  for (int n = 0; n < m; ++n) {
    double t = ((double)n - alpha) / alpha;
    double arg = beta * sqrt(1.0 - t * t);
    out[n] = (double)bessel_i0((long double)arg) / denom;
  }
}

void oracle_kaiser_nd(int ndim, const int *block, double beta, float *out) {
  /* w = kaiser(b0); w = outer(w, kaiser(b_d)) in float64; cast to float32 */
  int64_t total = prod_int(block, ndim);
  double *w = (double *)malloc(sizeof(double) * (size_t)total);
  double axis[64];
  oracle_kaiser(block[0], beta, axis);
  int64_t cur = block[0];
  for (int i = 0; i < block[0]; ++i) w[i] = axis[i];
  for (int d = 1; d < ndim; ++d) {
    oracle_kaiser(block[d], beta, axis);
    for (int64_t i = cur - 1; i >= 0; --i) {
      double wi = w[i];
      for (int j = 0; j < block[d]; ++j) w[i * block[d] + j] = wi * axis[j];
    }
    cur *= block[d];
  }
  for (int64_t i = 0; i < total; ++i) out[i] = (float)w[i];
  free(w);
}

/* ======================================================================== */
/* transform matrices (transforms.py:359-440, 699-856)                        */
/* ======================================================================== */

static int is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

/* hierarchical Haar, transforms.py:393-408: rows are [coarse; detail] built
 * by float32 Kronecker products with s = float32(1/sqrt(2)). */
int oracle_haar_matrix(int n, float *fwd) {
  if (!is_pow2(n)) return -1;
  float s = (float)(1.0 / sqrt(2.0));
  float *m = (float *)malloc(sizeof(float) * (size_t)n * n);
  float *nm = (float *)malloc(sizeof(float) * (size_t)n * n);
  m[0] = 1.0f;
  int cur = 1;
  while (cur < n) {
    int cols = 2 * cur;
    /* coarse = kron(m, [s, s]) : cur x 2cur */
    for (int r = 0; r < cur; ++r)
      for (int c = 0; c < cur; ++c) {
        nm[r * cols + 2 * c] = m[r * cur + c] * s;
        nm[r * cols + 2 * c + 1] = m[r * cur + c] * s;
      }
    /* detail = kron(eye(cur), [s, -s]) : cur x 2cur (zeros keep numpy's
     * signs: 0*s = +0, 0*(-s) = -0) */
    for (int r = 0; r < cur; ++r)
      for (int c = 0; c < cur; ++c) {
        float e = (r == c) ? 1.0f : 0.0f;
        nm[(cur + r) * cols + 2 * c] = e * s;
        nm[(cur + r) * cols + 2 * c + 1] = e * (-s);
      }
    cur = cols;
    memcpy(m, nm, sizeof(float) * (size_t)cur * cur);
  }
  memcpy(fwd, m, sizeof(float) * (size_t)n * n);
  free(m);
  free(nm);
  return 0;
}

/* DCT-II orthonormal, transforms.py:359-386 (also the non-power-of-two group
 * fallback, transforms.py:425-436).  fwd[f*n+s], inv[s*n+f]. */
void oracle_dct_matrix(int n, float *fwd, float *inv) {
  const double pi = 3.141592653589793;
  for (int f = 0; f < n; ++f)
    for (int s = 0; s < n; ++s) {
      double norm = (f == 0) ? sqrt(1.0 / n) : sqrt(2.0 / n);
      double arg = pi * ((double)s + 0.5) * (double)f / (double)n;
      float v = (float)(norm * cos(arg));
      if (fwd) fwd[f * n + s] = v;
      if (inv) inv[s * n + f] = v;
    }
}

/* LU with partial pivoting in float64 (the computation numpy.linalg.inv does
 * for a float32 input), result rounded to float32 (transforms.py:810). */
static int invert_f32_via_f64(int n, const float *a, float *out) {
  double lu[64 * 64];
  int piv[64];
  for (int i = 0; i < n * n; ++i) lu[i] = a[i];
  for (int i = 0; i < n; ++i) piv[i] = i;
  for (int k = 0; k < n; ++k) {
    int p = k;
    double best = fabs(lu[k * n + k]);
    for (int i = k + 1; i < n; ++i)
      if (fabs(lu[i * n + k]) > best) {
        best = fabs(lu[i * n + k]);
        p = i;
      }
    if (best == 0.0) return -1;
    if (p != k) {
      for (int j = 0; j < n; ++j) {
        double t = lu[k * n + j];
        lu[k * n + j] = lu[p * n + j];
        lu[p * n + j] = t;
      }
      int t = piv[k];
      piv[k] = piv[p];
      piv[p] = t;
    }
    for (int i = k + 1; i < n; ++i) {
      lu[i * n + k] /= lu[k * n + k];
      for (int j = k + 1; j < n; ++j) lu[i * n + j] -= lu[i * n + k] * lu[k * n + j];
    }
  }
  for (int col = 0; col < n; ++col) {
    double x[64];
    for (int i = 0; i < n; ++i) x[i] = (piv[i] == col) ? 1.0 : 0.0;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < i; ++j) x[i] -= lu[i * n + j] * x[j];
    for (int i = n - 1; i >= 0; --i) {
      for (int j = i + 1; j < n; ++j) x[i] -= lu[i * n + j] * x[j];
      x[i] /= lu[i * n + i];
    }
    for (int i = 0; i < n; ++i) out[i * n + col] = (float)x[i];
  }
  return 0;
}

/* Official bior1.5 analysis matrices for lengths 4 and 8, the literal values
 * of transforms.py:711-806 (same table; a format the reference fixes). */
int oracle_bior15_matrix(int n, float *fwd, float *inv) {
  if (n == 4) {
    const float a = 0.70710677f;
    const float m[16] = {0.5f, 0.5f, 0.5f,  0.5f,  0.5f, 0.5f, -0.5f, -0.5f,
                         a,    -a,   0.0f,  0.0f,  0.0f, 0.0f, a,     -a};
    memcpy(fwd, m, sizeof(m));
  } else if (n == 8) {
    const float c0 = (float)0.343550200747110, c1 = (float)-0.225454819240296,
                c2 = (float)-0.461645582253923, c3 = (float)0.569359398342840,
                c4 = (float)0.402347308162280, c5 = (float)-0.083506045090280,
                c6 = (float)0.707106781186550;
    const float m[64] = {
        c0,  c0,  c0,  c0,  c0,  c0,  c0,  c0,  /* row 0 */
        c1,  c2,  c2,  c1,  -c1, -c2, -c2, -c1, /* row 1 */
        c3,  c4,  -c4, -c3, c5,  -c5, c5,  -c5, /* row 2 */
        c5,  -c5, c5,  -c5, c3,  c4,  -c4, -c3, /* row 3 */
        c6,  -c6, 0,   0,   0,   0,   0,   0,   /* row 4 */
        0,   0,   c6,  -c6, 0,   0,   0,   0,   /* row 5 */
        0,   0,   0,   0,   c6,  -c6, 0,   0,   /* row 6 */
        0,   0,   0,   0,   0,   0,   c6,  -c6  /* row 7 */
    };
    memcpy(fwd, m, sizeof(m));
  } else {
    return -1;
  }
  return inv ? invert_f32_via_f64(n, fwd, inv) : 0;
}

/* group transform: Haar for powers of two, DCT otherwise (transforms.py:411-440) */
static void group_matrix(int n, float *fwd, float *inv) {
  if (is_pow2(n)) {
    oracle_haar_matrix(n, fwd);
  } else {
    oracle_dct_matrix(n, fwd, NULL);
  }
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) inv[i * n + j] = fwd[j * n + i];
}

/* ======================================================================== */
/* GENERATED reference schedule (shifts.py:229-266, 519-576)                  */
/* ======================================================================== */

static int axis_shifts(int block, int step, int allow_multiple, double density,
                       int *out) {
  if (!allow_multiple || block < 3 || step < 1 || step > block) {
    out[0] = 0;
    return 1;
  }
  double safe = density > 0.0 ? density : 0.0;
  int count = (int)ceil(safe * (double)(block - 1) / (double)step);
  if (count < 1) count = 1;
  if (count > block) count = block;
  if (count == 1) {
    out[0] = 0;
    return 1;
  }
  /* np.linspace(0, block-1, count): y[i] = i*step, last = stop exactly */
  double stop = (double)(block - 1);
  double lstep = stop / (double)(count - 1);
  unsigned char seen[256] = {0};
  seen[0] = 1;
  for (int i = 0; i < count; ++i) {
    double pos = (i == count - 1) ? stop : (double)i * lstep;
    int r = (int)rint(pos); /* Python round(): half to even */
    seen[r] = 1;
  }
  int n = 0;
  for (int v = 0; v < 256; ++v)
    if (seen[v]) out[n++] = v;
  return n;
}

int64_t oracle_schedule_generated(int ndim, const int64_t *counts,
                                  const int *block, const int *step,
                                  double shift_density, int schedule_density,
                                  int64_t *out_abs, int64_t cap) {
  if (ndim <= 0) return 0;
  int shifts[ORACLE_MAX_NDIM][256];
  int nshift[ORACLE_MAX_NDIM];
  int64_t last[ORACLE_MAX_NDIM], slots[ORACLE_MAX_NDIM], strides[ORACLE_MAX_NDIM];
  for (int d = 0; d < ndim; ++d) {
    nshift[d] = axis_shifts(block[d], step[d], d < (ndim - 1 > 0 ? ndim - 1 : 0),
                            shift_density, shifts[d]);
    last[d] = counts[d] - 1 > 0 ? counts[d] - 1 : 0;
    slots[d] = (last[d] + step[d] - 1) / step[d] + 1;
  }
  c_strides(ndim, counts, strides);
  int64_t total_grid = 1;
  for (int d = 0; d < ndim; ++d) total_grid *= counts[d];
  unsigned char *seen = (unsigned char *)calloc((size_t)total_grid, 1);
  int phases = schedule_density > 1 ? schedule_density : 1;
  int64_t nslot = 1;
  for (int d = 0; d < ndim; ++d) nslot *= slots[d];
  int64_t n = 0;
  int64_t slot[ORACLE_MAX_NDIM];
  for (int64_t sl = 0; sl < nslot; ++sl) {
    int64_t t = sl;
    for (int d = ndim - 1; d >= 0; --d) {
      slot[d] = t % slots[d];
      t /= slots[d];
    }
    for (int ph = 0; ph < phases; ++ph) {
      int64_t abs_pos[ORACLE_MAX_NDIM];
      int64_t lin = 0;
      for (int d = 0; d < ndim; ++d) {
        int64_t base = slot[d] * step[d];
        if (base > last[d]) base = last[d];
        int64_t offset = 0;
        if (d < ndim - 1) offset = shifts[d][(slot[d + 1] + ph) % nshift[d]];
        int64_t a = base - offset;
        if (a < 0) a = 0;
        if (a > last[d]) a = last[d];
        if (slot[d] == slots[d] - 1) a = last[d];
        abs_pos[d] = a;
        lin += a * strides[d];
      }
      if (seen[lin]) continue;
      seen[lin] = 1;
      if (out_abs && n < cap)
        for (int d = 0; d < ndim; ++d) out_abs[n * ndim + d] = abs_pos[d];
      ++n;
    }
  }
  free(seen);
  return n;
}

/* ======================================================================== */
/* block matching (blockmatching.py:322-371, 913-1205)                        */
/* ======================================================================== */

/* tie grid, blockmatching.py:957-964: q = 32 * eps32 * 2^floor(log2(max(|d|,1))) */
float oracle_canonical_key(float d) {
  if (!isfinite(d)) return d;
  double mag = fabs((double)d);
  if (mag < 1.0) mag = 1.0;
  int e = ilogb(mag);
  double q = 32.0 * F32_EPS * ldexp(1.0, e);
  return (float)(rint((double)d / q) * q);
}

/* The SSD branch accumulates in float32: its variable starts as
 * np.float32(0.0) and takes float32 sums only, which numba's per-assignment
 * (SSA) typing keeps float32; the other branches widen to float64
 * (blockmatching.py:320-371). */
float oracle_patch_distance(const float *ref, const float *cand, int p,
                            int code, double count_scale) {
  if (code == ORACLE_DIST_SSD) {
    float d32 = 0.0f;
    for (int i = 0; i < p; ++i) {
      float diff = cand[i] - ref[i];
      d32 = d32 + diff * diff;
    }
    return d32;
  }
  double dist = 0.0;
  double inv = 1.0 / count_scale;
  for (int i = 0; i < p; ++i) {
    double r = (double)ref[i] * inv;
    if (r < 0.0) r = 0.0;
    double c = (double)cand[i] * inv;
    if (c < 0.0) c = 0.0;
    if (code == ORACLE_DIST_DEVIANCE) {
      double comb = r + c;
      if (comb > 0.0) {
        double t = 0.0;
        if (r > 0.0) t += r * log(2.0 * r / comb);
        if (c > 0.0) t += c * log(2.0 * c / comb);
        dist += 2.0 * t;
      }
    } else if (code == ORACLE_DIST_PEARSON) {
      double comb = r + c;
      if (comb > 0.0) {
        double diff = c - r;
        dist += diff * diff / comb;
      }
    } else {
      double rs = 2.0 * sqrt(r + 0.375);
      double cs = 2.0 * sqrt(c + 0.375);
      double diff = cs - rs;
      dist += 0.5 * diff * diff;
    }
  }
  if (dist < 0.0) dist = 0.0;
  return (float)(dist / p);
}

/* Candidate-standardized distance, blockmatching.py:374-426: over the patch
 * voxels of positive pooled count r + c, (sum of the distance's terms - sum
 * of the null means of round(r + c)) / sqrt(sum of the null variances); the
 * null sums are float32, pooled counts above null_max take the chi-square(1)
 * moments 1 and 2; 0 when the variance sum is not positive. */
float oracle_standardized_distance(const float *ref, const float *cand, int p, int code,
                                   double count_scale, const float *null_means,
                                   const float *null_vars, int null_max) {
  double inv = 1.0 / count_scale;
  double dist = 0.0;
  float nm = 0.0f, nv = 0.0f;
  for (int i = 0; i < p; ++i) {
    double r = (double)ref[i] * inv;
    if (r < 0.0) r = 0.0;
    double c = (double)cand[i] * inv;
    if (c < 0.0) c = 0.0;
    double comb = r + c;
    if (comb <= 0.0) continue;
    if (code == ORACLE_DIST_DEVIANCE) {
      double t = 0.0;
      if (r > 0.0) t += r * log(2.0 * r / comb);
      if (c > 0.0) t += c * log(2.0 * c / comb);
      dist += 2.0 * t;
    } else if (code == ORACLE_DIST_PEARSON) {
      double diff = c - r;
      dist += diff * diff / comb;
    } else {
      double diff = 2.0 * sqrt(c + 0.375) - 2.0 * sqrt(r + 0.375);
      dist += 0.5 * diff * diff;
    }
    double pooled = rint(comb);
    if (pooled <= (double)null_max) {
      nm += null_means[(int)pooled];
      nv += null_vars[(int)pooled];
    } else {
      nm += 1.0f;
      nv += 2.0f;
    }
  }
  if (nv <= 0.0f) return 0.0f;
  return (float)((dist - (double)nm) / (double)sqrtf(nv));
}

static int precedes(int64_t cl, float cd, int64_t el, float ed, int stabilize) {
  if (stabilize) {
    cd = oracle_canonical_key(cd);
    ed = oracle_canonical_key(ed);
    if (cd == ed) return cl < el;
  }
  if (cd < ed) return 1;
  if (cd > ed) return 0;
  return cl < el;
}

/* bounded insertion, blockmatching.py:987-1029 */
static int insert_limited(int64_t *lins, float *dists, int limit, int size,
                          int64_t cl, float cd, int stabilize) {
  if (limit == 0) return size;
  int at;
  if (size < limit) {
    ++size;
    at = size - 1;
  } else {
    if (!precedes(cl, cd, lins[limit - 1], dists[limit - 1], stabilize))
      return size;
    at = limit - 1;
  }
  while (at > 0 && precedes(cl, cd, lins[at - 1], dists[at - 1], stabilize)) {
    lins[at] = lins[at - 1];
    dists[at] = dists[at - 1];
    --at;
  }
  lins[at] = cl;
  dists[at] = cd;
  return size;
}

/* insert_limited for a float64 candidate distance against the float32
 * distances the buffer stores: the Poisson SSD correction makes the
 * candidate distance a float64 (gamma times the float64 noise bias,
 * blockmatching.py:432-443, 1140-1145), while the buffers are float32 arrays
 * (blockmatching.py:1099-1102), so a candidate precedes an entry when its
 * float64 value is below the entry's stored float32 value, or equal to it
 * with the smaller index (blockmatching.py:967-1030, no tie grid for SSD). */
static int mixed_precedes(int64_t cl, double cd, int64_t el, float ed) {
  if (cd < (double)ed) return 1;
  if (cd > (double)ed) return 0;
  return cl < el;
}

static int insert_limited_mixed(int64_t *lins, float *dists, int limit, int size, int64_t cl,
                                double cd) {
  if (limit == 0) return size;
  int at;
  if (size < limit) {
    ++size;
    at = size - 1;
  } else {
    if (!mixed_precedes(cl, cd, lins[limit - 1], dists[limit - 1])) return size;
    at = limit - 1;
  }
  while (at > 0 && mixed_precedes(cl, cd, lins[at - 1], dists[at - 1])) {
    lins[at] = lins[at - 1];
    dists[at] = dists[at - 1];
    --at;
  }
  lins[at] = cl;
  dists[at] = (float)cd;
  return size;
}

/* Poisson SSD noise bias of a candidate (blockmatching.py:432-443): over the
 * patch in C order, the float32 pooled value r + c, and where it is positive
 * count_scale times it added to a float64 sum. */
static double poisson_ssd_bias(const float *ref, const float *cand, int p, double scale) {
  double bias = 0.0;
  for (int i = 0; i < p; ++i) {
    const float pooled = ref[i] + cand[i];
    if (pooled > 0.0f) bias += scale * (double)pooled;
  }
  return bias;
}

static int largest_pow2_at_most(int v) {
  if (v <= 0) return 0;
  int p = 1;
  while ((p << 1) <= v) p <<= 1;
  return p;
}

/* reference first, dedupe, cap, power-of-two truncation (blockmatching.py:913-954) */
static int finalize(const int64_t *sel, int sel_count, int64_t center,
                    int max_m, int min_m, int pow2, int64_t *out) {
  if (max_m <= 0) return 0;
  int cnt = 0;
  out[cnt++] = center;
  for (int i = 0; i < sel_count; ++i) {
    int64_t lin = sel[i];
    if (lin < 0 || lin == center) continue;
    int seen = 0;
    for (int j = 0; j < cnt; ++j)
      if (out[j] == lin) {
        seen = 1;
        break;
      }
    if (!seen) {
      out[cnt++] = lin;
      if (cnt >= max_m) break;
    }
  }
  if (pow2 && cnt > 0) {
    int target = largest_pow2_at_most(cnt);
    if (min_m > target) target = min_m;
    if (target < cnt) cnt = target;
  }
  return cnt;
}

void oracle_blockmatch(const float *patches_flat, const float *volume,
                       int ndim, const int64_t *vol_shape,
                       const int64_t *counts, const int *block,
                       const int64_t *refs, int64_t nrefs, const int *radius,
                       const float *thresholds, int max_matches,
                       int min_matches, int pow2, int stabilize, int code,
                       double count_scale, int nthreads, int64_t *sel_out,
                       int32_t *count_out, int64_t *accepted_out) {
  oracle_blockmatch_standardized(patches_flat, volume, ndim, vol_shape, counts, block, refs,
                                 nrefs, radius, thresholds, max_matches, min_matches, pow2,
                                 stabilize, code, count_scale, NULL, NULL, 0, nthreads, sel_out,
                                 count_out, accepted_out);
}

static void blockmatch_core(const float *patches_flat, const float *volume, int ndim,
                            const int64_t *vol_shape, const int64_t *counts, const int *block,
                            const int64_t *refs, int64_t nrefs, const int *radius,
                            const float *thresholds, int max_matches, int min_matches, int pow2,
                            int stabilize, int code, double count_scale,
                            const float *null_means, const float *null_vars, int null_max,
                            const float *noise_ssd, float gamma, float pgamma, double pscale,
                            int nthreads, int64_t *sel_out, int32_t *count_out,
                            int64_t *accepted_out);

/* oracle_blockmatch; with null_means the candidate-standardized distance
 * (blockmatching.py:1128-1145) */
void oracle_blockmatch_standardized(const float *patches_flat, const float *volume, int ndim,
                                    const int64_t *vol_shape, const int64_t *counts,
                                    const int *block, const int64_t *refs, int64_t nrefs,
                                    const int *radius, const float *thresholds, int max_matches,
                                    int min_matches, int pow2, int stabilize, int code,
                                    double count_scale, const float *null_means,
                                    const float *null_vars, int null_max, int nthreads,
                                    int64_t *sel_out, int32_t *count_out,
                                    int64_t *accepted_out) {
  blockmatch_core(patches_flat, volume, ndim, vol_shape, counts, block, refs, nrefs, radius,
                  thresholds, max_matches, min_matches, pow2, stabilize, code, count_scale,
                  null_means, null_vars, null_max, NULL, 0.0f, 0.0f, 1.0, nthreads, sel_out,
                  count_out, accepted_out);
}

/* SSD matching with the Gaussian noise-bias correction (blockmatching.py:
 * 1133-1139): every candidate's float32 SSD less gamma times the expected
 * noise SSD at its lag, noise_ssd being C-ordered over the (2 r + 1) search
 * box (NULL: no correction).  No tie grid (blockmatching.py:1108). */
void oracle_blockmatch_ssd(const float *volume, int ndim, const int64_t *vol_shape,
                           const int64_t *counts, const int *block, const int64_t *refs,
                           int64_t nrefs, const int *radius, const float *thresholds,
                           int max_matches, int min_matches, int pow2, const float *noise_ssd,
                           float gamma, int nthreads, int64_t *sel_out, int32_t *count_out,
                           int64_t *accepted_out) {
  blockmatch_core(NULL, volume, ndim, vol_shape, counts, block, refs, nrefs, radius, thresholds,
                  max_matches, min_matches, pow2, 0, ORACLE_DIST_SSD, 1.0, NULL, NULL, 0,
                  noise_ssd, gamma, 0.0f, 1.0, nthreads, sel_out, count_out, accepted_out);
}

/* SSD matching of the Poisson model's stage 1 with the Poisson gamma
 * correction (bmndalgo.py:292-294, 321-322; blockmatching.py:1140-1145):
 * every candidate other than the reference itself has gamma times its
 * Poisson noise bias (poisson_ssd_bias, count_scale = the profile's
 * poisson_count_scale) taken off its float32 SSD.  gamma is a float32 and the
 * bias a float64, so the corrected distance is a float64: it is compared with
 * the float32 threshold as a float64 and ordered against the float32 values
 * the buffers hold (insert_limited_mixed).  gamma == 0 is plain SSD. */
void oracle_blockmatch_ssd_poisson(const float *volume, int ndim, const int64_t *vol_shape,
                                   const int64_t *counts, const int *block, const int64_t *refs,
                                   int64_t nrefs, const int *radius, const float *thresholds,
                                   int max_matches, int min_matches, int pow2, float gamma,
                                   double count_scale, int nthreads, int64_t *sel_out,
                                   int32_t *count_out, int64_t *accepted_out) {
  blockmatch_core(NULL, volume, ndim, vol_shape, counts, block, refs, nrefs, radius, thresholds,
                  max_matches, min_matches, pow2, 0, ORACLE_DIST_SSD, 1.0, NULL, NULL, 0, NULL,
                  0.0f, gamma, count_scale, nthreads, sel_out, count_out, accepted_out);
}

static void blockmatch_core(const float *patches_flat, const float *volume, int ndim,
                            const int64_t *vol_shape, const int64_t *counts, const int *block,
                            const int64_t *refs, int64_t nrefs, const int *radius,
                            const float *thresholds, int max_matches, int min_matches, int pow2,
                            int stabilize, int code, double count_scale,
                            const float *null_means, const float *null_vars, int null_max,
                            const float *noise_ssd, float gamma, float pgamma, double pscale,
                            int nthreads, int64_t *sel_out, int32_t *count_out,
                            int64_t *accepted_out) {
  int P = (int)prod_int(block, ndim);
  int64_t cstr[ORACLE_MAX_NDIM], vstr[ORACLE_MAX_NDIM];
  c_strides(ndim, counts, cstr);
  int64_t *poff = (int64_t *)malloc(sizeof(int64_t) * (size_t)P);
  if (!patches_flat) {
    c_strides(ndim, vol_shape, vstr);
    for (int p = 0; p < P; ++p) {
      int lc[ORACLE_MAX_NDIM];
      local_coords(ndim, block, p, lc);
      int64_t o = 0;
      for (int d = 0; d < ndim; ++d) o += lc[d] * vstr[d];
      poff[p] = o;
    }
  }
  int stab = stabilize && code != ORACLE_DIST_SSD;
  int required = min_matches < max_matches ? min_matches : max_matches;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
#endif
  {
    float *refbuf = (float *)malloc(sizeof(float) * (size_t)P);
    float *candbuf = (float *)malloc(sizeof(float) * (size_t)P);
    int64_t *acc_l = (int64_t *)malloc(sizeof(int64_t) * (size_t)(max_matches + 1));
    float *acc_d = (float *)malloc(sizeof(float) * (size_t)(max_matches + 1));
    int64_t *rej_l = (int64_t *)malloc(sizeof(int64_t) * (size_t)(required + 1));
    float *rej_d = (float *)malloc(sizeof(float) * (size_t)(required + 1));
    int64_t *selb = (int64_t *)malloc(sizeof(int64_t) * (size_t)(max_matches + 1));
    int64_t *fin = (int64_t *)malloc(sizeof(int64_t) * (size_t)(max_matches + 1));
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 16)
#endif
    for (int64_t ri = 0; ri < nrefs; ++ri) {
      const int64_t *center = refs + ri * ndim;
      int64_t center_lin = 0;
      for (int d = 0; d < ndim; ++d) center_lin += center[d] * cstr[d];
      const float *refp;
      if (patches_flat) {
        refp = patches_flat + center_lin * P;
      } else {
        int64_t base = 0;
        for (int d = 0; d < ndim; ++d) base += center[d] * vstr[d];
        for (int p = 0; p < P; ++p) refbuf[p] = volume[base + poff[p]];
        refp = refbuf;
      }
      int64_t lo[ORACLE_MAX_NDIM], ext[ORACLE_MAX_NDIM], coord[ORACLE_MAX_NDIM];
      int64_t box = 1;
      for (int d = 0; d < ndim; ++d) {
        int64_t l = center[d] - radius[d];
        if (l < 0) l = 0;
        int64_t h = center[d] + radius[d];
        if (h > counts[d] - 1) h = counts[d] - 1;
        lo[d] = l;
        ext[d] = h - l + 1;
        coord[d] = l;
        box *= ext[d];
      }
      float thr = thresholds[ri];
      float thr_key = stab ? oracle_canonical_key(thr) : thr;
      int acc_size = 0, rej_size = 0;
      int64_t accepted = 0;
      for (int64_t b = 0; b < box; ++b) {
        int64_t lin = 0;
        for (int d = 0; d < ndim; ++d) lin += coord[d] * cstr[d];
        const float *cp;
        if (patches_flat) {
          cp = patches_flat + lin * P;
        } else {
          int64_t base = 0;
          for (int d = 0; d < ndim; ++d) base += coord[d] * vstr[d];
          for (int p = 0; p < P; ++p) candbuf[p] = volume[base + poff[p]];
          cp = candbuf;
        }
        float dist = null_means ? oracle_standardized_distance(refp, cp, P, code, count_scale,
                                                               null_means, null_vars, null_max)
                                : oracle_patch_distance(refp, cp, P, code, count_scale);
        if (noise_ssd) {
          int64_t ci = 0;
          for (int d = 0; d < ndim; ++d)
            ci = ci * (2 * radius[d] + 1) + (coord[d] - center[d] + radius[d]);
          dist = dist - gamma * noise_ssd[ci];
        }
        if (pgamma != 0.0f) {
          /* Poisson SSD correction: a float64 distance for every candidate
           * but the reference itself */
          double dd = (double)dist;
          if (lin != center_lin) dd = dd - (double)pgamma * poisson_ssd_bias(refp, cp, P, pscale);
          if (dd < (double)thr) {
            ++accepted;
            acc_size = insert_limited_mixed(acc_l, acc_d, max_matches, acc_size, lin, dd);
          } else {
            rej_size = insert_limited_mixed(rej_l, rej_d, required, rej_size, lin, dd);
          }
        } else {
          int ok = stab ? (oracle_canonical_key(dist) < thr_key) : (dist < thr);
          if (ok) {
            ++accepted;
            acc_size = insert_limited(acc_l, acc_d, max_matches, acc_size, lin, dist, stab);
          } else {
            rej_size = insert_limited(rej_l, rej_d, required, rej_size, lin, dist, stab);
          }
        }
        for (int d = ndim - 1; d >= 0; --d) {
          coord[d] += 1;
          if (coord[d] < lo[d] + ext[d]) break;
          coord[d] = lo[d];
        }
      }
      int sel_count = acc_size;
      for (int i = 0; i < acc_size; ++i) selb[i] = acc_l[i];
      if (sel_count < required) {
        int need = required - sel_count;
        int take = need < rej_size ? need : rej_size;
        for (int i = 0; i < take; ++i) selb[sel_count++] = rej_l[i];
      }
      int fc = finalize(selb, sel_count, center_lin, max_matches, min_matches, pow2, fin);
      for (int i = 0; i < max_matches; ++i) sel_out[ri * max_matches + i] = i < fc ? fin[i] : -1;
      count_out[ri] = fc;
      if (accepted_out) accepted_out[ri] = accepted;
    }
    free(refbuf);
    free(candbuf);
    free(acc_l);
    free(acc_d);
    free(rej_l);
    free(rej_d);
    free(selb);
    free(fin);
  }
  free(poff);
}

/* ======================================================================== */
/* separable transforms (transforms.py:209-255)                               */
/* ======================================================================== */

/* data has shape dims[0..ndim_total-1]; mats[a] is dims[a] x dims[a]
 * (row-major, out index first) or NULL to skip the axis.  Axes are applied in
 * order 0,1,...; each output is an fmaf chain over the input index. */
void oracle_apply_transform(float *data, int ndim_total, const int *dims,
                            const float *const *mats) {
  int64_t total = prod_int(dims, ndim_total);
  float *tmp = (float *)malloc(sizeof(float) * (size_t)total);
  for (int a = 0; a < ndim_total; ++a) {
    const float *m = mats[a];
    if (!m) continue;
    int n = dims[a];
    int64_t inner = 1;
    for (int d = a + 1; d < ndim_total; ++d) inner *= dims[d];
    int64_t outer = total / (inner * n);
    for (int64_t o = 0; o < outer; ++o)
      for (int r = 0; r < n; ++r)
        for (int64_t in = 0; in < inner; ++in) {
          float acc = 0.0f;
          const float *x = data + o * n * inner + in;
          for (int k = 0; k < n; ++k) acc = fmaf(m[r * n + k], x[k * inner], acc);
          tmp[o * n * inner + r * inner + in] = acc;
        }
    memcpy(data, tmp, sizeof(float) * (size_t)total);
  }
  free(tmp);
}

/* ======================================================================== */
/* exact overlap-aware Poisson variance (variance.py:652-975, 1634-1711)      */
/* ======================================================================== */

/* full separable spatial matrix, kron chain in float32 (variance.py:744-755) */
static float *spatial_kron(int ndim, const int *block, const float *const *mats) {
  int P = (int)prod_int(block, ndim);
  float *full = (float *)malloc(sizeof(float) * (size_t)P * P);
  float *tmp = (float *)malloc(sizeof(float) * (size_t)P * P);
  full[0] = 1.0f;
  int cur = 1;
  for (int d = 0; d < ndim; ++d) {
    int b = block[d];
    int nc = cur * b;
    for (int r = 0; r < cur; ++r)
      for (int c = 0; c < cur; ++c)
        for (int rr = 0; rr < b; ++rr)
          for (int cc = 0; cc < b; ++cc) {
            float mv = mats[d] ? mats[d][rr * b + cc] : (rr == cc ? 1.0f : 0.0f);
            tmp[(r * b + rr) * nc + (c * b + cc)] = full[r * cur + c] * mv;
          }
    cur = nc;
    memcpy(full, tmp, sizeof(float) * (size_t)cur * cur);
  }
  free(tmp);
  return full;
}

/* Order of the entries inside one union voxel.  0 (default): increasing
 * entry index, a stable sort.  1: the order numba's np.argsort leaves them in
 * (its default kind, quicksort: numba/misc/quicksort.py, median-of-three
 * partition, insertion sort below 15 elements, larger partition pushed on
 * the stack), which is what the reference's union segmentation uses
 * (variance.py:804).  Used to measure how far the two orders move the
 * float32 sums. */
static int g_exact_order = 0;

void oracle_set_exact_order(int quicksort) { g_exact_order = quicksort ? 1 : 0; }

static void numba_insertion_argsort(const int64_t *A, int32_t *R, int64_t low, int64_t high) {
  if (high <= low) return;
  for (int64_t i = low + 1; i <= high; ++i) {
    int32_t k = R[i];
    int64_t v = A[k];
    int64_t j = i;
    while (j > low && v < A[R[j - 1]]) {
      R[j] = R[j - 1];
      --j;
    }
    R[j] = k;
  }
}

static int64_t numba_partition_argsort(const int64_t *A, int32_t *R, int64_t low, int64_t high) {
  int64_t mid = (low + high) >> 1;
  int32_t t;
#define SWAPR(x, y) (t = R[x], R[x] = R[y], R[y] = t)
  if (A[R[mid]] < A[R[low]]) SWAPR(low, mid);
  if (A[R[high]] < A[R[mid]]) SWAPR(high, mid);
  if (A[R[mid]] < A[R[low]]) SWAPR(low, mid);
  int64_t pivot = A[R[mid]];
  SWAPR(high, mid);
  int64_t i = low, j = high - 1;
  for (;;) {
    while (i < high && A[R[i]] < pivot) ++i;
    while (j >= low && pivot < A[R[j]]) --j;
    if (i >= j) break;
    SWAPR(i, j);
    ++i;
    --j;
  }
  SWAPR(i, high);
#undef SWAPR
  return i;
}

/* R = argsort(A) in numba's quicksort order (R initialised here) */
void oracle_numba_argsort(const int64_t *A, int64_t n, int32_t *R) {
  for (int64_t i = 0; i < n; ++i) R[i] = (int32_t)i;
  if (n < 2) return;
  int64_t stack_lo[100], stack_hi[100];
  int top = 0;
  stack_lo[0] = 0;
  stack_hi[0] = n - 1;
  top = 1;
  while (top > 0) {
    --top;
    int64_t low = stack_lo[top], high = stack_hi[top];
    while (high - low >= 15) {
      int64_t i = numba_partition_argsort(A, R, low, high);
      if (high - i > i - low) {
        if (high > i) {
          stack_lo[top] = i + 1;
          stack_hi[top] = high;
          ++top;
        }
        high = i - 1;
      } else {
        if (i > low) {
          stack_lo[top] = low;
          stack_hi[top] = i - 1;
          ++top;
        }
        low = i + 1;
      }
    }
    numba_insertion_argsort(A, R, low, high);
  }
}

int oracle_exact_variance(const float *var_volume, int ndim,
                          const int64_t *vol_shape, const int *block, int n,
                          const int64_t *abs_pos, const float *group_mat,
                          const float *const *spatial_mats, int planes,
                          float *var_out) {
  int P = (int)prod_int(block, ndim);
  int64_t mn[ORACLE_MAX_NDIM], ushape[ORACLE_MAX_NDIM], ustr[ORACLE_MAX_NDIM],
      vstr[ORACLE_MAX_NDIM];
  for (int d = 0; d < ndim; ++d) {
    mn[d] = abs_pos[d];
    for (int i = 1; i < n; ++i)
      if (abs_pos[i * ndim + d] < mn[d]) mn[d] = abs_pos[i * ndim + d];
  }
  for (int d = 0; d < ndim; ++d) {
    int64_t mx = abs_pos[d] - mn[d];
    for (int i = 1; i < n; ++i)
      if (abs_pos[i * ndim + d] - mn[d] > mx) mx = abs_pos[i * ndim + d] - mn[d];
    ushape[d] = mx + block[d];
  }
  c_strides(ndim, ushape, ustr);
  c_strides(ndim, vol_shape, vstr);
  int64_t ucells = 1;
  for (int d = 0; d < ndim; ++d) ucells *= ushape[d];
  int entries = n * P;
  int64_t *label = (int64_t *)malloc(sizeof(int64_t) * (size_t)entries);
  for (int i = 0; i < n; ++i)
    for (int p = 0; p < P; ++p) {
      int lc[ORACLE_MAX_NDIM];
      local_coords(ndim, block, p, lc);
      int64_t l = 0;
      for (int d = 0; d < ndim; ++d) l += (abs_pos[i * ndim + d] - mn[d] + lc[d]) * ustr[d];
      label[i * P + p] = l;
    }
  /* stable counting sort by label: entries of one union voxel stay in
   * increasing entry (member) order */
  int32_t *cnt = (int32_t *)calloc((size_t)ucells + 1, sizeof(int32_t));
  for (int e = 0; e < entries; ++e) cnt[label[e] + 1]++;
  for (int64_t c = 0; c < ucells; ++c) cnt[c + 1] += cnt[c];
  int32_t *order = (int32_t *)malloc(sizeof(int32_t) * (size_t)entries);
  int32_t *fill = (int32_t *)malloc(sizeof(int32_t) * (size_t)(ucells + 1));
  memcpy(fill, cnt, sizeof(int32_t) * (size_t)(ucells + 1));
  for (int e = 0; e < entries; ++e) order[fill[label[e]]++] = e;
  if (g_exact_order == 1) oracle_numba_argsort(label, entries, order);

  float *S = spatial_kron(ndim, block, spatial_mats);
  int C = planes * P;
  float *acc = (float *)calloc((size_t)C, sizeof(float));
  float *contrib = (float *)malloc(sizeof(float) * (size_t)C);
  for (int64_t cell = 0; cell < ucells; ++cell) {
    int beg = cnt[cell], end = cnt[cell + 1];
    if (beg == end) continue;
    for (int c = 0; c < C; ++c) contrib[c] = 0.0f;
    for (int pos = beg; pos < end; ++pos) {
      int e = order[pos];
      int i = e / P, p = e % P;
      for (int q = 0; q < planes; ++q) {
        float g = group_mat[q * n + i];
        for (int s = 0; s < P; ++s) contrib[q * P + s] += g * S[s * P + p];
      }
    }
    /* union voxel variance at abs = union coords + group origin */
    int64_t t = cell, off = 0;
    for (int d = ndim - 1; d >= 0; --d) {
      int64_t u = t % ushape[d];
      t /= ushape[d];
      off += (u + mn[d]) * vstr[d];
    }
    float vv = var_volume[off];
    for (int c = 0; c < C; ++c) {
      float v = contrib[c];
      acc[c] += (vv * v) * v;
    }
  }
  memcpy(var_out, acc, sizeof(float) * (size_t)C);
  free(label);
  free(cnt);
  free(order);
  free(fill);
  free(S);
  free(acc);
  free(contrib);
  return 0;
}

static double np_pairwise_f64(const double *a, int64_t n);

/* Exact covariance-aware group risk of the windowed-synthesis, group-scope
 * weights (wiener.py:638-844; the plan of variance.py:174-450, 1413-1449;
 * the metrics of variance.py:1516-1569), for the diagonal group metric of
 * the orthonormal Haar / DCT member transforms and no mass functional:
 *   contributions C_x[q][s] = sum over the entries (i,p) of union voxel x,
 *     in the exact variance's entry order, of f32(G[q,i] S[s,p])
 *     (variance.py:1038-1077), q < planes;
 *   F = C * gain[q][s] (float32, wiener.py:675-680);
 *   quad[x][q] = sum_s' (double)F[s'] (double)(F @ SM)[s'] with F @ SM a
 *     float32 product (wiener.py:505-523);
 *   noise = sum_x var_x sum_q quad[x][q] gm[q] (float64, :526-535, 708-709);
 *   planes q >= `planes`: + sum f64((cm[q,s] gain^2) sigma^2) (:778-787);
 *   with the signal power S: + sum f64(cm (1 - gain)^2 S) (:788-795);
 * where SM = Sinv^T diag(w^2) Sinv (float32; Sinv the Kronecker synthesis
 * matrix, w the float32 window), gm = diag(Ginv^T Ginv) and cm[q,s] =
 * gm[q] SM[s][s] (float32).  gain, sigma and spow hold n*P values; spow may
 * be NULL (no bias term).  Returns the risk. */
double oracle_exact_group_risk(const float *var_volume, int ndim, const int64_t *vol_shape,
                               const int *block, int n, const int64_t *abs_pos,
                               const float *group_mat, const float *group_inv,
                               const float *const *spatial_mats,
                               const float *const *spatial_inv, const float *window,
                               int planes, const float *gain, const float *sigma,
                               const float *spow) {
  int P = (int)prod_int(block, ndim);
  if (planes > n) planes = n;
  if (planes < 0) planes = 0;
  /* metrics (variance.py:1543-1561), float64 sums rounded to float32 */
  float *Sinv = spatial_kron(ndim, block, spatial_inv);
  float *SM = (float *)malloc(sizeof(float) * (size_t)P * P);
  for (int a = 0; a < P; ++a)
    for (int b = 0; b < P; ++b) {
      double acc = 0.0;
      for (int x = 0; x < P; ++x) {
        const float w2 = window[x] * window[x];
        acc += (double)Sinv[x * P + a] * (double)(w2 * Sinv[x * P + b]);
      }
      SM[a * P + b] = (float)acc;
    }
  float *gm = (float *)malloc(sizeof(float) * (size_t)n);
  for (int q = 0; q < n; ++q) {
    double acc = 0.0;
    for (int x = 0; x < n; ++x)
      acc += (double)group_inv[x * n + q] * (double)group_inv[x * n + q];
    gm[q] = (float)acc;
  }
  double noise = 0.0;
  if (planes > 0) {
    int64_t mn[ORACLE_MAX_NDIM], ushape[ORACLE_MAX_NDIM], ustr[ORACLE_MAX_NDIM],
        vstr[ORACLE_MAX_NDIM];
    for (int d = 0; d < ndim; ++d) {
      mn[d] = abs_pos[d];
      for (int i = 1; i < n; ++i)
        if (abs_pos[i * ndim + d] < mn[d]) mn[d] = abs_pos[i * ndim + d];
    }
    for (int d = 0; d < ndim; ++d) {
      int64_t mx = abs_pos[d] - mn[d];
      for (int i = 1; i < n; ++i)
        if (abs_pos[i * ndim + d] - mn[d] > mx) mx = abs_pos[i * ndim + d] - mn[d];
      ushape[d] = mx + block[d];
    }
    c_strides(ndim, ushape, ustr);
    c_strides(ndim, vol_shape, vstr);
    int64_t ucells = 1;
    for (int d = 0; d < ndim; ++d) ucells *= ushape[d];
    int entries = n * P;
    int64_t *label = (int64_t *)malloc(sizeof(int64_t) * (size_t)entries);
    for (int i = 0; i < n; ++i)
      for (int p = 0; p < P; ++p) {
        int lc[ORACLE_MAX_NDIM];
        local_coords(ndim, block, p, lc);
        int64_t l = 0;
        for (int d = 0; d < ndim; ++d) l += (abs_pos[i * ndim + d] - mn[d] + lc[d]) * ustr[d];
        label[i * P + p] = l;
      }
    int32_t *cnt = (int32_t *)calloc((size_t)ucells + 1, sizeof(int32_t));
    for (int e = 0; e < entries; ++e) cnt[label[e] + 1]++;
    for (int64_t c = 0; c < ucells; ++c) cnt[c + 1] += cnt[c];
    int32_t *order = (int32_t *)malloc(sizeof(int32_t) * (size_t)entries);
    int32_t *fill = (int32_t *)malloc(sizeof(int32_t) * (size_t)(ucells + 1));
    memcpy(fill, cnt, sizeof(int32_t) * (size_t)(ucells + 1));
    for (int e = 0; e < entries; ++e) order[fill[label[e]]++] = e;
    if (g_exact_order == 1) oracle_numba_argsort(label, entries, order);
    float *S = spatial_kron(ndim, block, spatial_mats);
    float *F = (float *)malloc(sizeof(float) * (size_t)planes * P);
    float *proj = (float *)malloc(sizeof(float) * (size_t)P);
    for (int64_t cell = 0; cell < ucells; ++cell) {
      int beg = cnt[cell], end = cnt[cell + 1];
      if (beg == end) continue;
      for (int c = 0; c < planes * P; ++c) F[c] = 0.0f;
      for (int pos = beg; pos < end; ++pos) {
        int e = order[pos];
        int i = e / P, p = e % P;
        for (int q = 0; q < planes; ++q) {
          float g = group_mat[q * n + i];
          for (int s = 0; s < P; ++s) F[q * P + s] += g * S[s * P + p];
        }
      }
      int64_t t = cell, off = 0;
      for (int d = ndim - 1; d >= 0; --d) {
        int64_t u = t % ushape[d];
        t /= ushape[d];
        off += (u + mn[d]) * vstr[d];
      }
      double row = 0.0;
      for (int q = 0; q < planes; ++q) {
        float *f = F + q * P;
        for (int s = 0; s < P; ++s) f[s] = f[s] * gain[q * P + s];
        for (int b = 0; b < P; ++b) {
          float acc = 0.0f;
          for (int a = 0; a < P; ++a) acc += f[a] * SM[a * P + b];
          proj[b] = acc;
        }
        double quad = 0.0;
        for (int b = 0; b < P; ++b) quad += (double)f[b] * (double)proj[b];
        row += quad * (double)gm[q];
      }
      noise += row * (double)var_volume[off];
    }
    free(label);
    free(cnt);
    free(order);
    free(fill);
    free(S);
    free(F);
    free(proj);
  }
  const int rest = (n - planes) * P;
  double *tmp = (double *)malloc(sizeof(double) * (size_t)(n * P + 1));
  if (planes < n) {
    for (int e = 0; e < rest; ++e) {
      const int q = planes + e / P, s = e % P;
      const float cm = gm[q] * SM[s * P + s];
      const float g = gain[planes * P + e];
      const float nv = sigma[planes * P + e] * sigma[planes * P + e];
      tmp[e] = (double)((cm * (g * g)) * nv);
    }
    noise += np_pairwise_f64(tmp, rest);
  }
  double bias = 0.0;
  if (spow) {
    for (int e = 0; e < n * P; ++e) {
      const int q = e / P, s = e % P;
      const float cm = gm[q] * SM[s * P + s];
      const float r = 1.0f - gain[e];
      tmp[e] = (double)(cm * ((r * r) * spow[e]));
    }
    bias = np_pairwise_f64(tmp, (int64_t)n * P);
  }
  free(tmp);
  free(Sinv);
  free(SM);
  free(gm);
  return noise + bias;
}

/* plane limiting, variance.py:1572-1631 (float32 sequential plane sums) */
void oracle_limit_planes(const float *exact_std, const float *approx_std,
                         int n, int p, int k, float *out) {
  int epc = k < n ? k : n;
  if (epc == 0) {
    for (int i = 0; i < n * p; ++i) {
      float a = approx_std[i] > 0.0f ? approx_std[i] : 0.0f;
      out[i] = sqrtf(a * a);
    }
    return;
  }
  if (epc == n) {
    for (int i = 0; i < n * p; ++i) {
      float e = exact_std[i] > 0.0f ? exact_std[i] : 0.0f;
      out[i] = sqrtf(e * e);
    }
    return;
  }
  for (int s = 0; s < p; ++s) {
    float tot_a = 0.0f, tot_e = 0.0f, rem_a = 0.0f;
    for (int q = 0; q < n; ++q) {
      float a = approx_std[q * p + s] > 0.0f ? approx_std[q * p + s] : 0.0f;
      tot_a += a * a;
      if (q >= epc) rem_a += a * a;
    }
    for (int q = 0; q < epc; ++q) {
      float e = exact_std[q * p + s] > 0.0f ? exact_std[q * p + s] : 0.0f;
      tot_e += e * e;
    }
    float target = tot_a - tot_e;
    if (target < 0.0f) target = 0.0f;
    float scale = rem_a > 0.0f ? target / rem_a : 0.0f;
    float fallback = target / (float)(n - epc);
    for (int q = 0; q < n; ++q) {
      float v;
      if (q < epc) {
        float e = exact_std[q * p + s] > 0.0f ? exact_std[q * p + s] : 0.0f;
        v = e * e;
      } else {
        float a = approx_std[q * p + s] > 0.0f ? approx_std[q * p + s] : 0.0f;
        v = (a * a) * scale;
        if (!(rem_a > 0.0f)) v = fallback;
      }
      out[q * p + s] = sqrtf(v > 0.0f ? v : 0.0f);
    }
  }
}

int oracle_noise_std(const float *var_volume, int ndim,
                     const int64_t *vol_shape, const int *block, int n,
                     const int64_t *abs_pos, const float *group_mat,
                     const float *const *spatial_mats, int k,
                     float *sigma_out) {
  int P = (int)prod_int(block, ndim);
  int64_t vstr[ORACLE_MAX_NDIM];
  c_strides(ndim, vol_shape, vstr);
  /* approximate planes: squared transforms of the variance patches
   * (variance.py:613-623, 1647-1653) */
  int dims[ORACLE_MAX_NDIM + 1];
  dims[0] = n;
  for (int d = 0; d < ndim; ++d) dims[d + 1] = block[d];
  float *approx = (float *)malloc(sizeof(float) * (size_t)n * P);
  for (int i = 0; i < n; ++i) {
    int64_t base = 0;
    for (int d = 0; d < ndim; ++d) base += abs_pos[i * ndim + d] * vstr[d];
    for (int p = 0; p < P; ++p) {
      int lc[ORACLE_MAX_NDIM];
      local_coords(ndim, block, p, lc);
      int64_t o = base;
      for (int d = 0; d < ndim; ++d) o += lc[d] * vstr[d];
      float v = var_volume[o];
      approx[i * P + p] = v > 0.0f ? v : 0.0f;
    }
  }
  float *sq[ORACLE_MAX_NDIM + 1];
  sq[0] = (float *)malloc(sizeof(float) * (size_t)n * n);
  for (int i = 0; i < n * n; ++i) sq[0][i] = group_mat[i] * group_mat[i];
  for (int d = 0; d < ndim; ++d) {
    int b = block[d];
    if (!spatial_mats[d]) {
      sq[d + 1] = NULL;
      continue;
    }
    sq[d + 1] = (float *)malloc(sizeof(float) * (size_t)b * b);
    for (int i = 0; i < b * b; ++i) sq[d + 1][i] = spatial_mats[d][i] * spatial_mats[d][i];
  }
  oracle_apply_transform(approx, ndim + 1, dims, (const float *const *)sq);
  for (int i = 0; i < n * P; ++i) approx[i] = sqrtf(approx[i] > 0.0f ? approx[i] : 0.0f);
  for (int d = 0; d <= ndim; ++d) free(sq[d]);

  if (k == 0) {
    memcpy(sigma_out, approx, sizeof(float) * (size_t)n * P);
    free(approx);
    return 0;
  }
  int planes = k < n ? k : n;
  float *exact = (float *)malloc(sizeof(float) * (size_t)planes * P);
  oracle_exact_variance(var_volume, ndim, vol_shape, block, n, abs_pos, group_mat,
                        spatial_mats, planes, exact);
  for (int i = 0; i < planes * P; ++i) exact[i] = sqrtf(exact[i] > 0.0f ? exact[i] : 0.0f);
  if (planes == n) {
    memcpy(sigma_out, exact, sizeof(float) * (size_t)n * P);
  } else {
    oracle_limit_planes(exact, approx, n, P, k, sigma_out);
  }
  free(exact);
  free(approx);
  return 0;
}

/* Stationary Gaussian group noise variance (variance.py:1889-2106).
 * maps: P cross-covariance maps over the nf domain of D cells each (entry 0
 * the coefficient's variance); pair_lag: n x n flat domain lags of the member
 * pairs; gmat: the n x n member analysis matrix.  The first q = min(k, n)
 * planes: for each coefficient the serial float32 sum over the nonzero row
 * entries f (increasing) of w_f c[lag(f,f)] w_f, then of
 * w_f w_s (c[lag(f,s)] + c[lag(s,f)]) for the nonzero s < f, floored at 0;
 * the remaining planes share (n pv - sum of the exact planes) / (n - q),
 * floored at 0; k = 0 gives pv everywhere.  var_out is n x P. */
void oracle_gaussian_group_variance(const float *maps, int P, int64_t D, const int64_t *pair_lag,
                                    const float *gmat, int n, int k, float *var_out) {
  const int q = k < n ? k : n;
  for (int s = 0; s < P; ++s) {
    const float *cm = maps + (int64_t)s * D;
    const float pv = cm[0];
    if (q == 0) {
      for (int i = 0; i < n; ++i) var_out[(int64_t)i * P + s] = pv;
      continue;
    }
    float tot = 0.0f;
    for (int r = 0; r < q; ++r) {
      const float *row = gmat + (int64_t)r * n;
      float v = 0.0f;
      for (int f = 0; f < n; ++f) {
        const float wf = row[f];
        if (wf == 0.0f) continue;
        v = v + (wf * cm[pair_lag[(int64_t)f * n + f]]) * wf;
        for (int t = 0; t < f; ++t) {
          const float wt = row[t];
          if (wt == 0.0f) continue;
          const float pair = cm[pair_lag[(int64_t)f * n + t]] + cm[pair_lag[(int64_t)t * n + f]];
          v = v + (wf * wt) * pair;
        }
      }
      if (v < 0.0f) v = 0.0f;
      var_out[(int64_t)r * P + s] = v;
      tot = tot + v;  /* np.sum over the plane axis: sequential in the plane */
    }
    if (q < n) {
      float rem = ((float)n * pv - tot) / (float)(n - q);
      if (rem < 0.0f) rem = 0.0f;
      for (int i = q; i < n; ++i) var_out[(int64_t)i * P + s] = rem;
    }
  }
}

/* Gaussian hard threshold (bmndalgo.py:464-474): keep |C| >= lambda sigma,
 * the group's first coefficient always; atten gets the mask. */
void oracle_ht_gaussian(float *coef, const float *sigma, int64_t n, float lambda_,
                        float *atten) {
  for (int64_t i = 0; i < n; ++i) {
    const float thr = lambda_ * sigma[i];
    const float keep = (fabsf(coef[i]) >= thr || i == 0) ? 1.0f : 0.0f;
    coef[i] = coef[i] * keep;
    atten[i] = keep;
  }
}

/* Gaussian Wiener gain (wiener.py:56-75): |R|^2 / (|R|^2 + mu max(sigma^2,
 * 1e-8)), 0 where the denominator is not positive, clipped to [0, 1];
 * coef takes it; spow gets |R|^2, the risk model's signal power
 * (bmndalgo.py:889-899). */
void oracle_wiener_gaussian(float *coef, const float *ref_coef, const float *sigma, int64_t n,
                            float mu, float *atten, float *spow) {
  const float floor_ = (float)1e-8;
  for (int64_t i = 0; i < n; ++i) {
    const float r = fabsf(ref_coef[i]);
    const float pr = r * r;
    float nv = sigma[i] * sigma[i];
    if (nv < floor_) nv = floor_;
    const float den = pr + mu * nv;
    float gain = den > 0.0f ? pr / den : 0.0f;
    gain = gain < 0.0f ? 0.0f : (gain > 1.0f ? 1.0f : gain);
    coef[i] = coef[i] * gain;
    atten[i] = gain;
    if (spow) spow[i] = pr;
  }
}

/* per-patch inverse risk weights, VARIANCE/COEFFICIENT/PATCH
 * (aggregationweights.py:213-259, 616-653): risk_i = sum_j Ginv[i,j]^2 *
 * sum_c a[j,c]^2 sigma[j,c]^2 in float64; w_i = f32(1/(risk_i + eps)). */
void oracle_patch_weights(const float *atten, const float *sigma, int n,
                          int p, const float *group_inv, double eps,
                          float *w_out) {
  double plane[64];
  for (int j = 0; j < n; ++j) {
    double risk = 0.0;
    for (int c = 0; c < p; ++c) {
      double a = (double)atten[j * p + c];
      float nv = sigma[j * p + c] * sigma[j * p + c];
      double err = a * a;
      err *= (double)nv;
      risk += err;
    }
    plane[j] = risk;
  }
  for (int i = 0; i < n; ++i) {
    double risk = 0.0;
    for (int j = 0; j < n; ++j) {
      double m = (double)group_inv[i * n + j];
      risk += (m * m) * plane[j];
    }
    if (!isfinite(risk) || risk < 0.0)
      w_out[i] = 0.0f;
    else
      w_out[i] = (float)(1.0 / (risk + eps));
  }
}

/* numpy's float64 pairwise sum of a contiguous array (np.sum(x,
 * dtype=np.float64)): the float32 routine's tree in float64 */
static double np_pairwise_f64(const double *a, int64_t n) {
  if (n < 8) {
    double r = 0.0;
    for (int64_t i = 0; i < n; ++i) r += a[i];
    return r;
  }
  if (n <= 128) {
    double r[8];
    for (int j = 0; j < 8; ++j) r[j] = a[j];
    int64_t i = 8;
    for (; i < n - (n % 8); i += 8)
      for (int j = 0; j < 8; ++j) r[j] += a[i + j];
    double res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
    for (; i < n; ++i) res += a[i];
    return res;
  }
  int64_t n2 = n / 2;
  n2 -= n2 % 8;
  return np_pairwise_f64(a, n2) + np_pairwise_f64(a + n2, n - n2);
}

double oracle_np_sum_f64(const double *a, int64_t n) {
  double r = 0.0;
  for (int64_t i = 0; i < n; i += 8192) r += np_pairwise_f64(a + i, n - i < 8192 ? n - i : 8192);
  return r;
}

/* Aggregation risks of the diagonal filter models (aggregationweights.py:
 * 123-211 group scope, 214-259 patch scope): per coefficient the error
 * a^2 (model 1, classic), a^2 sigma^2 (0, variance) or a^2 sigma^2 +
 * (1 - a)^2 S (2, risk; S the signal power), in float64 from the float32
 * inputs; in the windowed-synthesis domain times the spatial metric
 * diagonal smd[c] (and, group scope, the group metric diagonal gmd[i] first).
 * Patch scope: risk_i = sum_j Ginv[i,j]^2 * sum_c error[j,c] (sequential
 * sums); group scope: one risk, numpy's float64 sum of all errors.  Returns
 * the number of risks written (n or 1). */
int oracle_weight_risks(const float *atten, const float *sigma, const float *spow, int n, int p,
                        const float *group_inv, const double *smd, const double *gmd,
                        int model, int windowed, int group_scope, double *risks) {
  double *err = (double *)malloc(sizeof(double) * (size_t)n * p);
  for (int j = 0; j < n; ++j) {
    for (int c = 0; c < p; ++c) {
      const int e = j * p + c;
      const double a = (double)atten[e];
      const float nv = sigma[e] * sigma[e];
      double x = a * a;
      if (model != 1) x *= (double)nv;
      if (model == 2) {
        const double rej = 1.0 - a;
        x += (rej * rej) * (double)spow[e];
      }
      if (windowed) {
        if (group_scope) x *= gmd[j];
        x *= smd[c];
      }
      err[e] = x;
    }
  }
  int out = 1;
  if (group_scope) {
    risks[0] = oracle_np_sum_f64(err, (int64_t)n * p);
  } else {
    double plane[64];
    for (int j = 0; j < n; ++j) {
      double r = 0.0;
      for (int c = 0; c < p; ++c) r += err[j * p + c];
      plane[j] = r;
    }
    for (int i = 0; i < n; ++i) {
      double r = 0.0;
      for (int j = 0; j < n; ++j) {
        const double m = (double)group_inv[i * n + j];
        r += (m * m) * plane[j];
      }
      risks[i] = r;
    }
    out = n;
  }
  free(err);
  return out;
}

/* weights from oracle_weight_risks: f32(1/(risk + eps)) per member, the
 * group's one risk for every member in group scope, 0 for a risk that is
 * not finite or negative (aggregationweights.py:617-687) */
void oracle_weights_ex(const float *atten, const float *sigma, const float *spow, int n, int p,
                       const float *group_inv, const double *smd, const double *gmd, int model,
                       int windowed, int group_scope, double eps, float *w_out) {
  double risks[64];
  const int nr = oracle_weight_risks(atten, sigma, spow, n, p, group_inv, smd, gmd, model,
                                     windowed, group_scope, risks);
  for (int i = 0; i < n; ++i) {
    const double r = risks[nr == 1 ? 0 : i];
    w_out[i] = (!isfinite(r) || r < 0.0) ? 0.0f : (float)(1.0 / (r + eps));
  }
}

/* ======================================================================== */
/* shrinkage and scatter components                                          */
/* ======================================================================== */

/* soft threshold (bmndalgo.py:475-483): s = max(|C| - thr, 0) with
 * thr = float32(lambda * sigma), atten = s / |C| (0 where |C| = 0), C =
 * sign(C) * s with numpy's sign (0 at +-0), all float32 */
void oracle_ht_soft(float *coef, const float *sigma, int64_t n, float lambda_, float *atten) {
  for (int64_t e = 0; e < n; ++e) {
    const float thr = lambda_ * sigma[e];
    const float mag = fabsf(coef[e]);
    float shr = mag - thr;
    shr = shr > 0.0f ? shr : 0.0f;
    atten[e] = mag > 0.0f ? shr / mag : 0.0f;
    const float c = coef[e];
    coef[e] = c > 0.0f ? shr : (c < 0.0f ? -shr : 0.0f);
  }
}

/* Poisson hard threshold (poisson.py:30-46 with bmndalgo.py:460-485): keep a
 * coefficient iff |C|/cs > thr/cs + 128*eps32*max(|C|/cs, thr/cs, 1), compared
 * in float64, with thr = float32(lambda * sigma). atten gets 1.0f / 0.0f and
 * coef is multiplied by it. */
void oracle_ht_shrink(float *coef, const float *sigma, int64_t n, float lambda_,
                      double count_scale, float *atten) {
  for (int64_t e = 0; e < n; ++e) {
    float t = lambda_ * sigma[e];
    double cc = (double)fabsf(coef[e]) / count_scale;
    double tc = (double)t / count_scale;
    double sc = cc > tc ? cc : tc;
    if (sc < 1.0) sc = 1.0;
    double tol = (128.0 * F32_EPS) * sc;
    float keep = (cc > tc + tol) ? 1.0f : 0.0f;
    atten[e] = keep;
    coef[e] = coef[e] * keep;
  }
}

/* variance-scaled Poisson Wiener gain (wiener.py:24-53, 78-95):
 * S = max(R^2 - mu*sigma^2, 0), g = clip(S / (S + mu*sigma^2 + 1e-10), 0, 1),
 * all float32; coef (the noisy coefficients) is multiplied by g. */
void oracle_wiener_shrink(float *coef, const float *ref_coef, const float *sigma,
                          int64_t n, float mu, float *atten) {
  oracle_wiener_shrink_mode(coef, ref_coef, sigma, n, mu, 0, atten);
}

/* wiener.py:24-53, 78-95: signal power |r|^2 less mu sigma^2 (mode 0,
 * variance-scaled), kept whole (1, classic) or less sigma^2 (2, noise-floor),
 * floored at 0; gain = power / (power + mu sigma^2 + eps) clipped to [0, 1]. */
static void wiener_shrink_power(float *coef, const float *ref_coef, const float *sigma,
                                int64_t n, float mu, int mode, float *atten, float *spow_out);

void oracle_wiener_shrink_mode(float *coef, const float *ref_coef, const float *sigma,
                               int64_t n, float mu, int mode, float *atten) {
  wiener_shrink_power(coef, ref_coef, sigma, n, mu, mode, atten, NULL);
}

/* as oracle_wiener_shrink_mode, also writing the signal power (the risk
 * weight model's S, wiener.py:24-53) into spow_out unless NULL */
static void wiener_shrink_power(float *coef, const float *ref_coef, const float *sigma,
                                int64_t n, float mu, int mode, float *atten, float *spow_out) {
  const float eps32 = (float)1e-10;
  for (int64_t e = 0; e < n; ++e) {
    float nv = sigma[e] * sigma[e];
    float eff = mu * nv;
    float pr = fabsf(ref_coef[e]) * fabsf(ref_coef[e]);
    float spow = pr;
    if (mode != 1) {
      spow = pr - (mode == 2 ? nv : eff);
      if (spow < 0.0f) spow = 0.0f;
    }
    if (spow_out) spow_out[e] = spow;
    float gain = spow / ((spow + eff) + eps32);
    if (gain < 0.0f) gain = 0.0f;
    if (gain > 1.0f) gain = 1.0f;
    atten[e] = gain;
    coef[e] = coef[e] * gain;
  }
}

/* sequential scatter of one group (patchaggregation.py:170-199):
 * num[x_i + p] += filt[i,p] * (w_i * k[p]); den[x_i + p] += w_i * k[p]. */
void oracle_scatter_add(float *num, float *den, int ndim, const int64_t *vol_shape,
                        const int *block, const float *kaiser, const float *filt,
                        const float *weights, const int64_t *lins, int n) {
  int64_t counts[ORACLE_MAX_NDIM], vstr[ORACLE_MAX_NDIM];
  for (int d = 0; d < ndim; ++d) counts[d] = vol_shape[d] - block[d] + 1;
  c_strides(ndim, vol_shape, vstr);
  int P = (int)prod_int(block, ndim);
  for (int i = 0; i < n; ++i) {
    int64_t lin = lins[i], base = 0;
    for (int d = ndim - 1; d >= 0; --d) {
      base += (lin % counts[d]) * vstr[d];
      lin /= counts[d];
    }
    for (int p = 0; p < P; ++p) {
      int lc[ORACLE_MAX_NDIM];
      local_coords(ndim, block, p, lc);
      int64_t v = base;
      for (int d = 0; d < ndim; ++d) v += lc[d] * vstr[d];
      float ww = weights[i] * kaiser[p];
      num[v] += filt[(int64_t)i * P + p] * ww;
      den[v] += ww;
    }
  }
}

/* numpy's float32 summation order for a contiguous array (np.sum / np.mean,
 * the pairwise sum of numpy's umath loops: blocks of up to 128 elements
 * summed with 8 interleaved accumulators, larger ranges halved at a multiple
 * of 8), the order the reference's np.mean calls use
 * (sigmaestimation.py:62-71, bmndalgo.py:875). */
static float np_pairwise_f32(const float *a, int64_t n) {
  if (n < 8) {
    float r = 0.0f;
    for (int64_t i = 0; i < n; ++i) r += a[i];
    return r;
  }
  if (n <= 128) {
    float r[8];
    for (int j = 0; j < 8; ++j) r[j] = a[j];
    int64_t i = 8;
    for (; i < n - (n % 8); i += 8)
      for (int j = 0; j < 8; ++j) r[j] += a[i + j];
    float res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
    for (; i < n; ++i) res += a[i];
    return res;
  }
  int64_t n2 = n / 2;
  n2 -= n2 % 8;
  return np_pairwise_f32(a, n2) + np_pairwise_f32(a + n2, n - n2);
}

/* np.sum: the reduction visits the array in buffer-sized chunks of 8192
 * elements, each summed pairwise and added in turn to the running total. */
float oracle_np_sum_f32(const float *a, int64_t n) {
  float r = 0.0f;
  for (int64_t i = 0; i < n; i += 8192) r += np_pairwise_f32(a + i, n - i < 8192 ? n - i : 8192);
  return r;
}

/* np.mean of a float32 array: the float32 sum divided by the count (in
 * float64 and rounded, which for a float32 sum equals the correctly rounded
 * float32 quotient) */
float oracle_np_mean_f32(const float *a, int64_t n) {
  return (float)((double)oracle_np_sum_f32(a, n) / (double)n);
}

/* oracle_scatter_add with the sigma-weighted sum of the reference's sigma map
 * (patchaggregation.py:229-294): sigma_num[x_i + p] += s(i,p) * (w_i * k[p]),
 * s(i,p) = group_sigma[i,p], or the scalar sigma when group_sigma is NULL. */
void oracle_scatter_add_sigma(float *num, float *den, float *sigma_num, int ndim,
                              const int64_t *vol_shape, const int *block, const float *kaiser,
                              const float *filt, const float *weights, const float *group_sigma,
                              float sigma, const int64_t *lins, int n) {
  int64_t counts[ORACLE_MAX_NDIM], vstr[ORACLE_MAX_NDIM];
  for (int d = 0; d < ndim; ++d) counts[d] = vol_shape[d] - block[d] + 1;
  c_strides(ndim, vol_shape, vstr);
  int P = (int)prod_int(block, ndim);
  for (int i = 0; i < n; ++i) {
    int64_t lin = lins[i], base = 0;
    for (int d = ndim - 1; d >= 0; --d) {
      base += (lin % counts[d]) * vstr[d];
      lin /= counts[d];
    }
    for (int p = 0; p < P; ++p) {
      int lc[ORACLE_MAX_NDIM];
      local_coords(ndim, block, p, lc);
      int64_t v = base;
      for (int d = 0; d < ndim; ++d) v += lc[d] * vstr[d];
      float ww = weights[i] * kaiser[p];
      num[v] += filt[(int64_t)i * P + p] * ww;
      den[v] += ww;
      const float sv = group_sigma ? group_sigma[(int64_t)i * P + p] : sigma;
      sigma_num[v] += sv * ww;
    }
  }
}

/* ======================================================================== */
/* full two-stage pipeline (bmndalgo.py:71-1063, Poisson branches)            */
/* ======================================================================== */

typedef struct {
  int ndim;
  int64_t shape[ORACLE_MAX_NDIM], vstr[ORACLE_MAX_NDIM];
  int64_t counts[ORACLE_MAX_NDIM], cstr[ORACLE_MAX_NDIM];
  int block[ORACLE_MAX_NDIM];
  int P;
  int64_t *poff;
  float *kaiser;
  float *sp_fwd[ORACLE_MAX_NDIM], *sp_inv[ORACLE_MAX_NDIM];
} stage_geom;

static void geom_init(stage_geom *g, const oracle_params *prm, const int *block,
                      double beta, int wiener) {
  g->ndim = prm->ndim;
  for (int d = 0; d < g->ndim; ++d) {
    g->shape[d] = prm->shape[d];
    g->block[d] = block[d];
    g->counts[d] = prm->shape[d] - block[d] + 1;
  }
  c_strides(g->ndim, g->shape, g->vstr);
  c_strides(g->ndim, g->counts, g->cstr);
  g->P = (int)prod_int(block, g->ndim);
  g->poff = (int64_t *)malloc(sizeof(int64_t) * (size_t)g->P);
  for (int p = 0; p < g->P; ++p) {
    int lc[ORACLE_MAX_NDIM];
    local_coords(g->ndim, block, p, lc);
    int64_t o = 0;
    for (int d = 0; d < g->ndim; ++d) o += lc[d] * g->vstr[d];
    g->poff[p] = o;
  }
  g->kaiser = (float *)malloc(sizeof(float) * (size_t)g->P);
  oracle_kaiser_nd(g->ndim, block, beta, g->kaiser);
  const float *given = wiener ? prm->wi_spatial : prm->ht_spatial;
  int64_t off = 0;
  for (int d = 0; d < g->ndim; ++d) {
    int b = block[d];
    g->sp_fwd[d] = (float *)malloc(sizeof(float) * (size_t)b * b);
    g->sp_inv[d] = (float *)malloc(sizeof(float) * (size_t)b * b);
    if (given) {
      memcpy(g->sp_fwd[d], given + off, sizeof(float) * (size_t)b * b);
      off += (int64_t)b * b;
    } else if (wiener) {
      oracle_dct_matrix(b, g->sp_fwd[d], g->sp_inv[d]);
    } else if (oracle_bior15_matrix(b, g->sp_fwd[d], g->sp_inv[d]) != 0) {
      /* bior1.5 lengths other than 4/8 go through pywt in the reference;
       * not part of this oracle */
      oracle_dct_matrix(b, g->sp_fwd[d], g->sp_inv[d]);
    }
  }
  if (given)
    for (int d = 0; d < g->ndim; ++d) {
      int b = block[d];
      memcpy(g->sp_inv[d], given + off, sizeof(float) * (size_t)b * b);
      off += (int64_t)b * b;
    }
}

static void geom_free(stage_geom *g) {
  free(g->poff);
  free(g->kaiser);
  for (int d = 0; d < g->ndim; ++d) {
    free(g->sp_fwd[d]);
    free(g->sp_inv[d]);
  }
}

static void lin_to_abs(const stage_geom *g, int64_t lin, int64_t *abs_pos) {
  for (int d = g->ndim - 1; d >= 0; --d) {
    abs_pos[d] = lin % g->counts[d];
    lin /= g->counts[d];
  }
}

static int64_t lin_to_vol_base(const stage_geom *g, int64_t lin) {
  int64_t base = 0;
  for (int d = g->ndim - 1; d >= 0; --d) {
    base += (lin % g->counts[d]) * g->vstr[d];
    lin /= g->counts[d];
  }
  return base;
}

/* variance model, variance.py:556-598: max(scale*max(v,0), floor) in float32 */
static void variance_volume(const float *v, int64_t n, double scale, double floor_,
                            float *out) {
  float s = (float)scale, f = (float)floor_;
  for (int64_t i = 0; i < n; ++i) {
    float x = v[i] > 0.0f ? v[i] : 0.0f;
    x = s * x;
    out[i] = x > f ? x : f;
  }
}

/* one stage: block matching + group filtering + sequential aggregation */
/* Group-mass correction of one group (poisson.py:166-232 with
 * patchaggregation.py:71-98): the output functional O = f32(w_i k_p) / den
 * at each member voxel (float32, as float64), its image under the adjoint of
 * the synthesis (M^T along every axis, float64), the target sum O * source,
 * and the correction of the coefficients along the constant-synthesis
 * direction D (the outer product of the per-axis directions dirs[a], scaled
 * by its largest magnitude) that makes the synthesised group carry the
 * target.  coef (float32, n x P) is corrected in place; sums are sequential
 * float64.  dims[0] = n, dims[1..ndim] the block; mats[a] the synthesis
 * matrix of axis a (NULL: identity). */
/* The correction itself for a given output functional O (float64, the
 * group's shape): oracle_mass_correct over nax axes. */
void oracle_mass_correct(int nax, const int *dims, const float *const *mats,
                         const double *const *dirs, const double *O, const float *src_vox,
                         float *coef, double eps) {
  int64_t total = prod_int(dims, nax);
  double *cf = (double *)malloc(sizeof(double) * (size_t)total);
  double *tmp = (double *)malloc(sizeof(double) * (size_t)total);
  double target = 0.0;
  for (int64_t e = 0; e < total; ++e) {
    cf[e] = O[e];
    target += cf[e] * (double)src_vox[e];
  }
  for (int a = 0; a < nax; ++a) {
    const float *m = mats[a];
    if (!m) continue;
    const int n = dims[a];
    int64_t inner = 1;
    for (int d = a + 1; d < nax; ++d) inner *= dims[d];
    const int64_t outer = total / (inner * n);
    for (int64_t o = 0; o < outer; ++o)
      for (int j = 0; j < n; ++j)
        for (int64_t in = 0; in < inner; ++in) {
          double acc = 0.0;
          const double *x = cf + o * n * inner + in;
          for (int k = 0; k < n; ++k) acc += (double)m[k * n + j] * x[k * inner];
          tmp[o * n * inner + j * inner + in] = acc;
        }
    memcpy(cf, tmp, sizeof(double) * (size_t)total);
  }
  double denom = 0.0;
  for (int64_t e = 0; e < total; ++e) denom += cf[e] * cf[e];
  if (denom <= eps) {
    double sc = 0.0;
    for (int64_t e = 0; e < total; ++e) sc = fabs(cf[e]) > sc ? fabs(cf[e]) : sc;
    if (!(sc > 0.0) || !isfinite(sc)) goto done;
    for (int64_t e = 0; e < total; ++e) cf[e] /= sc;
    target /= sc;
  }
  {
    double current = 0.0;
    for (int64_t e = 0; e < total; ++e) current += cf[e] * (double)coef[e];
    const double residual = target - current;
    if (residual == 0.0) goto done;
    double dscale = 1.0;
    for (int a = 0; a < nax; ++a) {
      double mx = 0.0;
      for (int k = 0; k < dims[a]; ++k) mx = fabs(dirs[a][k]) > mx ? fabs(dirs[a][k]) : mx;
      dscale *= mx;
    }
    if (!(dscale > 0.0) || !isfinite(dscale)) goto done;
    double response = 0.0;
    for (int64_t e = 0; e < total; ++e) {
      int64_t r = e;
      double dv = 1.0;
      int idx[ORACLE_MAX_NDIM + 1];
      for (int a = nax - 1; a >= 0; --a) {
        idx[a] = (int)(r % dims[a]);
        r /= dims[a];
      }
      dv = dirs[0][idx[0]];
      for (int a = 1; a < nax; ++a) dv *= dirs[a][idx[a]];
      tmp[e] = dv / dscale;
      response += cf[e] * tmp[e];
    }
    if (!isfinite(response) || response == 0.0) goto done;
    const double corr = residual / response;
    for (int64_t e = 0; e < total; ++e) coef[e] = (float)((double)coef[e] + corr * tmp[e]);
  }
done:
  free(cf);
  free(tmp);
}

static void mass_correct_group(int ndim, const int *dims, const float *const *mats,
                               const double *const *dirs, const float *w, const float *kaiser,
                               const float *den_vox, const float *src_vox, float *coef,
                               double eps) {
  const int nax = ndim + 1;
  const int64_t total = prod_int(dims, nax);
  const int P = (int)(total / dims[0]);
  double *O = (double *)malloc(sizeof(double) * (size_t)total);
  for (int64_t e = 0; e < total; ++e) {
    const int i = (int)(e / P), p = (int)(e % P);
    const float ww = w[i] * kaiser[p];
    O[e] = (double)(ww / den_vox[e]);
  }
  oracle_mass_correct(nax, dims, mats, dirs, O, src_vox, coef, eps);
  free(O);
}

static int run_stage(const oracle_params *prm, int wiener, const float *match_vol,
                     const float *pilot, const float *noisy, const float *var_vol,
                     const float *fallback,
                     float *est, int64_t *sel_dump, int32_t *cnt_dump, int64_t dump_cap,
                     float *sigma_map, float sigma_fill) {
  const int *block = wiener ? prm->wi_block : prm->ht_block;
  const int *step = wiener ? prm->wi_step : prm->ht_step;
  const int *radius = wiener ? prm->wi_radius : prm->ht_radius;
  int max_m = wiener ? prm->wi_max : prm->ht_max;
  int min_m = wiener ? prm->wi_min : prm->ht_min;
  float thr = wiener ? prm->wi_threshold : prm->ht_threshold;
  stage_geom g;
  geom_init(&g, prm, block, wiener ? prm->wi_beta : prm->ht_beta, wiener);
  /* group-mass conservation of this stage (poisson.py:49-265, bmndalgo.py:
   * 263-266, 537-548, 610-626, 704-711, 959-970, 1023-1038) */
  const int mass = wiener ? (prm->mass_mode & 2) != 0 : (prm->mass_mode & 1) != 0;
  int ndim = g.ndim, P = g.P;
  int64_t nvox = 1;
  for (int d = 0; d < ndim; ++d) nvox *= g.shape[d];

  int64_t nrefs = oracle_schedule_generated(ndim, g.counts, block, step, prm->shift_density,
                                            prm->schedule_density, NULL, 0);
  int64_t *refs = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nrefs * ndim + 1));
  oracle_schedule_generated(ndim, g.counts, block, step, prm->shift_density,
                            prm->schedule_density, refs, nrefs);
  float *thrs = (float *)malloc(sizeof(float) * (size_t)(nrefs + 1));
  if (prm->gaussian || prm->poisson_ssd) {
    /* SSD thresholds scale with the matched volume's squared value range
     * (blockmatching.py:168-184): float32 max - min, 1 when not positive,
     * the product in float64, stored as float32 */
    float mx = match_vol[0], mn = match_vol[0];
    for (int64_t v = 1; v < nvox; ++v) {
      if (match_vol[v] > mx) mx = match_vol[v];
      if (match_vol[v] < mn) mn = match_vol[v];
    }
    const float range = mx - mn;
    const double ds = range > 0.0f ? (double)range : 1.0;
    thr = (float)(((wiener ? prm->wi_match_d : prm->ht_match_d) * ds) * ds);
  }
  for (int64_t i = 0; i < nrefs; ++i)
    thrs[i] = (!wiener && prm->ht_thresholds) ? prm->ht_thresholds[i] : thr;
  int64_t *sel = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nrefs * max_m + 1));
  int32_t *cnt = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nrefs + 1));
  const int code = prm->distance ? prm->distance : ORACLE_DIST_DEVIANCE;
  const int std_ht = !wiener && prm->ht_standardize;
  if (prm->gaussian)
    oracle_blockmatch_ssd(match_vol, ndim, g.shape, g.counts, block, refs, nrefs, radius, thrs,
                          max_m, min_m, 1, wiener ? NULL : prm->ht_noise_ssd, prm->ht_gamma,
                          prm->nthreads, sel, cnt, NULL);
  else if (prm->poisson_ssd)
    /* stage 1 takes the Poisson gamma correction, stage 2 plain SSD
     * (bmndalgo.py:292-294, 321-322, 728-741) */
    oracle_blockmatch_ssd_poisson(match_vol, ndim, g.shape, g.counts, block, refs, nrefs, radius,
                                  thrs, max_m, min_m, 1, wiener ? 0.0f : prm->ht_gamma,
                                  prm->count_scale, prm->nthreads, sel, cnt, NULL);
  else
    oracle_blockmatch_standardized(NULL, match_vol, ndim, g.shape, g.counts, block, refs, nrefs,
                                   radius, thrs, max_m, min_m, 1, 1, code, prm->count_scale,
                                   std_ht ? prm->null_means : NULL,
                                   std_ht ? prm->null_vars : NULL, prm->null_max, prm->nthreads,
                                   sel, cnt, NULL);
  if (sel_dump && cnt_dump) {
    int64_t lim = nrefs < dump_cap ? nrefs : dump_cap;
    memcpy(sel_dump, sel, sizeof(int64_t) * (size_t)(lim * max_m));
    memcpy(cnt_dump, cnt, sizeof(int32_t) * (size_t)lim);
  }

  int64_t *off = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nrefs + 1));
  off[0] = 0;
  for (int64_t r = 0; r < nrefs; ++r) off[r + 1] = off[r] + cnt[r];
  float *filt = (float *)malloc(sizeof(float) * (size_t)(off[nrefs] * P + 1));
  float *wts = (float *)malloc(sizeof(float) * (size_t)(off[nrefs] + 1));
  /* sigma map: HT scatters each group's noise std element-wise
   * (bmndalgo.py:576-598), Wiener the mean of it as one value per group
   * (bmndalgo.py:874-875, 1001-1012) */
  float *gsig = NULL, *gsig1 = NULL;
  if (sigma_map) {
    if (wiener) gsig1 = (float *)malloc(sizeof(float) * (size_t)(nrefs + 1));
    else gsig = (float *)malloc(sizeof(float) * (size_t)(off[nrefs] * P + 1));
  }

#ifdef _OPENMP
#pragma omp parallel
#endif
  {
    int nmax = max_m;
    float *gfwd = (float *)malloc(sizeof(float) * (size_t)nmax * nmax);
    float *ginv = (float *)malloc(sizeof(float) * (size_t)nmax * nmax);
    float *X = (float *)malloc(sizeof(float) * (size_t)nmax * P);
    float *Y = (float *)malloc(sizeof(float) * (size_t)nmax * P);
    float *sig = (float *)malloc(sizeof(float) * (size_t)nmax * P);
    float *att = (float *)malloc(sizeof(float) * (size_t)nmax * P);
    float *spw = (float *)malloc(sizeof(float) * (size_t)nmax * P);
    int64_t *apos = (int64_t *)malloc(sizeof(int64_t) * (size_t)nmax * ndim);
    int64_t *lags = (int64_t *)malloc(sizeof(int64_t) * (size_t)nmax * nmax);
    int64_t cov_cells = 1;
    for (int d = 0; d < ndim; ++d) cov_cells *= prm->gaussian ? prm->cov_domain[d] : 1;
    int last_n = -1;
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 8)
#endif
    for (int64_t r = 0; r < nrefs; ++r) {
      int n = cnt[r];
      if (n == 0) continue;
      if (n != last_n) {
        group_matrix(n, gfwd, ginv);
        last_n = n;
      }
      int dims[ORACLE_MAX_NDIM + 1];
      dims[0] = n;
      for (int d = 0; d < ndim; ++d) dims[d + 1] = block[d];
      const float *fm[ORACLE_MAX_NDIM + 1], *im[ORACLE_MAX_NDIM + 1], *spf[ORACLE_MAX_NDIM];
      fm[0] = gfwd;
      im[0] = ginv;
      for (int d = 0; d < ndim; ++d) {
        fm[d + 1] = g.sp_fwd[d];
        im[d + 1] = g.sp_inv[d];
        spf[d] = g.sp_fwd[d];
      }
      for (int i = 0; i < n; ++i) {
        int64_t lin = sel[r * max_m + i];
        lin_to_abs(&g, lin, apos + i * ndim);
        int64_t base = lin_to_vol_base(&g, lin);
        for (int p = 0; p < P; ++p) X[i * P + p] = noisy[base + g.poff[p]];
        if (wiener)
          for (int p = 0; p < P; ++p) Y[i * P + p] = pilot[base + g.poff[p]];
      }
      oracle_apply_transform(X, ndim + 1, dims, fm);
      if (prm->gaussian) {
        /* pair lags of the members on the nf domain (variance.py:1889-1906),
         * the group variance, its square root (bmndalgo.py:443-451) */
        for (int i = 0; i < n; ++i)
          for (int j = 0; j < n; ++j) {
            int64_t l = 0;
            for (int d = 0; d < ndim; ++d) {
              const int64_t D = prm->cov_domain[d];
              int64_t x = (apos[i * ndim + d] - apos[j * ndim + d]) % D;
              if (x < 0) x += D;
              l = l * D + x;
            }
            lags[i * n + j] = l;
          }
        oracle_gaussian_group_variance(wiener ? prm->wi_cov : prm->ht_cov, P, cov_cells, lags,
                                       gfwd, n, prm->k, sig);
        for (int e = 0; e < n * P; ++e) sig[e] = sqrtf(sig[e] > 0.0f ? sig[e] : 0.0f);
      } else {
        oracle_noise_std(var_vol, ndim, g.shape, block, n, apos, gfwd, spf, prm->k, sig);
      }
      if (!wiener) {
        if (prm->ht_soft)
          oracle_ht_soft(X, sig, (int64_t)n * P, prm->ht_lambda, att);
        else if (prm->gaussian)
          oracle_ht_gaussian(X, sig, (int64_t)n * P, prm->ht_lambda, att);
        else
          oracle_ht_shrink(X, sig, (int64_t)n * P, prm->ht_lambda, prm->count_scale, att);
      } else {
        oracle_apply_transform(Y, ndim + 1, dims, fm);
        if (prm->gaussian)
          oracle_wiener_gaussian(X, Y, sig, (int64_t)n * P, prm->wi_mu, att, spw);
        else
          wiener_shrink_power(X, Y, sig, (int64_t)n * P, prm->wi_mu, prm->wi_gain_mode, att,
                              spw);
      }
      if (!mass) oracle_apply_transform(X, ndim + 1, dims, im);
      float *fo = filt + off[r] * P;
      memcpy(fo, X, sizeof(float) * (size_t)n * P);  /* mass: the shrunk coefficients */
      {
        const int wm = wiener ? prm->wi_wmodel : prm->ht_wmodel;
        const int wd = wiener ? prm->wi_windowed : prm->ht_windowed;
        const int ws = wiener ? prm->wi_group_scope : prm->ht_group_scope;
        const double *smd = wiener ? prm->wi_smd : prm->ht_smd;
        const double *gmd = wiener ? prm->wi_gmd : prm->ht_gmd;
        /* the exact covariance-aware group risk replaces the diagonal one
         * for the variance (and, in the Wiener stage, risk) model in the
         * windowed-synthesis domain with group scope and k > 0
         * (bmndalgo.py:359-377, 493-505, 752-758, 913-926) */
        const int exact_risk = !prm->gaussian && prm->k > 0 && wd && ws &&
                               (wm == 0 || (wiener && wm == 2));
        if (exact_risk) {
          const double risk = oracle_exact_group_risk(
              var_vol, ndim, g.shape, block, n, apos, gfwd, ginv, spf,
              (const float *const *)g.sp_inv, g.kaiser, prm->k < n ? prm->k : n, att, sig,
              (wiener && wm == 2) ? spw : NULL);
          float w = 0.0f;
          if (isfinite(risk) && risk >= 0.0) w = (float)(1.0 / (risk + 1e-10));
          for (int i = 0; i < n; ++i) wts[off[r] + i] = w;
        } else {
          oracle_weights_ex(att, sig, spw, n, P, ginv, smd, gmd ? gmd + n * (n - 1) / 2 : NULL,
                            wm, wd, ws, 1e-10, wts + off[r]);
        }
      }
      if (gsig) memcpy(gsig + off[r] * P, sig, sizeof(float) * (size_t)n * P);
      if (gsig1) gsig1[r] = oracle_np_mean_f32(sig, (int64_t)n * P);
    }
    free(gfwd);
    free(ginv);
    free(X);
    free(Y);
    free(sig);
    free(att);
    free(spw);
    free(apos);
    free(lags);
  }

  /* sequential aggregation in schedule order (patchaggregation.py:170-199) */
  float *num = (float *)calloc((size_t)nvox, sizeof(float));
  float *den = (float *)calloc((size_t)nvox, sizeof(float));
  float *snum = sigma_map ? (float *)calloc((size_t)nvox, sizeof(float)) : NULL;
  for (int64_t r = 0; r < nrefs; ++r) {
    if (snum)
      oracle_scatter_add_sigma(num, den, snum, ndim, g.shape, block, g.kaiser, filt + off[r] * P,
                               wts + off[r], gsig ? gsig + off[r] * P : NULL,
                               gsig1 ? gsig1[r] : 0.0f, sel + r * max_m, cnt[r]);
    else
      oracle_scatter_add(num, den, ndim, g.shape, block, g.kaiser, filt + off[r] * P,
                         wts + off[r], sel + r * max_m, cnt[r]);
  }
  if (mass) {
    /* the denominator stands; every group's coefficients are corrected
     * against it, synthesised and scattered into a fresh numerator
     * (poisson.py:235-265) */
    const double *mdir = wiener ? prm->wi_mdir : prm->ht_mdir;
    const double *sdir = wiener ? prm->wi_sdir : prm->ht_sdir;
#ifdef _OPENMP
#pragma omp parallel
#endif
    {
      float *dv = (float *)malloc(sizeof(float) * (size_t)max_m * P);
      float *sv = (float *)malloc(sizeof(float) * (size_t)max_m * P);
      float *gfwd = (float *)malloc(sizeof(float) * (size_t)max_m * max_m);
      float *ginv = (float *)malloc(sizeof(float) * (size_t)max_m * max_m);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 8)
#endif
      for (int64_t r = 0; r < nrefs; ++r) {
        const int n = cnt[r];
        if (n == 0) continue;
        group_matrix(n, gfwd, ginv);
        int dims[ORACLE_MAX_NDIM + 1];
        const float *im[ORACLE_MAX_NDIM + 1];
        const double *dirs[ORACLE_MAX_NDIM + 1];
        dims[0] = n;
        im[0] = ginv;
        dirs[0] = mdir + n * (n - 1) / 2;
        int so = 0;
        for (int d = 0; d < ndim; ++d) {
          dims[d + 1] = block[d];
          im[d + 1] = g.sp_inv[d];
          dirs[d + 1] = sdir + so;
          so += block[d];
        }
        for (int i = 0; i < n; ++i) {
          const int64_t base = lin_to_vol_base(&g, sel[r * max_m + i]);
          for (int p = 0; p < P; ++p) {
            dv[i * P + p] = den[base + g.poff[p]];
            sv[i * P + p] = noisy[base + g.poff[p]];
          }
        }
        float *c = filt + off[r] * P;
        mass_correct_group(ndim, dims, im, dirs, wts + off[r], g.kaiser, dv, sv, c, 1e-10);
        oracle_apply_transform(c, ndim + 1, dims, im);
      }
      free(dv);
      free(sv);
      free(gfwd);
      free(ginv);
    }
    memset(num, 0, sizeof(float) * (size_t)nvox);
    float *den2 = (float *)calloc((size_t)nvox, sizeof(float));
    for (int64_t r = 0; r < nrefs; ++r)
      oracle_scatter_add(num, den2, ndim, g.shape, block, g.kaiser, filt + off[r] * P,
                         wts + off[r], sel + r * max_m, cnt[r]);
    free(den2);
    /* np.divide(num, den, where=den > 0) over a copy of the input
     * (bmndalgo.py:618-626, 1031-1038): no eps */
    for (int64_t v = 0; v < nvox; ++v) est[v] = den[v] > 0.0f ? num[v] / den[v] : fallback[v];
  } else {
    /* normalisation, bmndalgo.py:626-631 / 1039-1044 */
    float eps32 = (float)1e-10;
    for (int64_t v = 0; v < nvox; ++v)
      est[v] = den[v] > 0.0f ? num[v] / (den[v] + eps32) : fallback[v];
  }
  /* sigma map, bmndalgo.py:1048-1057: the global sigma where nothing landed */
  if (snum)
    for (int64_t v = 0; v < nvox; ++v) sigma_map[v] = den[v] > 0.0f ? snum[v] / den[v] : sigma_fill;
  free(snum);
  free(gsig);
  free(gsig1);

  free(num);
  free(den);
  free(filt);
  free(wts);
  free(off);
  free(sel);
  free(cnt);
  free(thrs);
  free(refs);
  geom_free(&g);
  return 0;
}

/* Poisson global sigma (sigmaestimation.py:48-93) of a variance volume whose
 * every value is positive (the variance floor is positive): the square root
 * of its float32 mean */
static int cmp_f32(const void *a, const void *b) {
  const float x = *(const float *)a, y = *(const float *)b;
  return (x > y) - (x < y);
}

/* estimate_poisson_global_sigma (sigmaestimation.py:48-93) on the variance
 * volume with eps = max(floor, 1e-10) (variance.py:581-593): the float32 mean
 * of the finite positive values -- a 10%-trimmed mean of the sorted values
 * when they are fewer than a tenth of the finite ones (and at least 3) --
 * then sqrt(max(mean, eps)).  The reference's IQR / MAD fallbacks run only
 * for a result <= 0, which eps > 0 rules out. */
double oracle_poisson_global_sigma(const float *var_vol, int64_t n, double floor_) {
  const double eps = floor_ > 1e-10 ? floor_ : 1e-10;
  int64_t nfin = 0, npos = 0;
  for (int64_t i = 0; i < n; ++i)
    if (isfinite(var_vol[i])) {
      ++nfin;
      if (var_vol[i] > 0.0f) ++npos;
    }
  double mean = 0.0;  /* no positive value: the mean of zeros */
  if (npos == n) {
    mean = (double)oracle_np_mean_f32(var_vol, n);
  } else if (npos > 0) {
    float *pos = (float *)malloc(sizeof(float) * (size_t)npos);
    int64_t j = 0;
    for (int64_t i = 0; i < n; ++i)
      if (isfinite(var_vol[i]) && var_vol[i] > 0.0f) pos[j++] = var_vol[i];
    const double frac = (double)npos / (double)(nfin > 1 ? nfin : 1);
    if (frac < 0.1 && npos >= 3) {
      qsort(pos, (size_t)npos, sizeof(float), cmp_f32);
      int64_t trim = (int64_t)(0.1 * (double)npos);
      if (trim < 1) trim = 1;
      mean = npos - 2 * trim > 0 ? (double)oracle_np_mean_f32(pos + trim, npos - 2 * trim)
                                 : (double)oracle_np_mean_f32(pos, npos);
    } else {
      mean = (double)oracle_np_mean_f32(pos, npos);
    }
    free(pos);
  }
  const double g = sqrt(eps > mean ? eps : mean);
  return g > 0.0 ? g : 0.0;
}

int oracle_bmnd_poisson_sigma(const float *vol, const oracle_params *prm, float *out,
                              float *pilot, int64_t *ht_sel, int32_t *ht_cnt, int64_t ht_cap,
                              int64_t *wi_sel, int32_t *wi_cnt, int64_t wi_cap,
                              float *ht_sigma_map, float *wi_sigma_map, double *global_sigma) {
  if (prm->ndim < 1 || prm->ndim > ORACLE_MAX_NDIM) return -1;
  int64_t nvox = 1;
  for (int d = 0; d < prm->ndim; ++d) nvox *= prm->shape[d];
#ifdef _OPENMP
  if (prm->nthreads > 0) omp_set_num_threads(prm->nthreads);
#endif
  int any = 0;
  for (int64_t i = 0; i < nvox; ++i)
    if (vol[i] != 0.0f) {
      any = 1;
      break;
    }
  if (!any) { /* all-zero fast path, bmndalgo.py:160-171 */
    memcpy(out, vol, sizeof(float) * (size_t)nvox);
    if (pilot) memcpy(pilot, vol, sizeof(float) * (size_t)nvox);
    if (ht_sigma_map) memset(ht_sigma_map, 0, sizeof(float) * (size_t)nvox);
    if (wi_sigma_map) memset(wi_sigma_map, 0, sizeof(float) * (size_t)nvox);
    if (global_sigma) *global_sigma = 0.0;
    return 0;
  }
  if (prm->gaussian) {
    /* the stationary model: no variance volume; both stages match with SSD,
     * stage 2 on the pilot (bmndalgo.py:203-228, 714-743) */
    float *est1 = (float *)malloc(sizeof(float) * (size_t)nvox);
    const float fill = (float)prm->global_sigma;
    if (global_sigma) *global_sigma = prm->global_sigma;
    run_stage(prm, 0, vol, NULL, vol, NULL, vol, est1, ht_sel, ht_cnt, ht_cap, ht_sigma_map,
              fill);
    if (pilot) memcpy(pilot, est1, sizeof(float) * (size_t)nvox);
    run_stage(prm, 1, est1, est1, vol, NULL, vol, out, wi_sel, wi_cnt, wi_cap, wi_sigma_map,
              fill);
    free(est1);
    return 0;
  }
  float *var1 = (float *)malloc(sizeof(float) * (size_t)nvox);
  float *est1 = (float *)malloc(sizeof(float) * (size_t)nvox);
  variance_volume(vol, nvox, prm->count_scale, prm->variance_floor, var1);
  const double gs = oracle_poisson_global_sigma(var1, nvox, prm->variance_floor);
  if (global_sigma) *global_sigma = gs;
  run_stage(prm, 0, vol, NULL, vol, var1, vol, est1, ht_sel, ht_cnt, ht_cap, ht_sigma_map,
            (float)gs);
  if (pilot) memcpy(pilot, est1, sizeof(float) * (size_t)nvox);
  /* Wiener variance model from the pilot, or the observation's model kept
   * (bmndalgo.py:635-646); matching on the pilot or the observation
   * (bmndalgo.py:714-720); the gains always use the pilot */
  if (!prm->wi_var_observation)
    variance_volume(est1, nvox, prm->count_scale, prm->variance_floor, var1);
  run_stage(prm, 1, prm->wi_match_observation ? vol : est1, est1, vol, var1, vol, out, wi_sel,
            wi_cnt, wi_cap, wi_sigma_map, (float)gs);
  free(var1);
  free(est1);
  return 0;
}

int oracle_bmnd_poisson(const float *vol, const oracle_params *prm, float *out,
                        float *pilot, int64_t *ht_sel, int32_t *ht_cnt, int64_t ht_cap,
                        int64_t *wi_sel, int32_t *wi_cnt, int64_t wi_cap) {
  return oracle_bmnd_poisson_sigma(vol, prm, out, pilot, ht_sel, ht_cnt, ht_cap, wi_sel, wi_cnt,
                                   wi_cap, NULL, NULL, NULL);
}
