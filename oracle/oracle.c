/*
 * oracle.c -- CPU restatement of bone-morphometry's exact squared EDT and
 * Fiji-style local-thickness path (Tb.Th / Tb.Sp).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity checker and the CPU
 * baseline.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it; the product library (bone_morphometry_amd/libbm.so) never
 * links, loads or calls it.
 *
 * Every function follows the reference step for step and cites the Rust it
 * restates (paths relative to the reference crate root).  The arithmetic is
 * kept identical where the result depends on it:
 *   - the EDT is the same f64 Felzenszwalb-Huttenlocher pass structure with
 *     INF = 1e30 (src/edt.rs:6,57-223);
 *   - the clean-up pass 2 sums in f32 in NEIGHBORS_26 order and divides by
 *     (float)count (src/local_thickness.rs:243-277,285-312);
 *   - nonzero_stats accumulates sequentially in f64 in linear order
 *     (src/metrics.rs:204-241).
 * Build with -ffp-contract=off so no product is fused into an FMA.
 *
 * Parallelism mirrors rayon's decomposition with OpenMP: EDT lines, the ridge
 * filter and the ridge propagation (atomic max) run in parallel; distance_index,
 * the clean-up and the statistics are serial, as in the reference.
 *
 * Pinning status of each function (how the tests check it):
 *   oracle_phase_inside      known answer volume.rs:123-136 + closed form
 *   oracle_edt_squared       known answers edt.rs:240-274, brute force,
 *                            scipy-derived fixtures, single-site closed form
 *   oracle_templates         known answer local_thickness.rs:395-401 and a
 *                            brute-force lattice maximum
 *   oracle_local_thickness   known answers local_thickness.rs:364-393, the
 *                            propagate-from-all-voxels invariant, clean-up
 *                            order independence (reverse traversal), bounds
 *   oracle_nonzero_stats     known answer metrics.rs:315-326 + closed forms
 *   oracle_quantile          known answer metrics.rs:304-313 + numpy linear
 *                            quantiles
 *   oracle_gaussian_kernel   known answer sdf_curvature.rs:783-789 (sums to
 *                            1), scipy's truncate-4 Gaussian weights
 *   oracle_gaussian_smooth_nearest
 *                            known answer sdf_curvature.rs:791-796 (constant
 *                            field kept), impulse response = w[kz](w[ky]w[kx])
 *                            exactly, scipy.ndimage.gaussian_filter with
 *                            mode "nearest" (edges), brute-force order check
 *   oracle_signed_distance_field_um
 *                            brute-force distances on small volumes (sign,
 *                            phase and scale), the EDT's error precedence
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORACLE_INF 1.0e30 /* src/edt.rs:6 */

static __thread char g_err[512];

const char *oracle_last_error(void) { return g_err; }

static int fail(const char *msg) {
  snprintf(g_err, sizeof g_err, "%s", msg);
  return 1;
}

int oracle_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void oracle_set_num_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

/* Phase::is_selected, src/edt.rs:17-24.  phase 0 = Foreground, 1 = Background. */
static inline int is_selected(int phase, uint8_t v) { return phase == 0 ? v != 0 : v == 0; }

/* BinaryVolume::new dimension rule, src/volume.rs:27-34. */
static int check_dims(size_t w, size_t h, size_t d) {
  if (w < 2 || h < 2 || d < 2) {
    snprintf(g_err, sizeof g_err,
             "all dimensions must be at least 2 for 3-D morphometry; got %zux%zux%zu", w, h, d);
    return 1;
  }
  return 0;
}

/* BinaryVolume::phase_inside, src/volume.rs:84-105 (shape check done by caller). */
void oracle_phase_inside(const uint8_t *bone, const uint8_t *roi, size_t n, int select_bone,
                         uint8_t *out) {
  for (size_t i = 0; i < n; ++i)
    out[i] = (uint8_t)(roi[i] != 0 && ((bone[i] != 0) == (select_bone != 0)));
}

/* intersection, src/edt.rs:219-223 */
static inline double intersection(const double *f, size_t p, size_t q, double qf) {
  double pf = (double)p;
  return ((f[q] + qf * qf) - (f[p] + pf * pf)) / (2.0 * (qf - pf));
}

/* transform_1d (Felzenszwalb-Huttenlocher), src/edt.rs:167-217 */
static void transform_1d(const double *f, double *d, size_t *v, double *z, size_t n) {
  size_t k = 0;
  v[0] = 0;
  z[0] = -INFINITY;
  z[1] = INFINITY;
  for (size_t q = 1; q < n; ++q) {
    double qf = (double)q;
    size_t p = v[k];
    double s = intersection(f, p, q, qf);
    while (s <= z[k]) {
      if (k == 0) break;
      k -= 1;
      p = v[k];
      s = intersection(f, p, q, qf);
    }
    if (k == 0 && s <= z[0]) {
      v[0] = q;
      z[0] = -INFINITY;
      z[1] = INFINITY;
    } else {
      k += 1;
      v[k] = q;
      z[k] = s;
      z[k + 1] = INFINITY;
    }
  }
  size_t env = 0;
  for (size_t q = 0; q < n; ++q) {
    double qf = (double)q;
    while (z[env + 1] < qf) env += 1;
    size_t p = v[env];
    double delta = qf - (double)p;
    d[q] = delta * delta + f[p];
  }
}

/* euclidean_distance_transform_squared, src/edt.rs:57-156 */
int oracle_edt_squared(const uint8_t *vol, size_t w, size_t h, size_t dd, int phase,
                       uint32_t *out) {
  if (check_dims(w, h, dd)) return 1;
  size_t n = w * h * dd, slice = w * h;
  int all = 1;
  for (size_t i = 0; i < n && all; ++i) all = is_selected(phase, vol[i]);
  if (all) { /* src/edt.rs:61-66 */
    snprintf(g_err, sizeof g_err,
             "distance transform is undefined for an all-%s volume unless an exterior "
             "convention is specified",
             phase == 0 ? "foreground" : "background");
    return 1;
  }
  double *data = (double *)malloc(n * sizeof(double));
  double *zl = (double *)malloc(n * sizeof(double));
  if (!data || !zl) { free(data); free(zl); return fail("oracle: out of host memory"); }

#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < n; ++i) data[i] = is_selected(phase, vol[i]) ? ORACLE_INF : 0.0;

  size_t mx = w > h ? w : h;
  if (dd > mx) mx = dd;
#pragma omp parallel
  {
    double *in = (double *)malloc(mx * sizeof(double));
    double *ob = (double *)malloc(mx * sizeof(double));
    size_t *v = (size_t *)malloc(mx * sizeof(size_t));
    double *z = (double *)malloc((mx + 1) * sizeof(double));
    /* X pass, src/edt.rs:79-90 */
#pragma omp for schedule(static)
    for (size_t r = 0; r < h * dd; ++r) {
      double *row = data + r * w;
      memcpy(in, row, w * sizeof(double));
      transform_1d(in, row, v, z, w);
    }
    /* Y pass, src/edt.rs:92-113 */
#pragma omp for schedule(static)
    for (size_t zi = 0; zi < dd; ++zi) {
      double *sl = data + zi * slice;
      for (size_t x = 0; x < w; ++x) {
        for (size_t y = 0; y < h; ++y) in[y] = sl[y * w + x];
        transform_1d(in, ob, v, z, h);
        for (size_t y = 0; y < h; ++y) sl[y * w + x] = ob[y];
      }
    }
    /* Z pass into z-line-major buffer, src/edt.rs:115-133 */
#pragma omp for schedule(static)
    for (size_t line = 0; line < slice; ++line) {
      for (size_t zi = 0; zi < dd; ++zi) in[zi] = data[zi * slice + line];
      transform_1d(in, zl + line * dd, v, z, dd);
    }
    free(in); free(ob); free(v); free(z);
  }
  /* back-transpose, round, phase mask, src/edt.rs:137-153 */
#pragma omp parallel for schedule(static)
  for (size_t zi = 0; zi < dd; ++zi) {
    for (size_t line = 0; line < slice; ++line) {
      size_t idx = zi * slice + line;
      out[idx] = is_selected(phase, vol[idx]) ? (uint32_t)round(zl[line * dd + zi]) : 0u;
    }
  }
  free(data);
  free(zl);
  return 0;
}

/* integer_sqrt, src/local_thickness.rs:338-348 */
static inline uint32_t integer_sqrt(uint32_t value) {
  uint32_t root = (uint32_t)sqrt((double)value);
  while ((uint64_t)(root + 1) * (uint64_t)(root + 1) <= (uint64_t)value) root += 1;
  while ((uint64_t)root * (uint64_t)root > (uint64_t)value) root -= 1;
  return root;
}

/* ceiling_sqrt, src/local_thickness.rs:350-358 */
static inline uint32_t ceiling_sqrt(uint32_t value) {
  uint32_t fl = integer_sqrt(value);
  return fl * fl == value ? fl : fl + 1;
}

/* scan_cube, src/local_thickness.rs:75-102 (one radius) */
static uint32_t scan_cube_one(uint32_t dx, uint32_t dy, uint32_t dz, uint32_t r2) {
  uint32_t radius = integer_sqrt(r2) + 1, maximum = 0;
  for (uint32_t k = 0; k <= radius; ++k) {
    uint32_t sk = k * k, dk = (k + dz) * (k + dz);
    for (uint32_t j = 0; j <= radius; ++j) {
      uint32_t skj = sk + j * j;
      if (skj > r2) continue;
      uint32_t ip = integer_sqrt(r2 - skj) + dx;
      uint32_t cand = dk + (j + dy) * (j + dy) + ip * ip;
      if (cand > maximum) maximum = cand;
    }
  }
  return maximum;
}

/* create_templates, src/local_thickness.rs:67-73: out[c*count + i] for c = 0..2 */
void oracle_templates(const uint32_t *values, size_t count, uint32_t *out) {
#pragma omp parallel for schedule(dynamic, 4)
  for (size_t i = 0; i < count; ++i) {
    out[i] = scan_cube_one(1, 0, 0, values[i]);
    out[count + i] = scan_cube_one(1, 1, 0, values[i]);
    out[2 * count + i] = scan_cube_one(1, 1, 1, values[i]);
  }
}

/* NEIGHBORS_26 in Fiji's visiting order, src/local_thickness.rs:283-312 */
static const int NB26[26][3] = {
    {0, 0, -1},  {0, 0, 1},   {0, -1, 0},  {0, 1, 0},  {-1, 0, 0},  {1, 0, 0},  {0, 1, -1},
    {0, 1, 1},   {1, -1, 0},  {1, 1, 0},   {-1, 0, 1}, {1, 0, 1},   {0, -1, -1}, {0, -1, 1},
    {-1, -1, 0}, {-1, 1, 0},  {-1, 0, -1}, {1, 0, -1}, {1, 1, 1},   {1, -1, 1},  {-1, 1, 1},
    {-1, -1, 1}, {1, 1, -1},  {1, -1, -1}, {-1, 1, -1}, {-1, -1, -1}};

/* look, src/local_thickness.rs:314-331 */
static inline float look(const float *v, long x, long y, long z, long w, long h, long d) {
  if (x < 0 || y < 0 || z < 0 || x >= w || y >= h || z >= d) return -1.0f;
  return v[((size_t)z * (size_t)h + (size_t)y) * (size_t)w + (size_t)x];
}

/* clean_up_surface_jaggies, src/local_thickness.rs:210-281.  reverse != 0 runs
 * pass 2 in reverse z,y,x order (used only to test order independence). */
static void clean_up(const float *input, float *output, size_t w, size_t h, size_t d,
                     int reverse) {
  size_t n = w * h * d;
  memset(output, 0, n * sizeof(float));
  for (size_t z = 0; z < d; ++z)
    for (size_t y = 0; y < h; ++y)
      for (size_t x = 0; x < w; ++x) {
        size_t idx = (z * h + y) * w + x;
        float value = input[idx];
        if (value == 0.0f) continue;
        int borders = 0;
        for (int k = 0; k < 26 && !borders; ++k)
          borders = look(input, (long)x + NB26[k][0], (long)y + NB26[k][1],
                         (long)z + NB26[k][2], (long)w, (long)h, (long)d) == 0.0f;
        output[idx] = borders ? -1.0f : value;
      }
  for (size_t t = 0; t < n; ++t) {
    size_t idx = reverse ? n - 1 - t : t;
    if (output[idx] != -1.0f) continue;
    size_t z = idx / (w * h), rem = idx - z * w * h, y = rem / w, x = rem - y * w;
    size_t count = 0;
    float sum = 0.0f;
    for (int k = 0; k < 26; ++k) {
      float v = look(output, (long)x + NB26[k][0], (long)y + NB26[k][1], (long)z + NB26[k][2],
                     (long)w, (long)h, (long)d);
      if (v > 0.0f) {
        count += 1;
        sum += v;
      }
    }
    float corrected = count > 0 ? sum / (float)count : input[idx];
    output[idx] = -corrected;
  }
  for (size_t i = 0; i < n; ++i) output[i] = fabsf(output[i]);
}

static inline void atomic_max_u32(uint32_t *p, uint32_t v) {
  uint32_t cur = __atomic_load_n(p, __ATOMIC_RELAXED);
  while (cur < v &&
         !__atomic_compare_exchange_n(p, &cur, v, 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
  }
}

/* Optional intermediate outputs of oracle_local_thickness_ex. */
typedef struct {
  uint8_t *ridge_flags;     /* N bytes: 1 where distance_ridges selected the voxel */
  uint32_t *propagated;     /* N: max r^2 over covering ridge balls (before sqrt) */
  uint32_t *squared_edt;    /* N: the squared EDT */
  uint64_t ridge_count;     /* out */
  uint64_t fill_visits;     /* out: sum over ridges of |ball(r^2) within the array| */
  int propagate_from_all;   /* in: propagate from every r^2>0 voxel, not only ridges */
  int reverse_cleanup;      /* in: pass 2 traversal reversed */
} oracle_lt_debug;

/* local_thickness, src/local_thickness.rs:16-44, with the helpers it calls. */
int oracle_local_thickness_ex(const uint8_t *vol, size_t w, size_t h, size_t d, int phase,
                              float *out, oracle_lt_debug *dbg) {
  if (check_dims(w, h, d)) return 1;
  size_t n = w * h * d, slice = w * h;
  uint32_t *sq = (dbg && dbg->squared_edt) ? dbg->squared_edt
                                           : (uint32_t *)malloc(n * sizeof(uint32_t));
  if (!sq) return fail("oracle: out of host memory");
  if (oracle_edt_squared(vol, w, h, d, phase, sq)) {
    if (!(dbg && dbg->squared_edt)) free(sq);
    return 1;
  }
  /* distance_index, src/local_thickness.rs:46-65 (serial) */
  uint32_t maximum = 0;
  for (size_t i = 0; i < n; ++i)
    if (sq[i] > maximum) maximum = sq[i];
  uint8_t *occurs = (uint8_t *)calloc((size_t)maximum + 1, 1);
  for (size_t i = 0; i < n; ++i) occurs[sq[i]] = 1;
  size_t nvals = 0;
  for (size_t v = 0; v <= maximum; ++v) nvals += occurs[v];
  uint32_t *values = (uint32_t *)malloc(nvals * sizeof(uint32_t));
  size_t *rank = (size_t *)malloc(((size_t)maximum + 1) * sizeof(size_t));
  for (size_t v = 0, k = 0; v <= maximum; ++v) {
    rank[v] = (size_t)-1;
    if (occurs[v]) {
      values[k] = (uint32_t)v;
      rank[v] = k++;
    }
  }
  /* create_templates, src/local_thickness.rs:67-73 */
  uint32_t *tmpl = (uint32_t *)malloc(3 * nvals * sizeof(uint32_t));
  oracle_templates(values, nvals, tmpl);

  /* distance_ridges, src/local_thickness.rs:104-167.  The ridge list is kept in
   * ascending index order (rayon's indexed collect) via a flag array. */
  uint8_t *ridge = (uint8_t *)calloc(n, 1);
#pragma omp parallel for schedule(static)
  for (size_t idx = 0; idx < n; ++idx) {
    if (!is_selected(phase, vol[idx])) continue;
    uint32_t r2 = sq[idx];
    if (r2 == 0) continue;
    if (dbg && dbg->propagate_from_all) {
      ridge[idx] = 1;
      continue;
    }
    size_t z = idx / slice, rem = idx - z * slice, y = rem / w, x = rem - y * w;
    size_t ri = rank[r2];
    int is_ridge = 1;
    for (long dz = -1; dz <= 1 && is_ridge; ++dz) {
      long nz = (long)z + dz;
      if (nz < 0 || nz >= (long)d) continue;
      for (long dy = -1; dy <= 1 && is_ridge; ++dy) {
        long ny = (long)y + dy;
        if (ny < 0 || ny >= (long)h) continue;
        for (long dx = -1; dx <= 1; ++dx) {
          if (dx == 0 && dy == 0 && dz == 0) continue;
          long nx = (long)x + dx;
          if (nx < 0 || nx >= (long)w) continue;
          int comps = (dx != 0) + (dy != 0) + (dz != 0);
          size_t nb = ((size_t)nz * h + (size_t)ny) * w + (size_t)nx;
          if (sq[nb] >= tmpl[(size_t)(comps - 1) * nvals + ri]) {
            is_ridge = 0;
            break;
          }
        }
      }
    }
    ridge[idx] = (uint8_t)is_ridge;
  }

  /* propagate_ridges, src/local_thickness.rs:169-208 */
  uint32_t *prop = (dbg && dbg->propagated) ? dbg->propagated
                                            : (uint32_t *)malloc(n * sizeof(uint32_t));
  memset(prop, 0, n * sizeof(uint32_t));
  uint64_t ridge_count = 0, visits = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : ridge_count, visits)
  for (size_t idx = 0; idx < n; ++idx) {
    if (!ridge[idx]) continue;
    ridge_count += 1;
    uint32_t r2 = sq[idx];
    size_t z = idx / slice, rem = idx - z * slice, y = rem / w, x = rem - y * w;
    size_t radius = ceiling_sqrt(r2);
    size_t x0 = x > radius ? x - radius : 0, y0 = y > radius ? y - radius : 0,
           z0 = z > radius ? z - radius : 0;
    size_t x1 = x + radius < w - 1 ? x + radius : w - 1;
    size_t y1 = y + radius < h - 1 ? y + radius : h - 1;
    size_t z1 = z + radius < d - 1 ? z + radius : d - 1;
    uint64_t r2u = r2;
    for (size_t nz = z0; nz <= z1; ++nz) {
      uint64_t dzz = nz > z ? nz - z : z - nz;
      uint64_t dz2 = dzz * dzz;
      for (size_t ny = y0; ny <= y1; ++ny) {
        uint64_t dyy = ny > y ? ny - y : y - ny;
        uint64_t dyz = dz2 + dyy * dyy;
        if (dyz > r2u) continue;
        for (size_t nx = x0; nx <= x1; ++nx) {
          uint64_t dxx = nx > x ? nx - x : x - nx;
          if (dyz + dxx * dxx <= r2u) {
            atomic_max_u32(&prop[(nz * h + ny) * w + nx], r2);
            visits += 1;
          }
        }
      }
    }
  }
  float *diam = (float *)malloc(n * sizeof(float));
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < n; ++i) diam[i] = (float)(2.0 * sqrt((double)prop[i]));

  float *cleaned = (float *)malloc(n * sizeof(float));
  clean_up(diam, cleaned, w, h, d, dbg ? dbg->reverse_cleanup : 0);

  /* phase mask, src/local_thickness.rs:31-43 */
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < n; ++i) out[i] = is_selected(phase, vol[i]) ? cleaned[i] : 0.0f;

  if (dbg) {
    dbg->ridge_count = ridge_count;
    dbg->fill_visits = visits;
    if (dbg->ridge_flags) memcpy(dbg->ridge_flags, ridge, n);
  }
  free(cleaned);
  free(diam);
  if (!(dbg && dbg->propagated)) free(prop);
  free(ridge);
  free(tmpl);
  free(rank);
  free(values);
  free(occurs);
  if (!(dbg && dbg->squared_edt)) free(sq);
  return 0;
}

int oracle_local_thickness(const uint8_t *vol, size_t w, size_t h, size_t d, int phase,
                           float *out) {
  return oracle_local_thickness_ex(vol, w, h, d, phase, out, NULL);
}

/* DistributionStats, src/metrics.rs:26-38 */
typedef struct {
  double mean, standard_deviation, maximum, median, percentile_10, percentile_25,
      percentile_75, percentile_90, interquartile_range, coefficient_of_variation;
  uint64_t count;
} oracle_stats;

/* select_value, src/metrics.rs:295-299: the index-th smallest value (the
 * values here are finite and positive, so total_cmp is the plain order).  Like
 * the reference's select_nth_unstable_by it leaves the slice partitioned about
 * that position; the linear-time selection (Hoare partition about a median of
 * three, recursing into the side that holds the index) returns the same order
 * statistic whatever the partition order. */
static float select_value(float *v, size_t len, size_t index) {
  size_t lo = 0, hi = len - 1;
  while (hi > lo) {
    size_t mid = lo + (hi - lo) / 2;
    float a = v[lo], b = v[mid], c = v[hi];
    float pivot = a < b ? (b < c ? b : (a < c ? c : a)) : (a < c ? a : (b < c ? c : b));
    size_t i = lo, j = hi;
    for (;;) {
      while (v[i] < pivot) ++i;
      while (v[j] > pivot) --j;
      if (i >= j) break;
      float t = v[i];
      v[i] = v[j];
      v[j] = t;
      ++i;
      --j;
    }
    /* v[lo..j] <= pivot <= v[j+1..hi] */
    if (index <= j) hi = j;
    else lo = j + 1;
  }
  return v[index];
}

/* linear_quantile_unstable, src/metrics.rs:271-293. */
static double linear_quantile(float *values, size_t len, double q, float scale) {
  double position = q * (double)(len - 1);
  size_t lower = (size_t)floor(position), upper = (size_t)ceil(position);
  double fraction = position - (double)lower;
  double lo = (double)select_value(values, len, lower);
  double hi = upper == lower ? lo : (double)select_value(values, len, upper);
  double qv = lo + fraction * (hi - lo);
  return (double)((float)qv * scale);
}

double oracle_quantile(const float *values, size_t len, double q, float scale) {
  float *s = (float *)malloc(len * sizeof(float));
  memcpy(s, values, len * sizeof(float));
  double r = linear_quantile(s, len, q, scale);
  free(s);
  return r;
}

/* nonzero_stats, src/metrics.rs:204-269 */
int oracle_nonzero_stats(const float *values, size_t n, double scale, int include_quantiles,
                         oracle_stats *st) {
  uint64_t count = 0;
  double sum = 0.0, sumsq = 0.0, maximum = 0.0;
  float scale_f32 = (float)scale;
  for (size_t i = 0; i < n; ++i) {
    float raw = values[i];
    if (!isfinite(raw) || raw <= 0.0f) continue;
    double value = (double)(raw * scale_f32);
    count += 1;
    sum += value;
    sumsq += value * value;
    maximum = maximum > value ? maximum : value;
  }
  if (count == 0) return fail("cannot compute statistics because the selected phase is empty");
  double cf = (double)count;
  double mean = sum / cf;
  double centered = (cf * sumsq - sum * sum) / cf;
  double sd = (count > 1 && centered > 0.0) ? sqrt(centered / (cf - 1.0)) : 0.0;
  double med = NAN, p10 = NAN, p25 = NAN, p75 = NAN, p90 = NAN;
  if (include_quantiles) {
    float *kept = (float *)malloc(count * sizeof(float));
    size_t k = 0;
    for (size_t i = 0; i < n; ++i)
      if (isfinite(values[i]) && values[i] > 0.0f) kept[k++] = values[i];
    /* the reference's call order: median, P10, P25, P75, P90 (:245-251) */
    med = linear_quantile(kept, k, 0.5, scale_f32);
    p10 = linear_quantile(kept, k, 0.1, scale_f32);
    p25 = linear_quantile(kept, k, 0.25, scale_f32);
    p75 = linear_quantile(kept, k, 0.75, scale_f32);
    p90 = linear_quantile(kept, k, 0.9, scale_f32);
    free(kept);
  }
  st->mean = mean;
  st->standard_deviation = sd;
  st->maximum = maximum;
  st->median = med;
  st->percentile_10 = p10;
  st->percentile_25 = p25;
  st->percentile_75 = p75;
  st->percentile_90 = p90;
  st->interquartile_range = p75 - p25;
  st->coefficient_of_variation = sd / mean;
  st->count = count;
  return 0;
}

/* Thickness/spacing part of analyze, src/metrics.rs:64-132. */
typedef struct {
  oracle_stats thickness, spacing;
  double bone_volume, total_volume, bone_volume_fraction;
} oracle_metrics;

int oracle_analyze_thickness_spacing(const uint8_t *bone, const uint8_t *roi, size_t w, size_t h,
                                     size_t d, const double spacing[3], int bonej_mode,
                                     int include_quantiles, oracle_metrics *m) {
  if (check_dims(w, h, d)) return 1;
  for (int i = 0; i < 3; ++i)
    if (!isfinite(spacing[i]) || spacing[i] <= 0.0) {
      snprintf(g_err, sizeof g_err, "voxel spacings must be finite and positive; got [%g, %g, %g]",
               spacing[0], spacing[1], spacing[2]);
      return 1;
    }
  size_t n = w * h * d;
  size_t roi_voxels = 0;
  for (size_t i = 0; i < n; ++i) roi_voxels += roi[i] != 0;
  if (roi_voxels == 0) return fail("ROI mask contains no included voxels");
  if (bonej_mode && roi_voxels == n)
    return fail(
        "BoneJ compatibility mode requires an ROI mask with both inside and outside voxels");
  uint8_t *ab = (uint8_t *)malloc(n), *am = (uint8_t *)malloc(n);
  oracle_phase_inside(bone, roi, n, 1, ab);
  oracle_phase_inside(bone, roi, n, 0, am);
  size_t bone_voxels = 0, marrow_voxels = 0;
  for (size_t i = 0; i < n; ++i) {
    bone_voxels += ab[i] != 0;
    marrow_voxels += am[i] != 0;
  }
  int rc = 0;
  float *map = NULL;
  if (bone_voxels == 0) { rc = fail("ROI contains no bone voxels"); goto done; }
  if (marrow_voxels == 0) { rc = fail("ROI contains no marrow voxels"); goto done; }
  map = (float *)malloc(n * sizeof(float));
  if ((rc = oracle_local_thickness(ab, w, h, d, 0, map))) goto done;
  if ((rc = oracle_nonzero_stats(map, n, spacing[0], include_quantiles, &m->thickness))) goto done;
  if ((rc = oracle_local_thickness(am, w, h, d, 0, map))) goto done;
  if ((rc = oracle_nonzero_stats(map, n, spacing[0], include_quantiles, &m->spacing))) goto done;
  {
    double vv = spacing[0] * spacing[1] * spacing[2];
    m->bone_volume = (double)bone_voxels * vv;
    m->total_volume = (double)roi_voxels * vv;
    m->bone_volume_fraction = m->bone_volume / m->total_volume;
  }
done:
  free(map);
  free(ab);
  free(am);
  return rc;
}

/* ------------------------------------------------------------------------
 * The signed distance field and its Gaussian smoothing (the EDT consumers of
 * the curvature and Minkowski measures).
 */

/* gaussian_kernel, src/sdf_curvature.rs:182-198.  Writes min(len, cap)
 * weights; returns len = 2r+1 (0 with the error set when r exceeds 2^20).
 * `(4.0 * sigma + 0.5).floor() as isize` saturates in Rust and maps NaN to 0;
 * exp is the C library's, the one Rust's f64::exp calls on this platform. */
size_t oracle_gaussian_kernel(double sigma, double *out, size_t cap) {
  if (sigma <= 0.0) {
    if (cap) out[0] = 1.0;
    return 1;
  }
  double rf = floor(4.0 * sigma + 0.5);
  long radius = 0;
  if (!isnan(rf)) {
    if (rf > (double)(1 << 20)) {
      fail("oracle: gaussian radius too large");
      return 0;
    }
    radius = (long)rf;
  }
  size_t len = (size_t)(2 * radius + 1);
  double *k = (double *)malloc(len * sizeof(double));
  double sum = 0.0;
  for (long i = -radius; i <= radius; ++i) {
    double x = (double)i;
    double w = exp(-0.5 * x * x / (sigma * sigma));
    k[i + radius] = w;
    sum += w;
  }
  for (size_t i = 0; i < len; ++i) k[i] /= sum;
  if (cap) memcpy(out, k, (len < cap ? len : cap) * sizeof(double));
  free(k);
  return len;
}

static inline size_t clamp_index(long v, size_t n) {
  return v < 0 ? 0 : (v >= (long)n ? n - 1 : (size_t)v);
}

/* gaussian_smooth_nearest_in_place, src/sdf_curvature.rs:200-269: x, y, z
 * passes, each output acc = 0; acc += w[k] * data[clamp(p + k - r)] in tap
 * order.  The reference runs serially; the outputs are independent, so the
 * z-slices are spread over threads without changing any value. */
int oracle_gaussian_smooth_nearest(double *data, size_t width, size_t height, size_t depth,
                                   double sigma_vox) {
  size_t len = oracle_gaussian_kernel(sigma_vox, NULL, 0);
  if (len == 0) return 1;
  if (len == 1) return 0;
  double *kernel = (double *)malloc(len * sizeof(double));
  oracle_gaussian_kernel(sigma_vox, kernel, len);
  long radius = (long)(len / 2);
  size_t n = width * height * depth;
  double *tmp = (double *)malloc((n ? n : 1) * sizeof(double));
  if (!tmp || !kernel) {
    free(tmp);
    free(kernel);
    return fail("oracle: out of host memory");
  }
#define IDX(x, y, z) (((z) * height + (y)) * width + (x))
  for (int pass = 0; pass < 3; ++pass) {
    const double *src = pass == 1 ? tmp : data;
    double *dst = pass == 1 ? data : tmp;
#pragma omp parallel for schedule(static)
    for (long zl = 0; zl < (long)depth; ++zl) {
      size_t z = (size_t)zl;
      for (size_t y = 0; y < height; ++y)
        for (size_t x = 0; x < width; ++x) {
          double acc = 0.0;
          for (size_t ki = 0; ki < len; ++ki) {
            long dk = (long)ki - radius;
            size_t i;
            if (pass == 0) i = IDX(clamp_index((long)x + dk, width), y, z);
            else if (pass == 1) i = IDX(x, clamp_index((long)y + dk, height), z);
            else i = IDX(x, y, clamp_index((long)z + dk, depth));
            acc += kernel[ki] * src[i];
          }
          dst[IDX(x, y, z)] = acc;
        }
    }
  }
#undef IDX
  /* three swaps: the z pass's output (tmp) is the result */
  memcpy(data, tmp, n * sizeof(double));
  free(tmp);
  free(kernel);
  return 0;
}

/* signed_distance_field_um, src/sdf_curvature.rs:162-180: the foreground EDT
 * (its error first), +sqrt(d2) h on bone voxels; then the background EDT,
 * -sqrt(d2) h on the others. */
int oracle_signed_distance_field_um(const uint8_t *bone, size_t w, size_t h, size_t d,
                                    double hs, double *phi) {
  size_t n = w * h * d;
  uint32_t *sq = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
  if (!sq) return fail("oracle: out of host memory");
  for (size_t i = 0; i < n; ++i) phi[i] = 0.0;
  if (oracle_edt_squared(bone, w, h, d, 0, sq)) {
    free(sq);
    return 1;
  }
  for (size_t i = 0; i < n; ++i)
    if (bone[i] != 0) phi[i] = sqrt((double)sq[i]) * hs;
  if (oracle_edt_squared(bone, w, h, d, 1, sq)) {
    free(sq);
    return 1;
  }
  for (size_t i = 0; i < n; ++i)
    if (bone[i] == 0) phi[i] = -sqrt((double)sq[i]) * hs;
  free(sq);
  return 0;
}

/* ------------------------------------------------------------------------
 * The rest of analyze (src/metrics.rs:134-164): cubical measures, topology,
 * BoneJ connectivity, bone-marrow interface faces and the ROI-boundary bone
 * fraction, all on the analysis bone (bone inside the ROI).
 */
static inline int fg_at(const uint8_t *v, long x, long y, long z, long w, long h, long d) {
  if (x < 0 || y < 0 || z < 0 || x >= w || y >= h || z >= d) return 0;
  return v[((size_t)z * (size_t)h + (size_t)y) * (size_t)w + (size_t)x] != 0;
}

typedef struct {
  int64_t raw_euler;
  double mean_breadth_lattice_units;
  int64_t vertices, edges, faces, cubes; /* the cell counts behind them */
} oracle_cubical;

/* cubical_measures, src/topology.rs:53-171 (integer counts; the loops are
 * spread over threads, which changes no count). */
void oracle_cubical_measures(const uint8_t *v, size_t w_, size_t h_, size_t d_, oracle_cubical *out) {
  long w = (long)w_, h = (long)h_, d = (long)d_;
  int64_t cubes = 0, shared = 0, edges = 0, vertices = 0;
#pragma omp parallel for reduction(+ : cubes, shared) schedule(static)
  for (long z = 0; z < d; ++z)
    for (long y = 0; y < h; ++y)
      for (long x = 0; x < w; ++x) {
        if (!fg_at(v, x, y, z, w, h, d)) continue;
        cubes += 1;
        shared += fg_at(v, x + 1, y, z, w, h, d) + fg_at(v, x, y + 1, z, w, h, d) +
                  fg_at(v, x, y, z + 1, w, h, d);
      }
  int64_t faces = 6 * cubes - shared;
#pragma omp parallel for reduction(+ : edges) schedule(static)
  for (long zv = 0; zv <= d; ++zv) {
    for (long yv = 0; yv <= h; ++yv) /* x-directed */
      for (long x = 0; x < w; ++x)
        edges += fg_at(v, x, yv - 1, zv - 1, w, h, d) || fg_at(v, x, yv, zv - 1, w, h, d) ||
                 fg_at(v, x, yv - 1, zv, w, h, d) || fg_at(v, x, yv, zv, w, h, d);
    for (long y = 0; y < h; ++y) /* y-directed */
      for (long xv = 0; xv <= w; ++xv)
        edges += fg_at(v, xv - 1, y, zv - 1, w, h, d) || fg_at(v, xv, y, zv - 1, w, h, d) ||
                 fg_at(v, xv - 1, y, zv, w, h, d) || fg_at(v, xv, y, zv, w, h, d);
  }
#pragma omp parallel for reduction(+ : edges) schedule(static)
  for (long z = 0; z < d; ++z) /* z-directed */
    for (long yv = 0; yv <= h; ++yv)
      for (long xv = 0; xv <= w; ++xv)
        edges += fg_at(v, xv - 1, yv - 1, z, w, h, d) || fg_at(v, xv, yv - 1, z, w, h, d) ||
                 fg_at(v, xv - 1, yv, z, w, h, d) || fg_at(v, xv, yv, z, w, h, d);
#pragma omp parallel for reduction(+ : vertices) schedule(static)
  for (long zv = 0; zv <= d; ++zv)
    for (long yv = 0; yv <= h; ++yv)
      for (long xv = 0; xv <= w; ++xv) {
        int occ = 0;
        for (long dz = -1; dz <= 0 && !occ; ++dz)
          for (long dy = -1; dy <= 0 && !occ; ++dy)
            for (long dx = -1; dx <= 0 && !occ; ++dx) occ = fg_at(v, xv + dx, yv + dy, zv + dz, w, h, d);
        vertices += occ;
      }
  out->vertices = vertices;
  out->edges = edges;
  out->faces = faces;
  out->cubes = cubes;
  out->raw_euler = vertices - edges + faces - cubes;
  out->mean_breadth_lattice_units = 0.5 * (double)(edges - 2 * faces + 3 * cubes);
}

/* bonej_connectivity and its stack counts, src/topology.rs:175-427. */
typedef struct {
  double edge_correction, delta_chi, connectivity;
  int64_t f, e, c, d, a, b; /* stack_vertices, ... as combined at src/topology.rs:182-187 */
} oracle_bonej;

void oracle_bonej_connectivity(const uint8_t *v, size_t w_, size_t h_, size_t d_, int64_t raw_euler,
                               oracle_bonej *out) {
  long w = (long)w_, h = (long)h_, d = (long)d_;
#define FG(x, y, z) fg_at(v, (x), (y), (z), w, h, d)
  long zs[2] = {0, d - 1}, ys[2] = {0, h - 1}, xs[2] = {0, w - 1};
  int64_t sv = 0, se = 0, sf = 0, fv = 0, fe = 0, ev = 0;
  for (int i = 0; i < 2; ++i) /* stack_vertices */
    for (int j = 0; j < 2; ++j)
      for (int k = 0; k < 2; ++k) sv += FG(xs[k], ys[j], zs[i]);
  for (int i = 0; i < 2; ++i) /* stack_edges */
    for (int j = 0; j < 2; ++j)
      for (long x = 1; x < w - 1; ++x) se += FG(x, ys[j], zs[i]);
  for (int i = 0; i < 2; ++i)
    for (int k = 0; k < 2; ++k)
      for (long y = 1; y < h - 1; ++y) se += FG(xs[k], y, zs[i]);
  for (int j = 0; j < 2; ++j)
    for (int k = 0; k < 2; ++k)
      for (long z = 1; z < d - 1; ++z) se += FG(xs[k], ys[j], z);
  for (int i = 0; i < 2; ++i) /* stack_faces */
    for (long y = 1; y < h - 1; ++y)
      for (long x = 1; x < w - 1; ++x) sf += FG(x, y, zs[i]);
  for (int j = 0; j < 2; ++j)
    for (long z = 1; z < d - 1; ++z)
      for (long x = 1; x < w - 1; ++x) sf += FG(x, ys[j], z);
  for (int k = 0; k < 2; ++k)
    for (long y = 1; y < h - 1; ++y)
      for (long z = 1; z < d - 1; ++z) sf += FG(xs[k], y, z);
  for (int i = 0; i < 2; ++i) /* face_vertices */
    for (long y = 0; y <= h; ++y)
      for (long x = 0; x <= w; ++x) {
        long z = zs[i];
        fv += FG(x, y, z) || FG(x, y - 1, z) || FG(x - 1, y - 1, z) || FG(x - 1, y, z);
      }
  for (int k = 0; k < 2; ++k)
    for (long y = 0; y <= h; ++y)
      for (long z = 1; z < d; ++z) {
        long x = xs[k];
        fv += FG(x, y, z) || FG(x, y - 1, z) || FG(x, y - 1, z - 1) || FG(x, y, z - 1);
      }
  for (int j = 0; j < 2; ++j)
    for (long x = 1; x < w; ++x)
      for (long z = 1; z < d; ++z) {
        long y = ys[j];
        fv += FG(x, y, z) || FG(x, y, z - 1) || FG(x - 1, y, z - 1) || FG(x - 1, y, z);
      }
  for (int i = 0; i < 2; ++i) /* face_edges */
    for (long y = 0; y <= h; ++y)
      for (long x = 0; x <= w; ++x) {
        long z = zs[i];
        if (FG(x, y, z)) fe += 2;
        else fe += FG(x, y - 1, z) + FG(x - 1, y, z);
      }
  for (int j = 0; j < 2; ++j) {
    long y = ys[j];
    for (long z = 1; z < d; ++z)
      for (long x = 0; x < w; ++x) fe += FG(x, y, z) || FG(x, y, z - 1);
    for (long z = 0; z < d; ++z)
      for (long x = 0; x <= w; ++x) fe += FG(x, y, z) || FG(x - 1, y, z);
  }
  for (int k = 0; k < 2; ++k) {
    long x = xs[k];
    for (long z = 1; z < d; ++z)
      for (long y = 0; y < h; ++y) fe += FG(x, y, z) || FG(x, y, z - 1);
    for (long z = 0; z < d; ++z)
      for (long y = 1; y < h; ++y) fe += FG(x, y, z) || FG(x, y - 1, z);
  }
  for (int i = 0; i < 2; ++i) /* edge_vertices */
    for (int j = 0; j < 2; ++j)
      for (long x = 1; x < w; ++x) ev += FG(x, ys[j], zs[i]) || FG(x - 1, ys[j], zs[i]);
  for (int i = 0; i < 2; ++i)
    for (int k = 0; k < 2; ++k)
      for (long y = 1; y < h; ++y) ev += FG(xs[k], y, zs[i]) || FG(xs[k], y - 1, zs[i]);
  for (int k = 0; k < 2; ++k)
    for (int j = 0; j < 2; ++j)
      for (long z = 1; z < d; ++z) ev += FG(xs[k], ys[j], z) || FG(xs[k], ys[j], z - 1);
#undef FG
  int64_t f = sv, e = se + 3 * f, c = sf + 2 * e - 3 * f, dd = ev + f, a = fv, b = fe;
  out->f = f;
  out->e = e;
  out->c = c;
  out->d = dd;
  out->a = a;
  out->b = b;
  double chi_zero = (double)f, chi_one = (double)(dd - e), chi_two = (double)(a - b + c);
  out->edge_correction = chi_two / 2.0 + chi_one / 4.0 + chi_zero / 8.0;
  out->delta_chi = (double)raw_euler - out->edge_correction;
  out->connectivity = 1.0 - out->delta_chi;
}

/* flood, src/topology.rs:614-701: depth-first with an explicit stack */
static int flood_fill(const uint8_t *v, size_t w, size_t h, size_t d, size_t seed, int foreground,
                      int conn26, uint8_t *visited, uint32_t **stack, size_t *cap) {
  size_t top = 0, slice = w * h;
  (*stack)[top++] = (uint32_t)seed;
  visited[seed] = 1;
  while (top) {
    size_t index = (*stack)[--top];
    long z = (long)(index / slice), rem = (long)(index % slice);
    long y = rem / (long)w, x = rem % (long)w;
    for (long dz = -1; dz <= 1; ++dz)
      for (long dy = -1; dy <= 1; ++dy)
        for (long dx = -1; dx <= 1; ++dx) {
          int n6 = (dx != 0) + (dy != 0) + (dz != 0);
          if (n6 == 0 || (!conn26 && n6 != 1)) continue;
          long nx = x + dx, ny = y + dy, nz = z + dz;
          if (nx < 0 || ny < 0 || nz < 0 || nx >= (long)w || ny >= (long)h || nz >= (long)d) continue;
          size_t nb = ((size_t)nz * h + (size_t)ny) * w + (size_t)nx;
          if (!visited[nb] && ((v[nb] != 0) == foreground)) {
            visited[nb] = 1;
            if (top == *cap) {
              *cap *= 2;
              uint32_t *s = (uint32_t *)realloc(*stack, *cap * sizeof(uint32_t));
              if (!s) return 1;
              *stack = s;
            }
            (*stack)[top++] = (uint32_t)nb;
          }
        }
  }
  return 0;
}

typedef struct {
  int64_t raw_euler;
  uint64_t beta0, beta1, beta2;
} oracle_topology;

/* analyze_topology, src/topology.rs:29-51 with count_components (26,
 * foreground) and count_enclosed_background_components_6, :504-612. */
int oracle_analyze_topology(const uint8_t *v, size_t w, size_t h, size_t d, int64_t raw_euler,
                            oracle_topology *out) {
  size_t n = w * h * d;
  uint8_t *visited = (uint8_t *)calloc(n, 1);
  size_t cap = 1 << 16;
  uint32_t *stack = (uint32_t *)malloc(cap * sizeof(uint32_t));
  if (!visited || !stack) return fail("oracle: out of host memory");
  uint64_t beta0 = 0, beta2 = 0;
  int rc = 0;
  for (size_t seed = 0; seed < n && !rc; ++seed)
    if (v[seed] != 0 && !visited[seed]) {
      beta0 += 1;
      rc = flood_fill(v, w, h, d, seed, 1, 1, visited, &stack, &cap);
    }
  memset(visited, 0, n);
#define BOUNDARY(i)                                                              \
  do {                                                                           \
    size_t s_ = (i);                                                             \
    if (!rc && v[s_] == 0 && !visited[s_])                                       \
      rc = flood_fill(v, w, h, d, s_, 0, 0, visited, &stack, &cap);             \
  } while (0)
  for (size_t y = 0; y < h; ++y)
    for (size_t x = 0; x < w; ++x) {
      BOUNDARY((0 * h + y) * w + x);
      BOUNDARY(((d - 1) * h + y) * w + x);
    }
  for (size_t z = 1; z + 1 < d; ++z)
    for (size_t x = 0; x < w; ++x) {
      BOUNDARY((z * h + 0) * w + x);
      BOUNDARY((z * h + (h - 1)) * w + x);
    }
  for (size_t z = 1; z + 1 < d; ++z)
    for (size_t y = 1; y + 1 < h; ++y) {
      BOUNDARY((z * h + y) * w + 0);
      BOUNDARY((z * h + y) * w + (w - 1));
    }
#undef BOUNDARY
  for (size_t seed = 0; seed < n && !rc; ++seed)
    if (v[seed] == 0 && !visited[seed]) {
      beta2 += 1;
      rc = flood_fill(v, w, h, d, seed, 0, 0, visited, &stack, &cap);
    }
  free(visited);
  free(stack);
  if (rc) return fail("oracle: out of host memory");
  int64_t beta1 = (int64_t)beta0 + (int64_t)beta2 - raw_euler;
  if (beta1 < 0) {
    snprintf(g_err, sizeof g_err,
             "computed a negative beta1 (%lld); check the 26/6 topology convention",
             (long long)beta1);
    return 1;
  }
  out->raw_euler = raw_euler;
  out->beta0 = beta0;
  out->beta1 = (uint64_t)beta1;
  out->beta2 = beta2;
  return 0;
}

/* bone_marrow_interface_faces, src/topology.rs:442-490 */
int64_t oracle_interface_faces(const uint8_t *bone, const uint8_t *roi, size_t w_, size_t h_,
                               size_t d_) {
  long w = (long)w_, h = (long)h_, d = (long)d_;
  int64_t faces = 0;
#pragma omp parallel for reduction(+ : faces) schedule(static)
  for (long z = 0; z < d; ++z)
    for (long y = 0; y < h; ++y)
      for (long x = 0; x < w; ++x) {
        if (!fg_at(roi, x, y, z, w, h, d)) continue;
        int c = fg_at(bone, x, y, z, w, h, d);
        if (x + 1 < w && fg_at(roi, x + 1, y, z, w, h, d) && c != fg_at(bone, x + 1, y, z, w, h, d))
          faces += 1;
        if (y + 1 < h && fg_at(roi, x, y + 1, z, w, h, d) && c != fg_at(bone, x, y + 1, z, w, h, d))
          faces += 1;
        if (z + 1 < d && fg_at(roi, x, y, z + 1, w, h, d) && c != fg_at(bone, x, y, z + 1, w, h, d))
          faces += 1;
      }
  return faces;
}

/* roi_boundary_bone_fraction, src/metrics.rs:166-202: the two counts */
void oracle_roi_boundary_counts(const uint8_t *abone, const uint8_t *roi, size_t w_, size_t h_,
                                size_t d_, uint64_t *bone_voxels, uint64_t *boundary_voxels) {
  long w = (long)w_, h = (long)h_, d = (long)d_;
  uint64_t b = 0, t = 0;
  static const long off[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
#pragma omp parallel for reduction(+ : b, t) schedule(static)
  for (long z = 0; z < d; ++z)
    for (long y = 0; y < h; ++y)
      for (long x = 0; x < w; ++x) {
        if (!fg_at(abone, x, y, z, w, h, d)) continue;
        b += 1;
        int touches = 0;
        for (int k = 0; k < 6 && !touches; ++k)
          touches = !fg_at(roi, x + off[k][0], y + off[k][1], z + off[k][2], w, h, d);
        t += touches;
      }
  *bone_voxels = b;
  *boundary_voxels = t;
}

/* The fields of Metrics beyond thickness/spacing, src/metrics.rs:123-164. */
typedef struct {
  double bone_volume, total_volume, bone_volume_fraction;
  oracle_topology topology;
  oracle_bonej bonej;
  double bonej_connectivity_density, connectivity, connectivity_density,
      roi_boundary_bone_fraction;
  double bone_surface, bone_surface_to_bone_volume, bone_surface_to_total_volume, mean_breadth;
  int64_t interface_faces;
  oracle_cubical cubical;
} oracle_rest;

int oracle_analyze_rest(const uint8_t *bone, const uint8_t *roi, size_t w, size_t h, size_t d,
                        const double spacing[3], oracle_rest *m) {
  size_t n = w * h * d;
  uint8_t *ab = (uint8_t *)malloc(n);
  if (!ab) return fail("oracle: out of host memory");
  oracle_phase_inside(bone, roi, n, 1, ab);
  uint64_t bone_voxels = 0, roi_voxels = 0;
  for (size_t i = 0; i < n; ++i) {
    bone_voxels += ab[i] != 0;
    roi_voxels += roi[i] != 0;
  }
  double vv = spacing[0] * spacing[1] * spacing[2];
  m->bone_volume = (double)bone_voxels * vv;
  m->total_volume = (double)roi_voxels * vv;
  m->bone_volume_fraction = m->bone_volume / m->total_volume;
  oracle_cubical_measures(ab, w, h, d, &m->cubical);
  if (oracle_analyze_topology(ab, w, h, d, m->cubical.raw_euler, &m->topology)) {
    free(ab);
    return 1;
  }
  oracle_bonej_connectivity(ab, w, h, d, m->cubical.raw_euler, &m->bonej);
  m->bonej_connectivity_density = m->bonej.connectivity / m->total_volume;
  m->connectivity = (double)m->topology.beta1;
  m->connectivity_density = m->connectivity / m->total_volume;
  m->interface_faces = oracle_interface_faces(ab, roi, w, h, d);
  m->bone_surface = (double)m->interface_faces * spacing[0] * spacing[1];
  m->bone_surface_to_bone_volume = m->bone_surface / m->bone_volume;
  m->bone_surface_to_total_volume = m->bone_surface / m->total_volume;
  m->mean_breadth = m->cubical.mean_breadth_lattice_units * spacing[0];
  uint64_t bb = 0, bt = 0;
  oracle_roi_boundary_counts(ab, roi, w, h, d, &bb, &bt);
  m->roi_boundary_bone_fraction = (double)bt / (double)bb;
  free(ab);
  return 0;
}
