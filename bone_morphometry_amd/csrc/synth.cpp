// synth.cpp -- deterministic synthetic trabecular volumes and ROI masks.
//
// A seeded N(0,1) white-noise field (counter-based: voxel i's value depends only
// on (seed, i)), smoothed by a separable truncated Gaussian (radius
// int(4*sigma + 0.5), mirror-reflect boundary "d c b a | a b c d | d c b a"),
// thresholded so that round(bv_tv * N) voxels lie strictly above the
// threshold.  Bone voxels are stored as 255, the rest as 0, like the 8-bit
// binary TIFFs the reference reads (src/tiff_io.rs:131 binarises value > 0).
// The result depends only on the arguments: the smoothing sums in a fixed order
// in f32 and the build disables FMA contraction.
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

inline uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

// One standard normal per voxel index (Box-Muller on two 53-bit uniforms).
inline float normal_at(uint64_t seed, uint64_t i) {
  uint64_t a = splitmix64(seed * 0xD1B54A32D192ED03ull ^ (2 * i));
  uint64_t b = splitmix64(seed * 0xD1B54A32D192ED03ull ^ (2 * i + 1));
  double u1 = ((a >> 11) + 1.0) * (1.0 / 9007199254740993.0);  // (0, 1]
  double u2 = (b >> 11) * (1.0 / 9007199254740992.0);          // [0, 1)
  return (float)(std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2));
}

inline long reflect_index(long i, long n) {
  // mirror with edge repeat: -1 -> 0, n -> n-1; folded until inside
  while (i < 0 || i >= n) {
    if (i < 0) i = -i - 1;
    if (i >= n) i = 2 * n - i - 1;
  }
  return i;
}

std::vector<float> gaussian_weights(double sigma) {
  long r = (long)(4.0 * sigma + 0.5);
  std::vector<double> w(2 * r + 1);
  double s = 0.0;
  for (long k = -r; k <= r; ++k) {
    w[k + r] = std::exp(-0.5 * (double)(k * k) / (sigma * sigma));
    s += w[k + r];
  }
  std::vector<float> out(w.size());
  for (size_t k = 0; k < w.size(); ++k) out[k] = (float)(w[k] / s);
  return out;
}

// Smooth along one axis: element (line, t) lives at base(line) + t*stride.
void smooth_axis(const float* in, float* out, size_t w, size_t h, size_t d, int axis,
                 double sigma) {
  if (sigma <= 0.0) {
    std::memcpy(out, in, w * h * d * sizeof(float));
    return;
  }
  std::vector<float> k = gaussian_weights(sigma);
  long r = (long)(k.size() / 2);
  size_t n = axis == 0 ? w : axis == 1 ? h : d;
  size_t stride = axis == 0 ? 1 : axis == 1 ? w : w * h;
  size_t lines = w * h * d / n;
#pragma omp parallel
  {
    std::vector<float> buf(n + 2 * r);
#pragma omp for schedule(static)
    for (size_t line = 0; line < lines; ++line) {
      size_t base;
      if (axis == 0) base = line * w;
      else if (axis == 1) base = (line / w) * w * h + (line % w);
      else base = line;
      for (long t = -r; t < (long)n + r; ++t)
        buf[t + r] = in[base + (size_t)reflect_index(t, (long)n) * stride];
      for (size_t t = 0; t < n; ++t) {
        float acc = 0.0f;
        for (long j = 0; j < (long)k.size(); ++j) acc += k[j] * buf[t + j];
        out[base + t * stride] = acc;
      }
    }
  }
}

}  // namespace

extern "C" {

// Fills out[w*h*d] (z-major, x fastest) with 0/255.  Returns the threshold.
double bm_synth_volume(uint8_t* out, size_t w, size_t h, size_t d, double sigma_x,
                       double sigma_y, double sigma_z, double bv_tv, uint64_t seed,
                       int threads) {
  if (threads > 0) omp_set_num_threads(threads);
  size_t n = w * h * d;
  std::vector<float> a(n), b(n);
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < n; ++i) a[i] = normal_at(seed, i);
  smooth_axis(a.data(), b.data(), w, h, d, 0, sigma_x);
  smooth_axis(b.data(), a.data(), w, h, d, 1, sigma_y);
  smooth_axis(a.data(), b.data(), w, h, d, 2, sigma_z);
  size_t m = (size_t)std::llround(bv_tv * (double)n);
  if (m > n) m = n;
  float thr;
  if (m == 0) {
    thr = INFINITY;
  } else if (m == n) {
    thr = -INFINITY;
  } else {
    std::copy(b.begin(), b.end(), a.begin());
    std::nth_element(a.begin(), a.begin() + (n - m - 1), a.end());
    thr = a[n - m - 1];
  }
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < n; ++i) out[i] = b[i] > thr ? 255 : 0;
  return (double)thr;
}

// z-aligned cylinder: (x-cx)^2 + (y-cy)^2 <= radius^2, every z.  Values 0/1.
void bm_synth_roi_cylinder(uint8_t* out, size_t w, size_t h, size_t d, double cx, double cy,
                           double radius) {
  size_t slice = w * h;
  for (size_t y = 0; y < h; ++y)
    for (size_t x = 0; x < w; ++x) {
      double dx = (double)x - cx, dy = (double)y - cy;
      out[y * w + x] = (dx * dx + dy * dy <= radius * radius) ? 1 : 0;
    }
#pragma omp parallel for schedule(static)
  for (size_t z = 1; z < d; ++z) std::memcpy(out + z * slice, out, slice);
}

// Ellipsoid: sum(((p - c)/a)^2) <= 1.  Values 0/1.
void bm_synth_roi_ellipsoid(uint8_t* out, size_t w, size_t h, size_t d, double cx, double cy,
                            double cz, double ax, double ay, double az) {
#pragma omp parallel for schedule(static)
  for (size_t z = 0; z < d; ++z)
    for (size_t y = 0; y < h; ++y)
      for (size_t x = 0; x < w; ++x) {
        double u = ((double)x - cx) / ax, v = ((double)y - cy) / ay, t = ((double)z - cz) / az;
        out[(z * h + y) * w + x] = (u * u + v * v + t * t <= 1.0) ? 1 : 0;
      }
}

}  // extern "C"
