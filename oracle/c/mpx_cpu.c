/*
 * CPU restatement of NNEDI3 and RAVU-lite in C (FP32 arithmetic, OpenMP over rows) --
 * TEST INFRASTRUCTURE ONLY: the CPU baseline that bench.py times beside the GPU, and a
 * second, independent check of the NumPy oracle in tests/.  Never called by the product.
 *
 * nnedi3_double_y follows nnedi3-nns16-win8x4.hook:20-104 (double_y + combine_y);
 * ravu_lite follows ravu-lite-r3.hook:22-131 / ravu-lite-ar-r3.hook:90-182.
 * Build: oracle/cbuild.py (gcc -O3 -fopenmp -shared).
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* One doubling pass along y on a (h, w) float plane -> (2h, w): row 2y = input row y,
 * row 2y+1 = prediction between rows y and y+1.  wt = [nns][2K+2] rows (w1, w2, b1, b2),
 * window k = (dx+3)*wh + (dy-ymin), dx in [-3,4], dy in [ymin, ymin+wh-1]. */
void nnedi3_double_y(const float* in, int w, int h, float* out, const float* wt, int nns, int wh) {
  const int K = 8 * wh, ymin = -(wh / 2 - 1);
#pragma omp parallel for schedule(dynamic, 4)
  for (int y = 0; y < h; ++y) {
    float x[48];
    for (int xx = 0; xx < w; ++xx) {
      float sum = 0.f, sumsq = 0.f;
      for (int a = 0; a < 8; ++a)
        for (int b = 0; b < wh; ++b) {
          const float v = in[(size_t)clampi(y + ymin + b, 0, h - 1) * w + clampi(xx - 3 + a, 0, w - 1)];
          x[a * wh + b] = v;
          sum += v;
          sumsq += v * v;
        }
      const float m = sum / (float)K;
      const float var = sumsq / (float)K - m * m;
      const float is = var >= 1.192092896e-7f ? 1.0f / sqrtf(var) : 0.f;
      const float sd = var * is;
      float wsum = 0.f, vsum = 0.f;
      for (int n = 0; n < nns; ++n) {
        const float* r = wt + (size_t)n * (2 * K + 2);
        float s1 = 0.f, s2 = 0.f;
        for (int k = 0; k < K; ++k) {
          s1 += r[k] * x[k];
          s2 += r[K + k] * x[k];
        }
        s1 = expf(s1 * is + r[2 * K]);
        s2 = s2 * is + r[2 * K + 1];
        wsum += s1;
        vsum += s1 * (s2 / (1.0f + fabsf(s2)));
      }
      float p = m + 5.0f * vsum / wsum * sd;
      p = p < 0.f ? 0.f : (p > 1.f ? 1.f : p);
      out[(size_t)(2 * y) * w + xx] = in[(size_t)y * w + xx];
      out[(size_t)(2 * y + 1) * w + xx] = p;
    }
  }
}

/* Transpose helper for the double_x pass (the x-pass weights are the transposed y-pass). */
void transpose_f32(const float* in, int w, int h, float* out) {
#pragma omp parallel for
  for (int y = 0; y < h; ++y)
    for (int x = 0; x < w; ++x) out[(size_t)x * h + y] = in[(size_t)y * w + x];
}

/* RAVU-lite(-ar) 2x, luma.  lut = FP16-rounded weights as float, [288][ntex][4]. */
void ravu_lite(const float* in, int w, int h, float* out, const float* lut, int r, int ar,
               const double* gauss) {
  const int n = 2 * r - 1, c = (n * n - 1) / 2, ntex = c + 1, lo = -(r - 1);
  const int p0 = n <= 4 ? 0 : 1, p1 = n <= 4 ? n : n - 1;
#pragma omp parallel for schedule(dynamic, 4)
  for (int y = 0; y < h; ++y) {
    float s[49];
    for (int xx = 0; xx < w; ++xx) {
      for (int X = 0; X < n; ++X)
        for (int Y = 0; Y < n; ++Y)
          s[X * n + Y] = in[(size_t)clampi(y + lo + Y, 0, h - 1) * w + clampi(xx + lo + X, 0, w - 1)];
      float a = 0.f, b = 0.f, d = 0.f;
      int gi = 0;
      for (int X = p0; X < p1; ++X)
        for (int Y = p0; Y < p1; ++Y, ++gi) {
          const float gx = X == 0 ? s[n + Y] - s[Y]
                         : (X == n - 1 ? s[X * n + Y] - s[(X - 1) * n + Y] : (s[(X + 1) * n + Y] - s[(X - 1) * n + Y]) / 2.0f);
          const float gy = Y == 0 ? s[X * n + 1] - s[X * n]
                         : (Y == n - 1 ? s[X * n + Y] - s[X * n + Y - 1] : (s[X * n + Y + 1] - s[X * n + Y - 1]) / 2.0f);
          const float g = (float)gauss[gi];
          a += gx * gx * g;
          b += gx * gy * g;
          d += gy * gy * g;
        }
      const float T = a + d, D = a * d - b * b;
      const float delta = sqrtf(fmaxf(T * T / 4.0f - D, 0.f));
      const float L1 = T / 2.0f + delta, L2 = T / 2.0f - delta;
      const float sL1 = sqrtf(L1), sL2 = sqrtf(L2);
      float theta = fabsf(b) < 1.192092896e-7f ? 0.f : 0.f;
      if (!(fabsf(b) < 1.192092896e-7f)) {
        const float at = atan2f(L1 - a, b) + 3.141592653589793f;
        theta = at - 3.141592653589793f * floorf(at / 3.141592653589793f);
      }
      float mu = (sL1 - sL2) / (sL1 + sL2);
      if (sL1 + sL2 < 1.192092896e-7f) mu = 0.f;
      const int angle = (int)floorf(theta * 24.0f / 3.141592653589793f);
      const int strength = sL1 >= 0.016f ? (sL1 >= 0.05f ? 3 : 2) : (sL1 >= 0.004f ? 1 : 0);
      const int coh = mu >= 0.5f ? 2 : (mu >= 0.25f ? 1 : 0);
      const int row = clampi((angle * 4 + strength) * 3 + coh, 0, 287);
      const float* L = lut + (size_t)row * ntex * 4;
      float res[4] = {0, 0, 0, 0}, hi[4] = {0, 0, 0, 0}, lw[4] = {0, 0, 0, 0}, hi2[4] = {0, 0, 0, 0},
            lo2[4] = {0, 0, 0, 0};
      for (int i = 0; i <= c; ++i) {
        const float* wv = L + 4 * i;
        const float li = s[i], lj = i < c ? s[n * n - 1 - i] : 0.f;
        for (int k = 0; k < 4; ++k) res[k] += li * wv[k] + (i < c ? lj * wv[3 - k] : 0.f);
        if (ar) {
          const int X = i / n + lo, Y = i % n + lo;
          if (X * X + Y * Y <= 4) {
            for (int t = 0; t < (i < c ? 2 : 1); ++t) {
              const float l = t ? lj : li;
              const float pp = powf(0.1f + l, 32.f), mm = powf(1.1f - l, 32.f);
              for (int k = 0; k < 4; ++k) {
                const float wg = fmaxf(t ? wv[3 - k] : wv[k], 0.f);
                hi[k] += pp * wg;
                lw[k] += mm * wg;
                hi2[k] += pp * (0.1f + l) * wg;
                lo2[k] += mm * (1.1f - l) * wg;
              }
            }
          }
        }
      }
      for (int k = 0; k < 4; ++k) {
        float v = res[k];
        if (ar) {
          const float hv = hi2[k] / hi[k] - 0.1f, lv = 1.1f - lo2[k] / lw[k];
          const float cl = fminf(fmaxf(v, lv), hv);
          v = v * (1.0f - 0.8f) + cl * 0.8f;
        } else {
          v = fminf(fmaxf(v, 0.f), 1.f);
        }
        const int dx = k / 2, dy = k % 2;
        out[(size_t)(2 * y + dy) * (2 * w) + 2 * xx + dx] = v;
      }
    }
  }
}
