/* CPU oracle of the learned per-cell smoother's network (TEST INFRASTRUCTURE
 * ONLY, see oracle.h).  The network has no counterpart in the reference: it
 * is pinned by central finite differences (gradient), closed forms of special
 * weight sets (forward), the Adam update's published form and bit patterns of
 * bfloat16 rounding.
 *
 * Network: y = W3 g(W2 g(W1 x + b1) + b2) + b3 with g(a) = a/2 (1 + erf(a/sqrt2))
 * (GELU, Hendrycks & Gimpel 2016); parameters packed as
 * W1[64][20] b1[64] W2[64][64] b2[64] W3[10][64] b3[10].  Loss of n rows =
 * sum |y - t|^2 / n.  Adam: Kingma & Ba 2015, Algorithm 1.                  */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "oracle.h"

enum { NI = 20, NH = 64, NO = 10 };
enum { OW1 = 0, OB1 = OW1 + NH * NI, OW2 = OB1 + NH, OB2 = OW2 + NH * NH, OW3 = OB2 + NH,
       OB3 = OW3 + NO * NH, NP = OB3 + NO };

static double gelu(double a) { return 0.5 * a * (1.0 + erf(a / sqrt(2.0))); }
static double gelu_d(double a) {
  return 0.5 * (1.0 + erf(a / sqrt(2.0))) + a * exp(-0.5 * a * a) / sqrt(2.0 * M_PI);
}

int orc_mlp_n_params(void) { return NP; }

uint16_t orc_bf16(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40); /* quiet NaN */
  /* round to nearest, ties to even: add 0x7fff plus the lowest kept bit */
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

static double bf(double v) {
  const uint32_t u = (uint32_t)orc_bf16((float)v) << 16;
  float f;
  memcpy(&f, &u, 4);
  return f;
}

/* one row; q != 0 rounds weights and layer inputs to bfloat16 (the device's
 * MFMA inference), keeping biases and accumulation exact */
static void forward_row(const double* p, const double* x, int q, double* a1, double* h1, double* a2,
                        double* h2, double* y) {
  for (int o = 0; o < NH; ++o) {
    double s = 0;
    for (int k = 0; k < NI; ++k) s += (q ? bf(p[OW1 + o * NI + k]) * bf(x[k]) : p[OW1 + o * NI + k] * x[k]);
    a1[o] = s + p[OB1 + o];
    h1[o] = gelu(a1[o]);
  }
  for (int o = 0; o < NH; ++o) {
    double s = 0;
    for (int k = 0; k < NH; ++k) s += (q ? bf(p[OW2 + o * NH + k]) * bf(h1[k]) : p[OW2 + o * NH + k] * h1[k]);
    a2[o] = s + p[OB2 + o];
    h2[o] = gelu(a2[o]);
  }
  for (int o = 0; o < NO; ++o) {
    double s = 0;
    for (int k = 0; k < NH; ++k) s += (q ? bf(p[OW3 + o * NH + k]) * bf(h2[k]) : p[OW3 + o * NH + k] * h2[k]);
    y[o] = s + p[OB3 + o];
  }
}

int orc_mlp_forward(const double* params, const double* x, int64_t n, int bf16, double* y) {
  double a1[NH], h1[NH], a2[NH], h2[NH];
  for (int64_t r = 0; r < n; ++r) forward_row(params, x + r * NI, bf16, a1, h1, a2, h2, y + r * NO);
  return 0;
}

int orc_mlp_loss_grad(const double* p, const double* x, const double* t, int64_t n, double* grad, double* loss) {
  double a1[NH], h1[NH], a2[NH], h2[NH], y[NO], dy[NO], d2[NH], d1[NH];
  memset(grad, 0, sizeof(double) * NP);
  double l = 0;
  for (int64_t r = 0; r < n; ++r) {
    const double* xr = x + r * NI;
    forward_row(p, xr, 0, a1, h1, a2, h2, y);
    for (int o = 0; o < NO; ++o) {
      const double d = y[o] - t[r * NO + o];
      l += d * d;
      dy[o] = 2.0 * d / (double)n;
    }
    for (int k = 0; k < NH; ++k) {
      double s = 0;
      for (int o = 0; o < NO; ++o) s += p[OW3 + o * NH + k] * dy[o];
      d2[k] = s * gelu_d(a2[k]);
    }
    for (int k = 0; k < NH; ++k) {
      double s = 0;
      for (int o = 0; o < NH; ++o) s += p[OW2 + o * NH + k] * d2[o];
      d1[k] = s * gelu_d(a1[k]);
    }
    for (int o = 0; o < NO; ++o) {
      for (int k = 0; k < NH; ++k) grad[OW3 + o * NH + k] += dy[o] * h2[k];
      grad[OB3 + o] += dy[o];
    }
    for (int o = 0; o < NH; ++o) {
      for (int k = 0; k < NH; ++k) grad[OW2 + o * NH + k] += d2[o] * h1[k];
      grad[OB2 + o] += d2[o];
      for (int k = 0; k < NI; ++k) grad[OW1 + o * NI + k] += d1[o] * xr[k];
      grad[OB1 + o] += d1[o];
    }
  }
  *loss = l / (double)n;
  return 0;
}

int orc_mlp_loss(const double* p, const double* x, const double* t, int64_t n, double* loss) {
  double a1[NH], h1[NH], a2[NH], h2[NH], y[NO];
  double l = 0;
  for (int64_t r = 0; r < n; ++r) {
    forward_row(p, x + r * NI, 0, a1, h1, a2, h2, y);
    for (int o = 0; o < NO; ++o) l += (y[o] - t[r * NO + o]) * (y[o] - t[r * NO + o]);
  }
  *loss = l / (double)n;
  return 0;
}

/* Adam with bias correction, step = 1, 2, ... */
int orc_adam(double* p, const double* g, double* m, double* v, int64_t n, double lr, double b1, double b2,
             double eps, int64_t step) {
  const double c1 = 1.0 - pow(b1, (double)step), c2 = 1.0 - pow(b2, (double)step);
  for (int64_t e = 0; e < n; ++e) {
    m[e] = b1 * m[e] + (1.0 - b1) * g[e];
    v[e] = b2 * v[e] + (1.0 - b2) * g[e] * g[e];
    p[e] -= lr * (m[e] / c1) / (sqrt(v[e] / c2) + eps);
  }
  return 0;
}
