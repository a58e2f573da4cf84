// Host-side float64 cosmology tables: the 2LPT growth ODE with its Omega_m
// forward sensitivities, and the top-hat variance integral.  No device code:
// these run on the host in microseconds to milliseconds, so that a request's
// tables cost far less than its simulation.
//
// Growth (pmwd/boltzmann.py:163-229, cosmology.py:185-269): the state
// (G1, G1', G2, G2', S1, S1', S2, S2') in ln a, primes d/dln a, S = dG/dOmega_m
// at fixed Omega_k, w_0, w_a (Omega_de = 1 - Omega_m - Omega_k):
//   G1'' = -(3 + h - om) G1 - (4 + h) G1'
//   G2'' = om G1^2 - (8 + 2h - om) G2 - (6 + h) G2'
// with h = dlnH/dlna and om = 3/2 Omega_m(a), and the S equations their
// Omega_m derivatives.  Integrator: adaptive Dormand-Prince 5(4) as
// pmwd/ode_util.py:60-186 (RMS error ratio over atol + rtol max(|y0|, |y1|)
// of the first nerr components, step factor clip(0.9 ratio^-1/5, 0.2 or 1,
// 10), quartic dense output at the requested times).
//
// Top-hat variance (pmwd/boltzmann.py:272-291): sigma^2(R) =
// int k^3 P(k) W(kR)^2 / (2 pi^2) dln k on the uniform ln k grid, W(x) =
// 3 (sin x - x cos x) / x^3 (1 - x^2/10 below x = 1e-3), Simpson weights for
// an odd number of nodes, else the trapezoid rule.
#include <math.h>
#include <string.h>

#include "pmx_internal.h"

namespace {

struct Bg {
  double om, ok, ode, w0, wa;
};

struct BgA {
  double e2, e2a, h, oma, e2_om, h_om, oma_om;
};

// E2(a), dE2/da, H_deriv, Omega_m(a) and their Omega_m derivatives
BgA background(double a, const Bg& b) {
  const double p = -3.0 * (1.0 + b.w0 + b.wa);
  const double de = pow(a, p) * exp(-3.0 * b.wa * (1.0 - a));
  const double am3 = pow(a, -3.0), am4 = pow(a, -4.0);
  BgA r;
  r.e2 = b.om * am3 + b.ok * pow(a, -2.0) + b.ode * de;
  const double dde = (p / a + 3.0 * b.wa) * de;
  r.e2a = -3.0 * b.om * am4 - 2.0 * b.ok * am3 + b.ode * dde;
  r.h = 0.5 * a * r.e2a / r.e2;
  r.oma = b.om / (pow(a, 3.0) * r.e2);
  r.e2_om = am3 - de;
  const double e2a_om = -3.0 * am4 - dde;
  r.h_om = 0.5 * a * (e2a_om * r.e2 - r.e2a * r.e2_om) / (r.e2 * r.e2);
  r.oma_om = am3 * (r.e2 - b.om * r.e2_om) / (r.e2 * r.e2);
  return r;
}

void growth_rhs(const Bg& b, const double* y, double lna, double* f) {
  const BgA g = background(exp(lna), b);
  const double h = g.h, om = 1.5 * g.oma, dh = g.h_om, dom = 1.5 * g.oma_om;
  const double G1 = y[0], G1p = y[1], G2 = y[2], G2p = y[3];
  const double S1 = y[4], S1p = y[5], S2 = y[6], S2p = y[7];
  f[0] = G1p;
  f[1] = -(3.0 + h - om) * G1 - (4.0 + h) * G1p;
  f[2] = G2p;
  f[3] = om * (G1 * G1) - (8.0 + 2.0 * h - om) * G2 - (6.0 + h) * G2p;
  f[4] = S1p;
  f[5] = -(3.0 + h - om) * S1 - (4.0 + h) * S1p - (dh - dom) * G1 - dh * G1p;
  f[6] = S2p;
  f[7] = 2.0 * om * G1 * S1 - (8.0 + 2.0 * h - om) * S2 - (6.0 + h) * S2p + dom * (G1 * G1) -
         (2.0 * dh - dom) * G2 - dh * G2p;
}

constexpr int NY = 8;

// Dormand & Prince (1980) 5(4) coefficients
const double DP_C[7] = {0.0, 1.0 / 5, 3.0 / 10, 4.0 / 5, 8.0 / 9, 1.0, 1.0};
const double DP_A[7][6] = {
    {0, 0, 0, 0, 0, 0},
    {1.0 / 5, 0, 0, 0, 0, 0},
    {3.0 / 40, 9.0 / 40, 0, 0, 0, 0},
    {44.0 / 45, -56.0 / 15, 32.0 / 9, 0, 0, 0},
    {19372.0 / 6561, -25360.0 / 2187, 64448.0 / 6561, -212.0 / 729, 0, 0},
    {9017.0 / 3168, -355.0 / 33, 46732.0 / 5247, 49.0 / 176, -5103.0 / 18656, 0},
    {35.0 / 384, 0, 500.0 / 1113, 125.0 / 192, -2187.0 / 6784, 11.0 / 84},
};
const double DP_B5[7] = {35.0 / 384, 0, 500.0 / 1113, 125.0 / 192, -2187.0 / 6784, 11.0 / 84, 0};
// error weights: the 5th-order weights minus the embedded 4th-order ones
const double DP_BERR[7] = {35.0 / 384 - 1951.0 / 21600, 0, 500.0 / 1113 - 22642.0 / 50085,
                           125.0 / 192 - 451.0 / 720, -2187.0 / 6784 - -12231.0 / 42400,
                           11.0 / 84 - 649.0 / 6300, -1.0 / 60.0};
// dense output at the step midpoint (Shampine 1986), halved
const double DP_BMID[7] = {6025192743.0 / 30085553152 / 2, 0, 51252292925.0 / 65400821598 / 2,
                           -2691868925.0 / 45128329728 / 2, 187940372067.0 / 1594534317056 / 2,
                           -1776094331.0 / 19743644256 / 2, 11237099.0 / 235043384 / 2};

struct Step {
  double y1[NY], K[7][NY], err[NY];
};

void dp_step(const Bg& b, double t, const double* y, const double* f0, double h, Step& s) {
  memcpy(s.K[0], f0, sizeof(double) * NY);
  double yi[NY];
  for (int i = 1; i < 7; ++i) {
    for (int c = 0; c < NY; ++c) {
      double acc = 0.0;
      for (int j = 0; j < i; ++j) acc += DP_A[i][j] * s.K[j][c];
      yi[c] = y[c] + h * acc;
    }
    growth_rhs(b, yi, t + DP_C[i] * h, s.K[i]);
  }
  for (int c = 0; c < NY; ++c) {
    double a5 = 0.0, ae = 0.0;
    for (int j = 0; j < 7; ++j) {
      a5 += DP_B5[j] * s.K[j][c];
      ae += DP_BERR[j] * s.K[j][c];
    }
    s.y1[c] = h * a5 + y[c];
    s.err[c] = h * ae;
  }
}

// coefficients (highest power first) of the quartic through the step
void quartic(const double* y0, const Step& s, double h, double poly[5][NY]) {
  for (int c = 0; c < NY; ++c) {
    double am = 0.0;
    for (int j = 0; j < 7; ++j) am += DP_BMID[j] * s.K[j][c];
    const double ym = y0[c] + h * am, y1 = s.y1[c];
    const double d0 = h * s.K[0][c], d1 = h * s.K[6][c];
    poly[0][c] = -2 * d0 + 2 * d1 - 8 * y0[c] - 8 * y1 + 16 * ym;
    poly[1][c] = 5 * d0 - 3 * d1 + 18 * y0[c] + 14 * y1 - 32 * ym;
    poly[2][c] = -4 * d0 + d1 - 11 * y0[c] - 5 * y1 + 16 * ym;
    poly[3][c] = d0;
    poly[4][c] = y0[c];
  }
}

double l2_norm(const double* v, const double* scale, int n) {
  double s = 0.0;
  for (int i = 0; i < n; ++i) {
    const double r = v[i] / scale[i];
    s += r * r;
  }
  return sqrt(s);
}

// Hairer-Norsett-Wanner initial step (Solving ODEs I, II.4)
double initial_step(const Bg& b, double t0, const double* y0, const double* f0, int ne,
                    double rtol, double atol) {
  double scale[NY];
  for (int i = 0; i < ne; ++i) scale[i] = atol + fabs(y0[i]) * rtol;
  const double d0 = l2_norm(y0, scale, ne), d1 = l2_norm(f0, scale, ne);
  const double h0 = (d0 < 1e-5 || d1 < 1e-5) ? 1e-6 : 0.01 * d0 / d1;
  double y1[NY], f1[NY], df[NY];
  for (int i = 0; i < NY; ++i) y1[i] = y0[i] + h0 * f0[i];
  growth_rhs(b, y1, t0 + h0, f1);
  for (int i = 0; i < ne; ++i) df[i] = f1[i] - f0[i];
  const double d2 = l2_norm(df, scale, ne) / h0;
  double h1;
  if (d1 <= 1e-15 && d2 <= 1e-15)
    h1 = fmax(1e-6, h0 * 1e-3);
  else
    h1 = pow(0.01 / fmax(d1, d2), 1.0 / 5);
  return fmin(100 * h0, h1);
}

}  // namespace

extern "C" int pmx_growth_ode(double Omega_m, double Omega_k, double w_0, double w_a,
                              const double* lna, int na, const double* y0, double rtol,
                              double atol, double dt0, int nerr, double* Y, double* Yd) {
  PMX_REQUIRE(lna && y0 && Y && Yd && na >= 1, "pmx_growth_ode: bad argument");
  PMX_REQUIRE(nerr >= 1 && nerr <= NY, "pmx_growth_ode: nerr must be in 1..8");
  PMX_REQUIRE(rtol >= 0 && atol >= 0 && (rtol > 0 || atol > 0),
              "pmx_growth_ode: tolerances must be non-negative and not both zero");
  for (int i = 1; i < na; ++i)
    PMX_REQUIRE(lna[i] >= lna[i - 1], "pmx_growth_ode: times must not decrease");
  const Bg b = {Omega_m, Omega_k, 1.0 - (Omega_m + Omega_k), w_0, w_a};
  double y[NY], fy[NY];
  memcpy(y, y0, sizeof y);
  growth_rhs(b, y, lna[0], fy);
  double h = dt0 > 0 ? dt0 : initial_step(b, lna[0], y, fy, nerr, rtol, atol);
  double t = lna[0], t_last = lna[0];
  double poly[5][NY];
  for (int k = 0; k < 5; ++k) memcpy(poly[k], y, sizeof y);
  memcpy(Y, y, sizeof y);
  long nstep = 0;
  Step s;
  for (int i = 1; i < na; ++i) {
    const double target = lna[i];
    while (t < target && h > 0) {
      PMX_REQUIRE(++nstep <= 100000, "pmx_growth_ode: too many steps");
      dp_step(b, t, y, fy, h, s);
      double ssum = 0.0;
      for (int c = 0; c < nerr; ++c) {
        const double sc = atol + rtol * fmax(fabs(y[c]), fabs(s.y1[c]));
        const double r = s.err[c] / sc;
        ssum += r * r;
      }
      const double ratio = sqrt(ssum / nerr);
      double h_new;
      if (ratio == 0.0) {
        h_new = 10.0 * h;
      } else {
        const double lo = ratio < 1 ? 1.0 : 0.2;
        h_new = h * fmin(10.0, fmax(0.9 * pow(ratio, -0.2), lo));
      }
      if (ratio <= 1) {
        quartic(y, s, h, poly);
        t_last = t;
        t = t + h;
        memcpy(y, s.y1, sizeof y);
        memcpy(fy, s.K[6], sizeof fy);
      }
      h = h_new;
    }
    const double x = (target - t_last) / (t - t_last);
    for (int c = 0; c < NY; ++c) {
      double v = poly[0][c];
      for (int k = 1; k < 5; ++k) v = v * x + poly[k][c];
      Y[i * NY + c] = v;
    }
  }
  for (int i = 0; i < na; ++i) growth_rhs(b, Y + i * NY, lna[i], Yd + i * NY);
  return PMX_OK;
}

extern "C" int pmx_tophat_var(const double* k, const double* P, int nk, const double* R, int nR,
                              double* out) {
  PMX_REQUIRE(k && P && R && out && nk >= 2 && nR >= 0, "pmx_tophat_var: bad argument");
  const double h = log(k[1]) - log(k[0]);
  const double two_pi2 = 2.0 * M_PI * M_PI;
  for (int r = 0; r < nR; ++r) {
    double sum = 0.0, prev = 0.0;
    for (int j = 0; j < nk; ++j) {
      const double x = R[r] * k[j];
      double W;
      if (x < 1e-3) {
        W = 1.0 - x * x / 10.0;
      } else {
        double sx, cx;
        sincos(x, &sx, &cx);
        W = 3.0 * (sx - x * cx) / (x * x * x);
      }
      const double f = pow(k[j], 3.0) * P[j] / two_pi2 * (W * W);
      if (nk % 2 == 1) {
        const double w = (j == 0 || j == nk - 1) ? 1.0 : ((j & 1) ? 4.0 : 2.0);
        sum += f * w;
      } else if (j > 0) {
        sum += 0.5 * (prev + f) * (log(k[j]) - log(k[j - 1]));
      }
      prev = f;
    }
    out[r] = nk % 2 == 1 ? sum * (h / 3.0) : sum;
  }
  return PMX_OK;
}
