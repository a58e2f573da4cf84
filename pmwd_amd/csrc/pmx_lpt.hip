// 2LPT initial conditions on the particle grid (pmwd/lpt.py:13-212), with
// the library's FFT passes.
//
// Order 1: the linear modes over the cell volume, already a full 3D
// spectrum, enter the Poisson x pass directly (no forward x FFT): Green's
// function -1/k^2, -i k_d and the inverse x FFT, then the inverse y and z
// passes give the three components of the displacement field, and one pass
// over the particles adds D1 g and a^2 H D1' g to disp and vel.
// Order 2: each strain -k_i k_j pot_1 (lpt.py:13-37; Nyquist zeroed for
// i != j) is formed from the modes in one pass and inverted; the source
// L = s00 s22 + s00 s11 + s11 s22 - s01^2 - s02^2 - s12^2 (lpt.py:40-76, in
// that order) accumulates in the context's extra mesh, whose r2c goes
// through the forward Poisson solve like a density, and a second particle
// pass adds D2 g and a^2 H D2' g.
#include "pmx_internal.h"

struct LptK {
  const float* k[3];
  const float* km[3];
};

// dst (padded rows of n2p complex) from the modes (rows of n2h complex):
// i < 0: src = modes / cell_vol; else the strain spectrum
// (-k_i k_j) * pot, pot = -src / k^2 (0 at k = 0), k Nyquist-zeroed when i != j.
__global__ void __launch_bounds__(PMX_TPB)
    lpt_spec_k(const float2* __restrict__ modes, float2* __restrict__ dst, int n1, int n2h,
               int n2p, long long total, float cell_vol, LptK kt, int i, int j) {
  const long long t = (long long)blockIdx.x * PMX_TPB + threadIdx.x;
  if (t >= total) return;
  const long long row = t / n2h;
  const int kz = (int)(t - row * n2h);
  const int y = (int)(row % n1);
  const int x = (int)(row / n1);
  const float2 m = modes[t];
  float2 src = make_float2(m.x / cell_vol, m.y / cell_vol);
  if (i >= 0) {
    const float kx = kt.k[0][x], ky = kt.k[1][y], kk = kt.k[2][kz];
    float k2 = kx * kx;
    k2 = k2 + ky * ky;
    k2 = k2 + kk * kk;
    float2 pot = make_float2(0.f, 0.f);
    if (k2 != 0.f) pot = make_float2(-src.x / k2, -src.y / k2);
    const float mx = kt.km[0][x], my = kt.km[1][y], mz = kt.km[2][kz];
    auto sel = [](int a, float u, float w, float z) { return a == 0 ? u : (a == 1 ? w : z); };
    const float ki = i == j ? sel(i, kx, ky, kk) : sel(i, mx, my, mz);
    const float kj = i == j ? ki : sel(j, mx, my, mz);
    const float f = -ki * kj;
    src = make_float2(f * pot.x, f * pot.y);
  }
  dst[row * n2p + kz] = src;
}

// The source L over the padded real layout: mode 0 sets
// L = s0 s2 + s0 s1 + s1 s2 (the diagonal strains s00, s11, s22); mode 1
// subtracts s0^2 (an off-diagonal strain).
__global__ void __launch_bounds__(PMX_TPB)
    lpt_source_k(float* __restrict__ L, const float* __restrict__ s0, const float* __restrict__ s1,
                 const float* __restrict__ s2, int n2, int rs, long long total, int mode) {
  const long long t = (long long)blockIdx.x * PMX_TPB + threadIdx.x;
  if (t >= total) return;
  const long long row = t / n2;
  const long long i = row * rs + (t - row * n2);
  if (mode == 0) {
    const float a = s0[i], b = s1[i], c = s2[i];
    float l = a * c;
    l = l + a * b;
    l = l + b * c;
    L[i] = l;
  } else {
    const float s = s0[i];
    L[i] = L[i] - s * s;
  }
}

// disp += D g, vel += v g for particle p = grid cell p (C order), g the
// three displacement-field meshes in the padded real layout (lpt.py:196-208);
// the block's disp and vel rows move through LDS.
__global__ void __launch_bounds__(PMX_TPB)
    lpt_move_k(float* __restrict__ disp, float* __restrict__ vel, const float* __restrict__ g0,
               const float* __restrict__ g1, const float* __restrict__ g2, long long n, int n2,
               int rs, float D, float v) {
  constexpr int T3 = 3 * PMX_TPB;
  __shared__ __align__(16) float sh[2 * T3];  // disp | vel
  const long long b0 = (long long)blockIdx.x * PMX_TPB;
  const int cnt = (int)(n - b0 < PMX_TPB ? n - b0 : PMX_TPB);
  pmx_stage_in3<2, PMX_TPB>(sh, {{disp, vel}}, b0, cnt);
  __syncthreads();
  const int q = threadIdx.x;
  if (q < cnt) {
    const long long p = b0 + q;
    const long long row = p / n2;
    const long long i = row * rs + (p - row * n2);
    const float g[3] = {g0[i], g1[i], g2[i]};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      float x = sh[3 * q + d];
      x = x + D * g[d];
      sh[3 * q + d] = x;
      float u = sh[T3 + 3 * q + d];
      u = u + v * g[d];
      sh[T3 + 3 * q + d] = u;
    }
  }
  __syncthreads();
  pmx_stage_out3<2, PMX_TPB>({{disp, vel}}, sh, b0, cnt);
}

static LptK lpt_k(const pmx_ctx* c) {
  LptK kt;
  for (int a = 0; a < 3; ++a) {
    kt.k[a] = c->k[a];
    kt.km[a] = c->km[a];
  }
  return kt;
}

static int spec(pmx_ctx* c, const float* modes, float2* dst, float cell_vol, int i, int j) {
  const long long total = (long long)c->n[0] * c->n[1] * c->n2h;
  pmx_tic(c, PMX_T_OTHER);
  lpt_spec_k<<<pmx_blocks(total), PMX_TPB, 0, c->stream>>>(
      (const float2*)modes, dst, c->n[1], c->n2h, c->n2p, total, cell_vol, lpt_k(c), i, j);
  pmx_toc(c, PMX_T_OTHER);
  PMX_LAUNCH_CHECK();
  return PMX_OK;
}

// inverse 3D real FFT of a full spectrum in a padded buffer, in place
static int inverse3(pmx_ctx* c, float2* buf) {
  int rc;
  float2* f[1] = {buf};
  if ((rc = pmx_fft_strided(c, f, 1, 0, 1))) return rc;
  if ((rc = pmx_fft_strided(c, f, 1, 1, 1))) return rc;
  return pmx_fft_c2r_inplace(c, (float*)buf, (float)(1.0 / (double)c->M));
}

// -i k pot in f[0..2] (padded real, scaled 1/M) from f[0]: after the z and y
// forward passes (kin = 0) or a full spectrum (kin = 1)
static int neg_grad3(pmx_ctx* c, float2* const* f, int kin) {
  int rc;
  if ((rc = pmx_poisson_x(c, f, 0, kin))) return rc;
  if ((rc = pmx_fft_strided(c, f, 1, 1, 1))) return rc;
  if ((rc = pmx_fft_ypair(c, f, 1))) return rc;
  const float scale = (float)(1.0 / (double)c->M);
  for (int d = 0; d < 3; ++d)
    if ((rc = pmx_fft_c2r_inplace(c, (float*)f[d], scale))) return rc;
  return PMX_OK;
}

static int move(pmx_ctx* c, float* disp, float* vel, float2* const* g, float D, float v) {
  pmx_tic(c, PMX_T_OTHER);
  lpt_move_k<<<pmx_blocks(c->M), PMX_TPB, 0, c->stream>>>(
      disp, vel, (const float*)g[0], (const float*)g[1], (const float*)g[2], c->M, c->n[2],
      c->rs, D, v);
  pmx_toc(c, PMX_T_OTHER);
  PMX_LAUNCH_CHECK();
  return PMX_OK;
}

static int source(pmx_ctx* c, float2* const* s, int mode) {
  const long long total = c->M;
  pmx_tic(c, PMX_T_OTHER);
  lpt_source_k<<<pmx_blocks(total), PMX_TPB, 0, c->stream>>>(
      (float*)c->lbuf, (const float*)s[0], (const float*)s[1], (const float*)s[2], c->n[2], c->rs,
      total, mode);
  pmx_toc(c, PMX_T_OTHER);
  PMX_LAUNCH_CHECK();
  return PMX_OK;
}

extern "C" int pmx_lpt(pmx_ctx* c, const float* modes, float cell_vol, int order,
                       const float* fac, float* disp, float* vel) {
  PMX_ENTRY(c);
  PMX_REQUIRE(c && order >= 0 && order <= 2, "pmx_lpt: bad argument (order %d)", order);
  if (order == 0) return PMX_OK;
  PMX_REQUIRE(modes && fac && disp && vel, "pmx_lpt: NULL pointer");
  PMX_REQUIRE(cell_vol > 0.f, "pmx_lpt: cell volume must be positive");
  int rc;
  // every allocation before the first write to disp and vel
  if (order > 1 && !c->lbuf &&
      (rc = pmx_dev_alloc(c, (void**)&c->lbuf, (size_t)c->Mc * sizeof(float2))))
    return rc;
  float2* f[3] = {c->cbuf[0], c->cbuf[1], c->cbuf[2]};
  // first order: -i k (-(modes / V_cell) / k^2), inverted
  if ((rc = spec(c, modes, f[0], cell_vol, -1, -1))) return rc;
  if ((rc = neg_grad3(c, f, 1))) return rc;
  if ((rc = move(c, disp, vel, f, fac[0], fac[1]))) return rc;
  if (order == 1) return PMX_OK;
  // second order: the diagonal strains, then the off-diagonal ones
  for (int d = 0; d < 3; ++d) {
    if ((rc = spec(c, modes, f[d], cell_vol, d, d))) return rc;
    if ((rc = inverse3(c, f[d]))) return rc;
  }
  if ((rc = source(c, f, 0))) return rc;
  const int off[3][2] = {{0, 1}, {0, 2}, {1, 2}};
  for (const auto& ij : off) {
    if ((rc = spec(c, modes, f[0], cell_vol, ij[0], ij[1]))) return rc;
    if ((rc = inverse3(c, f[0]))) return rc;
    if ((rc = source(c, f, 1))) return rc;
  }
  // pot_2 = laplace(rfftn(L)) and its -i k gradient, as the forward Poisson
  // solve of a density (lpt.py:171-176)
  float2* g[3] = {c->lbuf, c->cbuf[1], c->cbuf[2]};
  if ((rc = pmx_fft_r2c_inplace(c, (float*)g[0], 0, 0.f))) return rc;
  if ((rc = pmx_fft_strided(c, g, 1, 1, -1))) return rc;
  if ((rc = neg_grad3(c, g, 0))) return rc;
  return move(c, disp, vel, g, fac[2], fac[3]);
}

// ------------------------------------------------------- linear modes --
// modes.py:67-117 on an r2c spectrum in place: m *= scale, then m *=
// sqrt(V P(|k|)) in float32, P evaluated in float64 as boltzmann.py:399-455
// with the transfer table interpolated linearly (numpy.interp, clamped):
// P = amp (k / k_pivot)^n_s (q T(k))^2 d2, amp = 0.32 A_s k_pivot,
// q = pi (c/H_0)^2 / Omega_m, d2 = D(a)^2 or 1; |k| is formed in float32 from
// the float32 wavevectors.
struct PowerPar {
  const double* tk;   // transfer table k (ascending) and T on the device
  const double* tT;
  int ntab;
  double amp, k_pivot, n_s, q, vol, d2;
};

__device__ __forceinline__ double interp_table(double x, const double* __restrict__ xp,
                                               const double* __restrict__ fp, int n) {
  if (x <= xp[0]) return fp[0];
  if (x >= xp[n - 1]) return fp[n - 1];
  // first index with xp[i] > x (searchsorted right), at least 1
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (xp[mid] > x) hi = mid;
    else lo = mid + 1;
  }
  const int i = lo < 1 ? 1 : (lo > n - 1 ? n - 1 : lo);
  const double x0 = xp[i - 1], x1 = xp[i], f0 = fp[i - 1], f1 = fp[i];
  const double t = (x - x0) / (x1 - x0);
  return f0 + t * (f1 - f0);
}

__global__ void __launch_bounds__(PMX_TPB)
    linear_modes_k(float2* __restrict__ modes, int n1, int n2h, long long total, float scale,
                   LptK kt, PowerPar pp) {
  const long long t = (long long)blockIdx.x * PMX_TPB + threadIdx.x;
  if (t >= total) return;
  const long long row = t / n2h;
  const int kz = (int)(t - row * n2h);
  const int y = (int)(row % n1);
  const int x = (int)(row / n1);
  const float kx = kt.k[0][x], ky = kt.k[1][y], kk = kt.k[2][kz];
  float k2 = kx * kx;
  k2 = k2 + ky * ky;
  k2 = k2 + kk * kk;
  const double k = (double)sqrtf(k2);
  const double T = interp_table(k, pp.tk, pp.tT, pp.ntab);
  const double qT = pp.q * T;
  double P = pp.amp * pow(k / pp.k_pivot, pp.n_s) * (qT * qT);
  if (pp.d2 != 1.0) P = P * pp.d2;
  const float fac = (float)sqrt(P * pp.vol);
  float2 m = modes[t];
  m.x = m.x * scale;
  m.y = m.y * scale;
  modes[t] = make_float2(m.x * fac, m.y * fac);
}

extern "C" int pmx_linear_modes(pmx_ctx* c, float* modes, float scale, const double* tk,
                                const double* tT, int ntab, const double* par) {
  PMX_ENTRY(c);
  PMX_REQUIRE(c && modes && tk && tT && par && ntab >= 2, "pmx_linear_modes: bad argument");
  // the transfer table goes to the context's table scratch, then the stream
  // waits for nothing else (the copy is from pageable host memory)
  const size_t tb = (size_t)2 * ntab * sizeof(double);
  if (c->tab_cap < tb) {
    if (c->tab) pmx_dev_free(c, (void**)&c->tab, c->tab_cap);
    c->tab_cap = 0;
    if (int rc = pmx_dev_alloc(c, (void**)&c->tab, tb)) return rc;
    c->tab_cap = tb;
  }
  PMX_CHECK_HIP(hipMemcpyAsync(c->tab, tk, ntab * sizeof(double), hipMemcpyHostToDevice,
                               c->stream));
  PMX_CHECK_HIP(hipMemcpyAsync(c->tab + ntab, tT, ntab * sizeof(double), hipMemcpyHostToDevice,
                               c->stream));
  PowerPar pp = {c->tab, c->tab + ntab, ntab, par[0], par[1], par[2], par[3], par[4], par[5]};
  const long long total = (long long)c->n[0] * c->n[1] * c->n2h;
  pmx_tic(c, PMX_T_OTHER);
  linear_modes_k<<<pmx_blocks(total), PMX_TPB, 0, c->stream>>>((float2*)modes, c->n[1], c->n2h,
                                                               total, scale, lpt_k(c), pp);
  pmx_toc(c, PMX_T_OTHER);
  PMX_LAUNCH_CHECK();
  return PMX_OK;
}
