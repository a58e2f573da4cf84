// Hand-written 3D real FFTs and the fused Poisson/gradient solve (gfx950).
//
// Layout: a real mesh lives in "padded" rows of rs = 2*(n2/2+1) floats so the
// r2c/c2r row passes run in place; the complex mesh is (n0, n1, n2h) float2 in rows of n2p >= n2h.
// Each kernel stages P pencils of length n in LDS (pencil-major for the
// contiguous row passes, interleaved for the strided passes) and runs a
// Stockham autosort FFT with radix 8/4/2 stages, or a direct DFT for lengths
// that are not powers of two.
//
// Pass order (forward): z r2c (rows, contiguous) -> y (strided) -> x
// (strided).  The x pass of the Poisson solve does forward x FFT, Green's
// function, -ik_d and the inverse x FFT for the three components, so the
// inverse ends with y (strided) and z c2r (rows).
#include "pmx_internal.h"

#define FFT_TPB 512                  // threads per FFT workgroup
#define FFT_CAP 8192                 // complex values per workgroup
#define FFT_MAXV (FFT_CAP / FFT_TPB) // values per thread
#define RFFT_CAP 4096                // row passes: fewer values per thread
#define RFFT_MAXV (RFFT_CAP / FFT_TPB)
// Poisson x pass: 8 columns of 512 per 512-thread workgroup, three such
// workgroups (40 KB of LDS each) per CU; 16 columns per 1024 or per 512
// threads and 8 per 256 measured slower, two waves per SIMD no faster
#ifndef PFFT_TPB
#define PFFT_TPB 512
#endif
#ifndef PFFT_CAP
#define PFFT_CAP 4096
#endif
#ifndef PFFT_WPE                     // waves per SIMD the register budget allows
#define PFFT_WPE 1
#endif
#define PFFT_MAXV (PFFT_CAP / PFFT_TPB)

// Complex products inside the transforms may fuse into FMAs (the file is
// otherwise built with separate roundings).
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
#pragma clang fp contract(fast)
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a w (DIR < 0) or a conj(w) (DIR > 0)
template <int DIR>
__device__ __forceinline__ float2 cmul_dir(float2 a, float2 w) {
#pragma clang fp contract(fast)
  return DIR < 0 ? make_float2(a.x * w.x - a.y * w.y, a.x * w.y + a.y * w.x)
                 : make_float2(a.x * w.x + a.y * w.y, a.y * w.x - a.x * w.y);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }

// multiply by -i (forward, DIR<0) or +i (inverse)
template <int DIR>
__device__ __forceinline__ float2 mul_mi(float2 a) {
  return DIR < 0 ? make_float2(a.y, -a.x) : make_float2(-a.y, a.x);
}

template <int DIR>
__device__ __forceinline__ float2 twiddle(const float2* __restrict__ tw, int m) {
  float2 w = tw[m];
  if (DIR > 0) w.y = -w.y;
  return w;
}

template <int DIR>
__device__ __forceinline__ void dft2(float2* v) {
  float2 a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}

template <int DIR>
__device__ __forceinline__ void dft4(float2* v) {
  float2 t0 = cadd(v[0], v[2]);
  float2 t1 = csub(v[0], v[2]);
  float2 t2 = cadd(v[1], v[3]);
  float2 t3 = mul_mi<DIR>(csub(v[1], v[3]));
  v[0] = cadd(t0, t2);
  v[2] = csub(t0, t2);
  v[1] = cadd(t1, t3);
  v[3] = csub(t1, t3);
}

template <int DIR>
__device__ __forceinline__ void dft8(float2* v) {
  const float h = 0.70710678118654752440f;
  float2 e[4] = {v[0], v[2], v[4], v[6]};
  float2 o[4] = {v[1], v[3], v[5], v[7]};
  dft4<DIR>(e);
  dft4<DIR>(o);
  // w8^q for q = 0..3 with w8 = exp(DIR * i pi/4)
  float2 o1 = DIR < 0 ? make_float2(h * (o[1].x + o[1].y), h * (o[1].y - o[1].x))
                      : make_float2(h * (o[1].x - o[1].y), h * (o[1].y + o[1].x));
  float2 o2 = mul_mi<DIR>(o[2]);
  float2 o3 = DIR < 0 ? make_float2(h * (o[3].y - o[3].x), -h * (o[3].x + o[3].y))
                      : make_float2(-h * (o[3].x + o[3].y), h * (o[3].x - o[3].y));
  v[0] = cadd(e[0], o[0]);
  v[4] = csub(e[0], o[0]);
  v[1] = cadd(e[1], o1);
  v[5] = csub(e[1], o1);
  v[2] = cadd(e[2], o2);
  v[6] = csub(e[2], o2);
  v[3] = cadd(e[3], o3);
  v[7] = csub(e[3], o3);
}

template <int R, int DIR>
__device__ __forceinline__ void dftR(float2* v) {
  if (R == 2) dft2<DIR>(v);
  else if (R == 4) dft4<DIR>(v);
  else dft8<DIR>(v);
}

// LDS layouts of P pencils of length n:
//   ROWS (pencil-major):  element (p, j) at p*(n + n/8) + j + j/8; the gap
//                         after every 8 elements spreads the stride-R stage
//                         writes over the banks
//   COLS (interleaved):   element (p, j) at j*P + p
// Work items are mapped so that consecutive lanes touch consecutive words.
enum { ROWS = 0, COLS = 1 };

template <int LAY>
__device__ __forceinline__ int lidx(int p, int j, int n, int P) {
  return LAY == ROWS ? p * (n + (n >> 3)) + j + (j >> 3) : j * P + p;
}

// LDS elements needed for P pencils of length n in a layout
__host__ __device__ constexpr int lds_elems(int lay, int n, int P) {
  return lay == ROWS ? P * (n + (n >> 3)) : P * n;
}

__host__ __device__ constexpr int ilog2c(int n) { return n <= 1 ? 0 : 1 + ilog2c(n >> 1); }

// pencils per workgroup for length n: a power of two with P*n <= cap
__host__ __device__ constexpr int fft_pencils(int n, int cap = FFT_CAP) {
  int P = 1;
  while (P * 2 * n <= cap && P * 2 <= 256) P *= 2;
  return P;
}

// rows per workgroup for the row passes: P*n <= RFFT_CAP and the swizzled
// ROWS block within FFT_CAP elements (64 KB)
__host__ __device__ constexpr int row_pencils(int n) {
  int P = 1;
  while (P * 2 * n <= RFFT_CAP && lds_elems(ROWS, n, P * 2) <= FFT_CAP && P * 2 <= 256) P *= 2;
  return P;
}

// One Stockham stage of radix R (n a power of two, Ns the size of the
// sub-transforms already done): reads x[j + r n/R], twiddles by
// w_{Ns R}^{(j mod Ns) r}, and writes the radix-R DFT to
// (j - j mod Ns) R + j mod Ns + r Ns.  Twiddle powers are built from at most
// three table reads (w, w^2, w^4) with at most two extra products; all reads
// of a stage are issued before the arithmetic.
template <int R, int DIR, int LAY, int MAXV, int TPB>
__device__ __forceinline__ void fft_stage(float2* s, int n, int P, int Ns, int lnb, int lP,
                                          const float2* __restrict__ tw, int tws) {
  constexpr int NG = MAXV / R;
  const int nb = n / R;
  const int total = P * nb;
  const int step = n / (Ns * R);
  float2 v[MAXV];
  float2 w1[NG], w2[NG], w4[NG];
  int mm[NG];
#pragma unroll
  for (int it = 0; it < NG; ++it) {
    int idx = threadIdx.x + it * TPB;
    mm[it] = 0;
    if (idx < total) {
      int p = LAY == ROWS ? (idx >> lnb) : (idx & (P - 1));
      int j = LAY == ROWS ? (idx & (nb - 1)) : (idx >> lP);
      int m = (j & (Ns - 1)) * step * tws;
      mm[it] = m;
      w1[it] = twiddle<DIR>(tw, m);
      if (R >= 4) w2[it] = twiddle<DIR>(tw, 2 * m);
      if (R == 8) w4[it] = twiddle<DIR>(tw, 4 * m);
#pragma unroll
      for (int r = 0; r < R; ++r) v[it * R + r] = s[lidx<LAY>(p, j + r * nb, n, P)];
    }
  }
#pragma unroll
  for (int it = 0; it < NG; ++it) {
    int idx = threadIdx.x + it * TPB;
    if (idx < total) {
      if (mm[it] != 0) {
        float2* u = &v[it * R];
        u[1] = cmul(u[1], w1[it]);
        if (R >= 4) {
          float2 w3 = cmul(w2[it], w1[it]);
          u[2] = cmul(u[2], w2[it]);
          u[3] = cmul(u[3], w3);
          if (R == 8) {
            u[4] = cmul(u[4], w4[it]);
            u[5] = cmul(u[5], cmul(w4[it], w1[it]));
            u[6] = cmul(u[6], cmul(w4[it], w2[it]));
            u[7] = cmul(u[7], cmul(w4[it], w3));
          }
        }
      }
      dftR<R, DIR>(&v[it * R]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int it = 0; it < NG; ++it) {
    int idx = threadIdx.x + it * TPB;
    if (idx < total) {
      int p = LAY == ROWS ? (idx >> lnb) : (idx & (P - 1));
      int j = LAY == ROWS ? (idx & (nb - 1)) : (idx >> lP);
      int k = j & (Ns - 1);
      int o = (j - k) * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) s[lidx<LAY>(p, o + r * Ns, n, P)] = v[it * R + r];
    }
  }
  __syncthreads();
}

// Direct DFT of P pencils of any length n (P*n <= FFT_CAP).
template <int DIR, int LAY, int MAXV, int TPB>
__device__ __forceinline__ void fft_direct(float2* s, int n, int P, const float2* __restrict__ tw,
                                           int tws) {
  float2 v[MAXV];
  const int total = P * n;
#pragma unroll
  for (int it = 0; it < MAXV; ++it) {
    int idx = threadIdx.x + it * TPB;
    if (idx < total) {
      int p = LAY == ROWS ? idx / n : idx % P;
      int k = LAY == ROWS ? idx % n : idx / P;
      float2 acc = make_float2(0.f, 0.f);
      int m = 0;
      for (int j = 0; j < n; ++j) {
        float2 w = twiddle<DIR>(tw, m * tws);
        float2 x = s[lidx<LAY>(p, j, n, P)];
        acc.x += x.x * w.x - x.y * w.y;
        acc.y += x.x * w.y + x.y * w.x;
        m += k;
        if (m >= n) m -= n;
      }
      v[it] = acc;
    }
  }
  __syncthreads();
#pragma unroll
  for (int it = 0; it < MAXV; ++it) {
    int idx = threadIdx.x + it * TPB;
    if (idx < total) {
      int p = LAY == ROWS ? idx / n : idx % P;
      int k = LAY == ROWS ? idx % n : idx / P;
      s[lidx<LAY>(p, k, n, P)] = v[it];
    }
  }
  __syncthreads();
}

// Unnormalised in-LDS FFT of P pencils (exp(DIR 2 pi i jk/n)).  POW2
// selects the Stockham path (n and P powers of two) or the direct DFT.  tw
// is a table of exp(-2 pi i m/(n tws)); tws > 1 reuses a longer table.
// Callers synchronise before it; it ends with a barrier.
template <int DIR, int LAY, bool POW2, int MAXV = FFT_MAXV, int TPB = FFT_TPB>
__device__ __forceinline__ void lds_fft(float2* s, int n, int P, const float2* __restrict__ tw,
                                        int tws = 1) {
  if (n == 1) return;
  if (!POW2) {
    fft_direct<DIR, LAY, MAXV, TPB>(s, n, P, tw, tws);
    return;
  }
  const int lP = __ffs(P) - 1;
  int Ns = 1;
  while (Ns < n) {
    int rem = n / Ns;
    if (rem % 8 == 0) {
      fft_stage<8, DIR, LAY, MAXV, TPB>(s, n, P, Ns, __ffs(n / 8) - 1, lP, tw, tws);
      Ns *= 8;
    } else if (rem % 4 == 0) {
      fft_stage<4, DIR, LAY, MAXV, TPB>(s, n, P, Ns, __ffs(n / 4) - 1, lP, tw, tws);
      Ns *= 4;
    } else {
      fft_stage<2, DIR, LAY, MAXV, TPB>(s, n, P, Ns, __ffs(n / 2) - 1, lP, tw, tws);
      Ns *= 2;
    }
  }
}

// ----------------------------------------- compile-time length FFTs --
// The same Stockham stages with N, P and the stage size NS known at compile
// time (N and P powers of two): index arithmetic folds to shifts and
// immediate LDS offsets, and the bounds tests go when the butterflies fill
// the workgroup.  COLS blocks read every twiddle power from the table (the
// pencils of a wave share a twiddle, so the reads broadcast); ROWS blocks
// build w^3, w^5..w^7 from w, w^2, w^4 as lds_fft does.  In both layouts
// element x of pencil p sits at lidx(p, x); the R inputs of a butterfly are
// a fixed stride apart, and so are its R outputs except for ROWS stages
// with 1 < NS < 8, which compute each address.
template <int R, int DIR, int LAY, int N, int P, int NS, int TPB, int TWS>
__device__ __forceinline__ void ct_stage(float2* s, const float2* __restrict__ tw) {
  constexpr int NB = N / R, TOTAL = P * NB;
  constexpr int NG = (TOTAL + TPB - 1) / TPB;
  constexpr bool FULL = TOTAL % TPB == 0;
  constexpr int LP = ilog2c(P), LNB = ilog2c(NB);
  constexpr int STEP = N / (NS * R) * TWS;
  constexpr bool SEP_IN = LAY == COLS || NB % 8 == 0;
  constexpr bool SEP_OUT = LAY == COLS || NS % 8 == 0 || NS == 1;
  constexpr int SIN = LAY == ROWS ? NB + (NB >> 3) : NB * P;
  constexpr int SOUT = LAY == ROWS ? NS + (NS >> 3) : NS * P;
  constexpr bool ALLTW = LAY == COLS;
  float2 v[NG * R];
  float2 w[NG][R > 1 ? R - 1 : 1];
  int ob[NG];
#pragma unroll
  for (int it = 0; it < NG; ++it) {
    const int idx = threadIdx.x + it * TPB;
    if (FULL || idx < TOTAL) {
      const int p = LAY == ROWS ? idx >> LNB : idx & (P - 1);
      const int j = LAY == ROWS ? idx & (NB - 1) : idx >> LP;
      const int k = j & (NS - 1);
      ob[it] = lidx<LAY>(p, (j - k) * R + k, N, P);
      const int ib = lidx<LAY>(p, j, N, P);
#pragma unroll
      for (int r = 0; r < R; ++r)
        v[it * R + r] = SEP_IN ? s[ib + r * SIN] : s[lidx<LAY>(p, j + r * NB, N, P)];
      if (NS > 1) {
        const int m = k * STEP;
        if (ALLTW) {
#pragma unroll
          for (int r = 1; r < R; ++r) w[it][r - 1] = tw[r * m];
        } else {
          w[it][0] = tw[m];
          if constexpr (R >= 4) w[it][1] = tw[2 * m];
          if constexpr (R == 8) w[it][3] = tw[4 * m];
        }
      }
    }
  }
#pragma unroll
  for (int it = 0; it < NG; ++it) {
    const int idx = threadIdx.x + it * TPB;
    if (FULL || idx < TOTAL) {
      float2* u = &v[it * R];
      if (NS > 1) {
        if constexpr (!ALLTW) {
          if constexpr (R >= 4) w[it][2] = cmul(w[it][1], w[it][0]);
          if constexpr (R == 8) {
            w[it][4] = cmul(w[it][3], w[it][0]);
            w[it][5] = cmul(w[it][3], w[it][1]);
            w[it][6] = cmul(w[it][3], w[it][2]);
          }
        }
#pragma unroll
        for (int r = 1; r < R; ++r) u[r] = cmul_dir<DIR>(u[r], w[it][r - 1]);
      }
      dftR<R, DIR>(u);
    }
  }
  __syncthreads();
#pragma unroll
  for (int it = 0; it < NG; ++it) {
    const int idx = threadIdx.x + it * TPB;
    if (FULL || idx < TOTAL) {
      if (SEP_OUT) {
#pragma unroll
        for (int r = 0; r < R; ++r) s[ob[it] + r * SOUT] = v[it * R + r];
      } else {
        const int p = idx >> LNB, j = idx & (NB - 1);
        const int k = j & (NS - 1), o = (j - k) * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) s[lidx<LAY>(p, o + r * NS, N, P)] = v[it * R + r];
      }
    }
  }
  __syncthreads();
}

template <int DIR, int LAY, int N, int P, int NS, int TPB, int TWS>
__device__ __forceinline__ void ct_fft(float2* s, const float2* __restrict__ tw) {
  if constexpr (NS < N) {
    constexpr int REM = N / NS;
    constexpr int R = REM % 8 == 0 ? 8 : (REM % 4 == 0 ? 4 : 2);
    ct_stage<R, DIR, LAY, N, P, NS, TPB, TWS>(s, tw);
    ct_fft<DIR, LAY, N, P, NS * R, TPB, TWS>(s, tw);
  }
}

// FFT of P pencils of length n: the compile-time path when LN > 0
// (n = 2^LN, P = 2^LP), else lds_fft.
template <int DIR, int LAY, bool POW2, int LN, int LP, int MAXV, int TPB, int TWS = 1>
__device__ __forceinline__ void fft_run(float2* s, int n, int P, const float2* __restrict__ tw) {
  if constexpr (LN > 0)
    ct_fft<DIR, LAY, 1 << LN, 1 << LP, 1, TPB, TWS>(s, tw);
  else
    lds_fft<DIR, LAY, POW2, MAXV, TPB>(s, n, P, tw, TWS);
}

static bool is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

// -------------------------------------------------- batched transfers --
// Global <-> LDS traffic goes through registers in unrolled batches of up to
// MAXV values per thread, so a thread's loads are all in flight at once
// instead of waiting out the memory latency one element at a time.
template <int MAXV, int TPB, class LD, class ST>
__device__ __forceinline__ void batch_copy(int total, LD ld, ST st) {
  float2 r[MAXV];
#pragma unroll
  for (int it = 0; it < MAXV; ++it) {
    int idx = threadIdx.x + it * TPB;
    if (idx < total) r[it] = ld(idx);
  }
#pragma unroll
  for (int it = 0; it < MAXV; ++it) {
    int idx = threadIdx.x + it * TPB;
    if (idx < total) st(idx, r[it]);
  }
}

// Copies the twiddle table tw[0, n) to LDS (no barrier).
template <int TPB>
__device__ __forceinline__ void stage_twiddles(float2* stw, const float2* __restrict__ tw, int n) {
  for (int base = 0; base < n; base += 4 * TPB)
    batch_copy<4, TPB>(n - base, [&](int i) { return tw[base + i]; },
                       [&](int i, float2 v) { stw[base + i] = v; });
}

// Logical block of a pass: the XCD-contiguous remap of the hardware block
// index, in reverse when rev is set.  Consecutive passes of a transform run
// in opposite directions, so that each first reads what the pass before it
// wrote last, while that is still in the last-level cache.
__device__ __forceinline__ unsigned pass_block(unsigned lin, unsigned nb, int rev) {
  const unsigned l = pmx_xcd_block(lin, nb);
  return rev ? nb - 1 - l : l;
}

// ------------------------------------------------------------ z passes --
// In-place r2c of padded rows as a length-n2 complex FFT of the real row
// (any n2); PRE applies (x - 1) * coef on load (gravity.py:52-54).
// LDS: the row block, then the n2-entry twiddle table.
template <int PRE, bool POW2>
__global__ void __launch_bounds__(FFT_TPB, 4) z_r2c_k(float* __restrict__ data, long long nrows, int n2,
                                                      int rs, int P, float coef,
                                                      const float2* __restrict__ tw,
                                                      const float* __restrict__ fcoef) {
  extern __shared__ float2 s[];
  if (PRE && fcoef) coef = fcoef[0];
  float2* stw = s + lds_elems(ROWS, n2, P);
  long long row0 = (long long)blockIdx.x * P;
  int np = (int)((nrows - row0) < P ? (nrows - row0) : P);
  stage_twiddles<FFT_TPB>(stw, tw, n2);
  batch_copy<RFFT_MAXV, FFT_TPB>(
      np * n2,
      [&](int idx) {
        int p = idx / n2, j = idx - (idx / n2) * n2;
        float x = data[(row0 + p) * rs + j];
        if (PRE) {
          x = x - 1.f;
          x = x * coef;
        }
        return make_float2(x, 0.f);
      },
      [&](int idx, float2 v) {
        int p = idx / n2, j = idx - (idx / n2) * n2;
        s[lidx<ROWS>(p, j, n2, P)] = v;
      });
  __syncthreads();
  lds_fft<-1, ROWS, POW2, RFFT_MAXV>(s, n2, np, stw);
  const int n2h = n2 / 2 + 1;
  float2* out = reinterpret_cast<float2*>(data);
  const int rsc = rs / 2;
  batch_copy<RFFT_MAXV, FFT_TPB>(
      np * n2h,
      [&](int idx) {
        int p = idx / n2h, k = idx - (idx / n2h) * n2h;
        return s[lidx<ROWS>(p, k, n2, P)];
      },
      [&](int idx, float2 v) {
        int p = idx / n2h, k = idx - (idx / n2h) * n2h;
        out[(row0 + p) * rsc + k] = v;
      });
}

// In-place c2r of padded rows with a final scale (any n2).  The full
// spectrum is rebuilt from the half (X[n-k] = conj X[k]) and the real part
// of the inverse is kept, which discards Im X[0] and Im X[n/2] as irfft does.
template <bool POW2>
__global__ void __launch_bounds__(FFT_TPB, 4) z_c2r_k(float* __restrict__ data, long long nrows, int n2,
                                                      int rs, int P, float scale,
                                                      const float2* __restrict__ tw) {
  extern __shared__ float2 s[];
  float2* stw = s + lds_elems(ROWS, n2, P);
  long long row0 = (long long)blockIdx.x * P;
  int np = (int)((nrows - row0) < P ? (nrows - row0) : P);
  const float2* in = reinterpret_cast<const float2*>(data);
  const int rsc = rs / 2;
  const int half = n2 / 2;
  stage_twiddles<FFT_TPB>(stw, tw, n2);
  batch_copy<RFFT_MAXV, FFT_TPB>(
      np * n2,
      [&](int idx) {
        int p = idx / n2, k = idx - (idx / n2) * n2;
        int kk = k <= half ? k : n2 - k;
        float2 v = in[(row0 + p) * rsc + kk];
        if (k > half) v.y = -v.y;
        return v;
      },
      [&](int idx, float2 v) {
        int p = idx / n2, k = idx - (idx / n2) * n2;
        s[lidx<ROWS>(p, k, n2, P)] = v;
      });
  __syncthreads();
  lds_fft<1, ROWS, POW2, RFFT_MAXV>(s, n2, np, stw);
  batch_copy<RFFT_MAXV, FFT_TPB>(
      np * n2,
      [&](int idx) {
        int p = idx / n2, j = idx - (idx / n2) * n2;
        return s[lidx<ROWS>(p, j, n2, P)];
      },
      [&](int idx, float2 v) {
        int p = idx / n2, j = idx - (idx / n2) * n2;
        data[(row0 + p) * rs + j] = v.x * scale;
      });
}

// Half-length r2c for even n2 with H = n2/2 a power of two: the row is read
// as H complex values z[m] = x[2m] + i x[2m+1], Z = FFT_H(z), and for
// k = 0..H:  X[k] = E + w^k O,  E = (Z[k] + conj Z[H-k])/2,
// O = -i (Z[k] - conj Z[H-k])/2,  w = exp(-2 pi i/n2)  (Z[H] = Z[0]).
// PRE applies (x - 1) * coef on load.  LH > 0 fixes H = 2^LH and
// P = row_pencils(H) at compile time (the launch then holds whole blocks).
__device__ __forceinline__ float2 r2c_half_x(const float2* s, const float2* stw, int p, int k, int H,
                                             int P) {
  float2 zk = s[lidx<ROWS>(p, k & (H - 1), H, P)];
  float2 zm = s[lidx<ROWS>(p, (H - k) & (H - 1), H, P)];
  zm.y = -zm.y;
  float2 e = make_float2(0.5f * (zk.x + zm.x), 0.5f * (zk.y + zm.y));
  float2 dd = csub(zk, zm);
  float2 o = make_float2(0.5f * dd.y, -0.5f * dd.x);
  return cadd(e, cmul(stw[k], o));
}

// The deposit's halo shells of one channel for a row pass that folds them
// in on load (16^3 tiles; n1 the y extent of the mesh).
struct FoldIn {
  const float* halo;
  int shell, ntx, nty, ntz, n1;
};

// Row (x, y), z cells 2m and 2m+1: each cell on a low face of its tile gets
// the fold's shell sum added, as fold_k adds it to the mesh.
__device__ __forceinline__ float2 fold_pair(float2 v, const FoldIn& fi, int x, int y, int m) {
  const int lx = x & 15, ly = y & 15, z = 2 * m, lz = z & 15;
  if (lx == 0 || ly == 0 || lz == 0)
    v.x = v.x + pmx_fold_sum16(fi.halo, fi.shell, fi.ntx, fi.nty, fi.ntz, x >> 4, y >> 4, z >> 4,
                               lx, ly, lz);
  if (lx == 0 || ly == 0)
    v.y = v.y + pmx_fold_sum16(fi.halo, fi.shell, fi.ntx, fi.nty, fi.ntz, x >> 4, y >> 4, z >> 4,
                               lx, ly, lz + 1);
  return v;
}

// The fold terms of a block of P rows (x plane x, rows y = y0 + p, pairs
// z = 2m, 2m + 1; 2H cells a row), staged in LDS.  A cell takes
// pmx_fold_sum16's terms; by the faces its row lies on (lx = x % 16,
// ly = y % 16, lz = z % 16):
//   lx != 0, ly != 0: S = 4 on cells with lz = 0;
//   lx == 0, ly != 0 (a face row): S = 1 on every cell, then S = 4, 5 when lz = 0;
//   lx != 0, ly == 0 (a face row): S = 2 on every cell, then S = 4, 6 when lz = 0;
//   lx == 0, ly == 0 (a corner row, 1 in 256): fold_pair.
// T[p * 2H + z] holds the face rows' S = 1 or S = 2 term, loaded straight into
// LDS (the rows' FFT buffer, before the rows go there), and Z0[p + P tz] the
// S = 4 term and Z1[p + P tz] the S = 5 or S = 6 term of cell (p, 16 tz),
// one each per thread (P * H / 8 <= TPB of them).  All of it is issued before
// the row loads, so no wait falls between two of them.
template <int P, int TPB>
__device__ __forceinline__ void fold_tables(const FoldIn& fi, int x, int y0, int lH, int H, float* T,
                                            float& z0, float& z1) {
  typedef __attribute__((address_space(1))) const void gptr;
  typedef __attribute__((address_space(3))) void lptr;
  constexpr int lP = ilog2c(P);
  const int lx = x & 15, tx = x >> 4, txm = tx == 0 ? fi.ntx - 1 : tx - 1;
  auto tile = [&](int nx, int ny, int nz) {
    return fi.halo + (long long)((nx * fi.nty + ny) * fi.ntz + nz) * fi.shell;
  };
  const int j = (int)threadIdx.x;
  z0 = z1 = 0.f;
  if (j < P * (H >> 3)) {
    const int y = y0 + (j & (P - 1)), tz = j >> lP, ly = y & 15, ty = y >> 4;
    const int tym = ty == 0 ? fi.nty - 1 : ty - 1, tzm = tz == 0 ? fi.ntz - 1 : tz - 1;
    if (lx != 0 || ly != 0) {
      z0 = tile(tx, ty, tzm)[561 + lx * 16 + ly];
      if (lx == 0)
        z1 = tile(txm, ty, tzm)[ly * 17 + 16];
      else if (ly == 0)
        z1 = tile(tx, tym, tzm)[289 + lx * 17 + 16];
    }
  }
  // a wave's 64 entries lie in one row (2H >= 128), so the face-row test is
  // wave-uniform; lane l of a wave writes the LDS word at its base plus l
  const int n2 = 2 * H, wave0 = j & ~63;
  for (int e0 = 0; e0 < P * n2; e0 += TPB) {
    const int e = e0 + j;
    if (e >= P * n2) break;
    const int p = e >> (lH + 1), z = e & (n2 - 1);
    const int y = y0 + p, ly = y & 15, ty = y >> 4;
    if ((lx == 0) == (ly == 0)) continue;
    const float* src = lx == 0 ? tile(txm, ty, z >> 4) + ly * 17 + (z & 15)
                               : tile(tx, ty == 0 ? fi.nty - 1 : ty - 1, z >> 4) + 289 + lx * 17 + (z & 15);
    __builtin_amdgcn_global_load_lds((gptr*)src, (lptr*)(T + e0 + wave0), 4, 0, 0);
  }
}

// Adds the fold (fold_tables' terms, in pmx_fold_sum16's order) to a block's
// NV values per thread (row p = idx >> lH, pair m = idx & (H - 1)).
template <int NV, int TPB, int P>
__device__ __forceinline__ void fold_rows(float2 (&r)[NV], const FoldIn& fi, int x, int y0, int lH,
                                          int H, int total, const float* T, const float* Z0,
                                          const float* Z1) {
  const int lx = x & 15;
#pragma unroll
  for (int it = 0; it < NV; ++it) {
    const int idx = threadIdx.x + it * TPB;
    if (idx >= total) continue;
    const int p = idx >> lH, y = y0 + p, m = idx & (H - 1);
    const int ly = y & 15, z = 2 * m, lz = z & 15, k = (m >> 3) * P + p;
    if (lx == 0 && ly == 0) {
      r[it] = fold_pair(r[it], fi, x, y, m);
    } else if (lx == 0 || ly == 0) {
      const float* t = T + (p << (lH + 1)) + z;
      float sx = 0.f;
      sx += t[0];
      if (lz == 0) {
        sx += Z0[k];
        sx += Z1[k];
      }
      r[it].x = r[it].x + sx;
      float sy = 0.f;
      sy += t[1];
      r[it].y = r[it].y + sy;
    } else if (lz == 0) {
      float sx = 0.f;
      sx += Z0[k];
      r[it].x = r[it].x + sx;
    }
  }
}

// FOLD (with LH > 0, 16^3 tiles, P | n1): the density is the deposit's tile
// interiors plus the fold of fi.halo, added here on load.
template <int PRE, int LH, bool FOLD = false>
__global__ void __launch_bounds__(FFT_TPB, 4) z_r2c_half_k(float* __restrict__ data, long long nrows,
                                                           int n2, int rs, int P_, float coef,
                                                           const float2* __restrict__ tw,
                                                           FoldIn fi, const float* __restrict__ fcoef,
                                                           int rev) {
  extern __shared__ float2 s[];
  // 1.5 Omega_m from the device factor table when the call has one
  if (PRE && fcoef) coef = fcoef[0];
  constexpr int PC = LH ? row_pencils(1 << LH) : 1;
  const int H = LH ? 1 << LH : n2 >> 1;
  const int lH = LH ? LH : __ffs(H) - 1;
  const int P = LH ? PC : P_;
  float2* stw = s + lds_elems(ROWS, H, P);
  const long long row0 = (long long)pass_block(blockIdx.x, gridDim.x, rev) * P;
  const int np = LH ? PC : (int)((nrows - row0) < P ? (nrows - row0) : P);
  const int rsc = rs / 2;
  float2* blk = reinterpret_cast<float2*>(data) + row0 * rsc;
  const int total = np * H;
  stage_twiddles<FFT_TPB>(stw, tw, 2 * H);
  // the block's rows lie in one x plane when folding
  const int fx = FOLD ? (int)row0 / fi.n1 : 0;
  const int fy0 = FOLD ? (int)row0 - fx * fi.n1 : 0;
  {
    // the block's fold terms (fold_tables), issued before the row loads, in
    // the rows' buffer: T (P * 2H floats), then Z0 and Z1 (P * H / 8 each) in
    // what the rows' padding leaves (P * (H + H / 8) float2 in all)
    float* T = reinterpret_cast<float*>(s);
    float* Z0 = T + PC * 2 * H;
    float* Z1 = Z0 + PC * (H >> 3);
    float z0 = 0.f, z1 = 0.f;
    if (FOLD) fold_tables<PC, FFT_TPB>(fi, fx, fy0, lH, H, T, z0, z1);
    float2 r[RFFT_MAXV];
#pragma unroll
    for (int it = 0; it < RFFT_MAXV; ++it) {
      const int idx = threadIdx.x + it * FFT_TPB;
      if (idx < total) r[it] = blk[__umul24(idx >> lH, rsc) + (idx & (H - 1))];
    }
    if (FOLD) {
      if ((int)threadIdx.x < PC * (H >> 3)) {
        Z0[threadIdx.x] = z0;
        Z1[threadIdx.x] = z1;
      }
      __builtin_amdgcn_s_waitcnt(0);  // T's loads into LDS have landed
      __syncthreads();
      fold_rows<RFFT_MAXV, FFT_TPB, PC>(r, fi, fx, fy0, lH, H, total, T, Z0, Z1);
      __syncthreads();  // T is read before the rows take its place
    }
    if (PRE) {
#pragma unroll
      for (int it = 0; it < RFFT_MAXV; ++it) {
        r[it].x = r[it].x - 1.f;
        r[it].x = r[it].x * coef;
        r[it].y = r[it].y - 1.f;
        r[it].y = r[it].y * coef;
      }
    }
#pragma unroll
    for (int it = 0; it < RFFT_MAXV; ++it) {
      const int idx = threadIdx.x + it * FFT_TPB;
      if (idx < total) s[lidx<ROWS>(idx >> lH, idx & (H - 1), H, P)] = r[it];
    }
  }
  __syncthreads();
  fft_run<-1, ROWS, true, LH, ilog2c(PC), RFFT_MAXV, FFT_TPB, 2>(s, H, np, stw);
  {
    float2 r[RFFT_MAXV];
#pragma unroll
    for (int it = 0; it < RFFT_MAXV; ++it) {
      const int idx = threadIdx.x + it * FFT_TPB;
      if (idx < total) r[it] = r2c_half_x(s, stw, idx >> lH, idx & (H - 1), H, P);
    }
#pragma unroll
    for (int it = 0; it < RFFT_MAXV; ++it) {
      const int idx = threadIdx.x + it * FFT_TPB;
      if (idx < total) blk[__umul24(idx >> lH, rsc) + (idx & (H - 1))] = r[it];
    }
    // X[H] of each row
    if ((int)threadIdx.x < np)
      blk[__umul24(threadIdx.x, rsc) + H] = r2c_half_x(s, stw, threadIdx.x, H, H, P);
  }
}

// Half-length c2r (inverse of the above) with a final scale:
// Z[k] = (X[k] + conj X[H-k]) + i w^-k (X[k] - conj X[H-k]) for k < H with
// Im X[0] = Im X[H] = 0 (as irfft), z = IFFT_H(Z), x[2m] + i x[2m+1] = z[m].
// LH as in z_r2c_half_k.
template <int LH>
__global__ void __launch_bounds__(FFT_TPB, 4) z_c2r_half_k(float* __restrict__ data, long long nrows,
                                                           int n2, int rs, int P_, float scale,
                                                           const float2* __restrict__ tw, int rev) {
  extern __shared__ float2 s[];
  constexpr int PC = LH ? row_pencils(1 << LH) : 1;
  const int H = LH ? 1 << LH : n2 >> 1;
  const int lH = LH ? LH : __ffs(H) - 1;
  const int P = LH ? PC : P_;
  float2* stw = s + lds_elems(ROWS, H, P);
  const long long row0 = (long long)pass_block(blockIdx.x, gridDim.x, rev) * P;
  const int np = LH ? PC : (int)((nrows - row0) < P ? (nrows - row0) : P);
  const int rsc = rs / 2;
  float2* blk = reinterpret_cast<float2*>(data) + row0 * rsc;
  const int total = np * H;
  stage_twiddles<FFT_TPB>(stw, tw, 2 * H);
  // X[k] and X[H-k] per thread: both loads of the batch are issued first
  float2 xa[RFFT_MAXV], xb[RFFT_MAXV];
#pragma unroll
  for (int it = 0; it < RFFT_MAXV; ++it) {
    const int idx = threadIdx.x + it * FFT_TPB;
    if (idx < total) {
      const float2* row = blk + __umul24(idx >> lH, rsc);
      const int k = idx & (H - 1);
      xa[it] = row[k];
      xb[it] = row[H - k];
    }
  }
  __syncthreads();  // the twiddle table is read below
#pragma unroll
  for (int it = 0; it < RFFT_MAXV; ++it) {
    const int idx = threadIdx.x + it * FFT_TPB;
    if (idx < total) {
      const int p = idx >> lH, k = idx & (H - 1);
      float2 a = xa[it], b = xb[it];
      if (k == 0) {
        a.y = 0.f;
        b.y = 0.f;
      }
      b.y = -b.y;
      float2 sum = cadd(a, b);
      float2 dif = csub(a, b);
      float2 t = cmul_dir<1>(dif, stw[k]);
      s[lidx<ROWS>(p, k, H, P)] = make_float2(sum.x - t.y, sum.y + t.x);
    }
  }
  __syncthreads();
  fft_run<1, ROWS, true, LH, ilog2c(PC), RFFT_MAXV, FFT_TPB, 2>(s, H, np, stw);
  float2 r[RFFT_MAXV];
#pragma unroll
  for (int it = 0; it < RFFT_MAXV; ++it) {
    const int idx = threadIdx.x + it * FFT_TPB;
    if (idx < total) r[it] = s[lidx<ROWS>(idx >> lH, idx & (H - 1), H, P)];
  }
#pragma unroll
  for (int it = 0; it < RFFT_MAXV; ++it) {
    const int idx = threadIdx.x + it * FFT_TPB;
    if (idx < total)
      blk[__umul24(idx >> lH, rsc) + (idx & (H - 1))] = make_float2(r[it].x * scale, r[it].y * scale);
  }
}

// ------------------------------------------------------ strided passes --
// rev: the pass runs its blocks in the reverse order (pass_block)
struct Fields3 {
  float2* f[3];
  int rev;
};

// Block coordinates of a (gx, gy, gz) grid after the XCD-contiguous remap
// of the linear block index: neighbouring column blocks, which share the
// 128-byte lines of a row, then run on one XCD close together and meet in
// its L2.
__device__ __forceinline__ void xcd_block3(int& bx, int& by, int& bz, int rev) {
  const unsigned nb = gridDim.x * gridDim.y * gridDim.z;
  const unsigned lin = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
  const unsigned l = pass_block(lin, nb, rev);
  bx = l % gridDim.x;
  by = (l / gridDim.x) % gridDim.y;
  bz = l / (gridDim.x * gridDim.y);
}

// A thread's share of an n x B block of columns (B a power of two dividing
// TPB): its it-th element is (row j0 + it*js, column b), at LDS index
// threadIdx.x + it*TPB in the COLS layout.
struct ColMap {
  int b, j0, js;
};
template <int TPB>
__device__ __forceinline__ ColMap col_map(int lB) {
  ColMap m;
  m.b = threadIdx.x & ((1 << lB) - 1);
  m.j0 = threadIdx.x >> lB;
  m.js = TPB >> lB;
  return m;
}

// In-place FFT along a strided axis of a complex (outer, n, ncol) view:
// element j of pencil (o, c) is at o*ostride + j*es + c.  A workgroup takes
// B consecutive columns, so each of its loads is B contiguous float2.
// LDS: the n x B block, then the n-entry twiddle table.  LN > 0 fixes
// n = 2^LN and B = fft_pencils(n) at compile time.
template <int DIR, bool POW2, int LN>
__global__ void __launch_bounds__(FFT_TPB, 4) strided_fft_k(Fields3 fl, int n_, long long es,
                                                         long long ostride, int ncol, int B_,
                                                         const float2* __restrict__ tw) {
  extern __shared__ float2 s[];
  constexpr int BC = LN ? fft_pencils(1 << LN) : 1;
  const int n = LN ? 1 << LN : n_;
  const int B = LN ? BC : B_;
  const int lB = LN ? ilog2c(BC) : __ffs(B_) - 1;
  float2* stw = s + n * B;
  int bx, by, bz;
  xcd_block3(bx, by, bz, fl.rev);
  const int c0 = bx * B;
  const int nb = ncol - c0 < B ? ncol - c0 : B;
  const ColMap cm = col_map<FFT_TPB>(lB);
  const bool live = cm.b < nb;
  float2* g = (bz == 0 ? fl.f[0] : (bz == 1 ? fl.f[1] : fl.f[2])) + (long long)by * ostride + c0 +
              cm.b + (long long)cm.j0 * es;
  const long long gs = (long long)cm.js * es;
  stage_twiddles<FFT_TPB>(stw, tw, n);
  float2 r[FFT_MAXV];
#pragma unroll
  for (int it = 0; it < FFT_MAXV; ++it) {
    r[it] = make_float2(0.f, 0.f);
    if (live && cm.j0 + it * cm.js < n) r[it] = g[it * gs];
  }
#pragma unroll
  for (int it = 0; it < FFT_MAXV; ++it)
    if (cm.j0 + it * cm.js < n) s[threadIdx.x + it * FFT_TPB] = r[it];
  __syncthreads();
  fft_run<DIR, COLS, POW2, LN, ilog2c(BC), FFT_MAXV, FFT_TPB>(s, n, B, stw);
#pragma unroll
  for (int it = 0; it < FFT_MAXV; ++it)
    if (cm.j0 + it * cm.js < n) r[it] = s[threadIdx.x + it * FFT_TPB];
#pragma unroll
  for (int it = 0; it < FFT_MAXV; ++it)
    if (live && cm.j0 + it * cm.js < n) g[it * gs] = r[it];
}

struct KTab {
  const float* k[3];
  const float* km[3];
};

// x pass of the Poisson solve (gravity.py:56-66): forward FFT along x of
// c0, pot = -src/k^2 (k=0 -> 0), g_d = (-i km_d) pot, inverse FFT along x.
// km_1 (one value per y row) and km_2 (one per kz column) are constant along
// x, so the y and z components share one inverse transform h =
// IFFT_x(-i pot), stored to c_1 (y_inv_seq_k scales it by km_1 and km_2 for
// the two components); the x component (-i km_0 pot) is transformed on its
// own and stored to c_0.  The two inverse transforms run one after the
// other in the same LDS block, h kept in registers meanwhile.  Pencils along
// axis 0 of (n0,n1,n2h); blockIdx.y = y index, blockIdx.x = block of B kz
// columns.  LDS: the n0 x B block, the n0-entry twiddle table, then the k_0
// and km_0 tables.
template <bool POW2, int LN, bool KIN = false>
__global__ void __launch_bounds__(PFFT_TPB) __attribute__((amdgpu_waves_per_eu(PFFT_WPE))) poisson_x_k(Fields3 fl, int n0_, int n1, int n2h, int n2p,
                                                       int B_, KTab kt, const float2* __restrict__ tw) {
  extern __shared__ float2 s[];
  constexpr int BC = LN ? fft_pencils(1 << LN, PFFT_CAP) : 1;
  const int n0 = LN ? 1 << LN : n0_;
  const int B = LN ? BC : B_;
  const int lB = LN ? ilog2c(BC) : __ffs(B_) - 1;
  float2* stw = s + n0 * B;
  float* sk = reinterpret_cast<float*>(stw + n0);
  float* skm = sk + n0;
  const long long es = (long long)n1 * n2p;
  int bx, y, bz;
  xcd_block3(bx, y, bz, fl.rev);
  const int c0 = bx * B;
  const int nb = n2h - c0 < B ? n2h - c0 : B;
  const long long off = (long long)y * n2p + c0;
  stage_twiddles<PFFT_TPB>(stw, tw, n0);
  for (int i = threadIdx.x; i < n0; i += PFFT_TPB) {
    sk[i] = kt.k[0][i];
    skm[i] = kt.km[0][i];
  }
  const ColMap cm = col_map<PFFT_TPB>(lB);
  const bool live = cm.b < nb;
  float2* g0 = fl.f[0] + off + cm.b + (long long)cm.j0 * es;
  float2* g1 = fl.f[1] + off + cm.b + (long long)cm.j0 * es;
  const long long gs = (long long)cm.js * es;
  float2 r[PFFT_MAXV];
#pragma unroll
  for (int it = 0; it < PFFT_MAXV; ++it) {
    r[it] = make_float2(0.f, 0.f);
    if (live && cm.j0 + it * cm.js < n0) r[it] = g0[it * gs];
  }
#pragma unroll
  for (int it = 0; it < PFFT_MAXV; ++it)
    if (cm.j0 + it * cm.js < n0) s[threadIdx.x + it * PFFT_TPB] = r[it];
  __syncthreads();
  // KIN: the input is already transformed along x (a full 3D spectrum)
  if constexpr (!KIN) fft_run<-1, COLS, POW2, LN, ilog2c(BC), PFFT_MAXV, PFFT_TPB>(s, n0, B, stw);
  const float ky = kt.k[1][y];
  const float kz = live ? kt.k[2][c0 + cm.b] : 0.f;
  float2 h[PFFT_MAXV];
#pragma unroll
  for (int it = 0; it < PFFT_MAXV; ++it) {
    const int j = cm.j0 + it * cm.js;
    h[it] = make_float2(0.f, 0.f);
    if (live && j < n0) {
      float2 v = s[threadIdx.x + it * PFFT_TPB];
      float kx = sk[j];
      float k2 = kx * kx;
      k2 = k2 + ky * ky;
      k2 = k2 + kz * kz;
      // pot = -v / k^2, h = -i pot = (pot.y, -pot.x)
      if (k2 != 0.f) {
        const float q = 1.f / k2;
        h[it] = make_float2(-(v.y * q), v.x * q);
      }
    }
  }
  __syncthreads();
  // x component: IFFT_x(km_0 h) to c_0
#pragma unroll
  for (int it = 0; it < PFFT_MAXV; ++it) {
    const int j = cm.j0 + it * cm.js;
    if (j < n0) {
      const float kmx = skm[j];
      s[threadIdx.x + it * PFFT_TPB] = make_float2(kmx * h[it].x, kmx * h[it].y);
    }
  }
  __syncthreads();
  fft_run<1, COLS, POW2, LN, ilog2c(BC), PFFT_MAXV, PFFT_TPB>(s, n0, B, stw);
#pragma unroll
  for (int it = 0; it < PFFT_MAXV; ++it)
    if (cm.j0 + it * cm.js < n0) r[it] = s[threadIdx.x + it * PFFT_TPB];
  __syncthreads();
  // h to c_1
#pragma unroll
  for (int it = 0; it < PFFT_MAXV; ++it)
    if (cm.j0 + it * cm.js < n0) s[threadIdx.x + it * PFFT_TPB] = h[it];
#pragma unroll
  for (int it = 0; it < PFFT_MAXV; ++it)
    if (live && cm.j0 + it * cm.js < n0) g0[it * gs] = r[it];
  __syncthreads();
  fft_run<1, COLS, POW2, LN, ilog2c(BC), PFFT_MAXV, PFFT_TPB>(s, n0, B, stw);
#pragma unroll
  for (int it = 0; it < PFFT_MAXV; ++it)
    if (cm.j0 + it * cm.js < n0) r[it] = s[threadIdx.x + it * PFFT_TPB];
#pragma unroll
  for (int it = 0; it < PFFT_MAXV; ++it)
    if (live && cm.j0 + it * cm.js < n0) g1[it * gs] = r[it];
}

// Adjoint x pass: acc = sum_d (-i km_d) laplace(FFT_x c_d); inverse FFT
// along x; stored to c0.  Transpose of the three force components.  As km_1
// and km_2 are constant along x, FFT_x(km_1 c_1 + km_2 c_2) replaces the two
// transforms of c_1 and c_2; y_fwd_seq_k leaves that combination in c_1.
// The forward transforms of c_0 and c_1 run one after the other in the same
// LDS block (the first result kept in registers), then the inverse of the
// sum.  LDS as poisson_x_k.
template <bool POW2, int LN>
__global__ void __launch_bounds__(PFFT_TPB) __attribute__((amdgpu_waves_per_eu(PFFT_WPE))) poisson_x_adj_k(Fields3 fl, int n0_, int n1, int n2h, int n2p,
                                                           int B_, KTab kt, const float2* __restrict__ tw) {
  extern __shared__ float2 s[];
  constexpr int BC = LN ? fft_pencils(1 << LN, PFFT_CAP) : 1;
  const int n0 = LN ? 1 << LN : n0_;
  const int B = LN ? BC : B_;
  const int lB = LN ? ilog2c(BC) : __ffs(B_) - 1;
  float2* stw = s + n0 * B;
  float* sk = reinterpret_cast<float*>(stw + n0);
  float* skm = sk + n0;
  const long long es = (long long)n1 * n2p;
  int bx, y, bz;
  xcd_block3(bx, y, bz, fl.rev);
  const int c0 = bx * B;
  const int nb = n2h - c0 < B ? n2h - c0 : B;
  const long long off = (long long)y * n2p + c0;
  stage_twiddles<PFFT_TPB>(stw, tw, n0);
  for (int i = threadIdx.x; i < n0; i += PFFT_TPB) {
    sk[i] = kt.k[0][i];
    skm[i] = kt.km[0][i];
  }
  const ColMap cm = col_map<PFFT_TPB>(lB);
  const bool live = cm.b < nb;
  float2* g0 = fl.f[0] + off + cm.b + (long long)cm.j0 * es;
  const float2* g1 = fl.f[1] + off + cm.b + (long long)cm.j0 * es;
  const long long gs = (long long)cm.js * es;
  float2 x0[PFFT_MAXV];
#pragma unroll
  for (int it = 0; it < PFFT_MAXV; ++it) {
    x0[it] = make_float2(0.f, 0.f);
    if (live && cm.j0 + it * cm.js < n0) x0[it] = g0[it * gs];
  }
#pragma unroll
  for (int it = 0; it < PFFT_MAXV; ++it)
    if (cm.j0 + it * cm.js < n0) s[threadIdx.x + it * PFFT_TPB] = x0[it];
  __syncthreads();
  fft_run<-1, COLS, POW2, LN, ilog2c(BC), PFFT_MAXV, PFFT_TPB>(s, n0, B, stw);
  {
    float2 yz[PFFT_MAXV];
#pragma unroll
    for (int it = 0; it < PFFT_MAXV; ++it) {
      yz[it] = make_float2(0.f, 0.f);
      if (live && cm.j0 + it * cm.js < n0) yz[it] = g1[it * gs];
    }
#pragma unroll
    for (int it = 0; it < PFFT_MAXV; ++it)
      if (cm.j0 + it * cm.js < n0) x0[it] = s[threadIdx.x + it * PFFT_TPB];
    __syncthreads();
#pragma unroll
    for (int it = 0; it < PFFT_MAXV; ++it)
      if (cm.j0 + it * cm.js < n0) s[threadIdx.x + it * PFFT_TPB] = yz[it];
  }
  __syncthreads();
  fft_run<-1, COLS, POW2, LN, ilog2c(BC), PFFT_MAXV, PFFT_TPB>(s, n0, B, stw);
  const float ky = kt.k[1][y];
  const float kz = live ? kt.k[2][c0 + cm.b] : 0.f;
  float2 acc[PFFT_MAXV];
#pragma unroll
  for (int it = 0; it < PFFT_MAXV; ++it) {
    const int j = cm.j0 + it * cm.js;
    acc[it] = make_float2(0.f, 0.f);
    if (live && j < n0) {
      float kx = sk[j];
      float k2 = kx * kx;
      k2 = k2 + ky * ky;
      k2 = k2 + kz * kz;
      if (k2 != 0.f) {
        const float kmx = skm[j];
        const float2 w = s[threadIdx.x + it * PFFT_TPB];
        float sx = kmx * x0[it].x + w.x, sy = kmx * x0[it].y + w.y;
        // pot = -sum / k^2, acc = -i pot = (pot.y, -pot.x)
        const float q = 1.f / k2;
        acc[it] = make_float2(-(sy * q), sx * q);
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int it = 0; it < PFFT_MAXV; ++it)
    if (cm.j0 + it * cm.js < n0) s[threadIdx.x + it * PFFT_TPB] = acc[it];
  __syncthreads();
  fft_run<1, COLS, POW2, LN, ilog2c(BC), PFFT_MAXV, PFFT_TPB>(s, n0, B, stw);
#pragma unroll
  for (int it = 0; it < PFFT_MAXV; ++it)
    if (cm.j0 + it * cm.js < n0) acc[it] = s[threadIdx.x + it * PFFT_TPB];
#pragma unroll
  for (int it = 0; it < PFFT_MAXV; ++it)
    if (live && cm.j0 + it * cm.js < n0) g0[it * gs] = acc[it];
}

// y passes of the gradient components, paired (pencils along axis 1 of
// (n0,n1,n2h); blockIdx.y = x row, blockIdx.x = block of B = fft_pencils(n)
// kz columns, 128-byte segments at n = 512; LDS: the n x B block, then the
// twiddle table).  The two transforms run one after the other in the same
// LDS block.
//
// Inverse: from h = IFFT_x(-i pot) in c_1, the y and z force components
// IFFT_y(km_1 h) to c_1 and IFFT_y(km_2 h) to c_2 -- the products and
// transforms of a separate y pass over km_1 h and km_2 h, with h read once
// and kept in registers.
//
// Forward (adjoint path, y_fwd_seq_k): FFT_y of c_1 (kept in registers) and
// of c_2, then their combination km_1 FFT_y(c_1) + km_2 FFT_y(c_2) to c_1
// for poisson_x_adj_k.
template <bool POW2, int LN>
__global__ void __launch_bounds__(FFT_TPB, 2) y_inv_seq_k(Fields3 fl, int n_, long long es,
                                                        long long ostride, int ncol, int B_,
                                                        KTab kt, const float2* __restrict__ tw) {
  extern __shared__ float2 s[];
  constexpr int BC = LN ? fft_pencils(1 << LN) : 1;
  const int n = LN ? 1 << LN : n_;
  const int B = LN ? BC : B_;
  const int lB = LN ? ilog2c(BC) : __ffs(B_) - 1;
  float2* stw = s + n * B;
  int bx, by, bz;
  xcd_block3(bx, by, bz, fl.rev);
  const int c0 = bx * B;
  const int nb = ncol - c0 < B ? ncol - c0 : B;
  const long long off = (long long)by * ostride + c0;
  stage_twiddles<FFT_TPB>(stw, tw, n);
  const ColMap cm = col_map<FFT_TPB>(lB);
  const bool live = cm.b < nb;
  const float kmz = live ? kt.km[2][c0 + cm.b] : 0.f;
  float2* g1 = fl.f[1] + off + cm.b + (long long)cm.j0 * es;
  float2* g2 = fl.f[2] + off + cm.b + (long long)cm.j0 * es;
  const long long gs = (long long)cm.js * es;
  float2 h[FFT_MAXV], r[FFT_MAXV];
#pragma unroll
  for (int it = 0; it < FFT_MAXV; ++it) {
    h[it] = make_float2(0.f, 0.f);
    if (live && cm.j0 + it * cm.js < n) h[it] = g1[it * gs];
  }
#pragma unroll
  for (int it = 0; it < FFT_MAXV; ++it) {
    const int j = cm.j0 + it * cm.js;
    if (j < n) {
      const float kmy = kt.km[1][j];
      s[threadIdx.x + it * FFT_TPB] = make_float2(kmy * h[it].x, kmy * h[it].y);
    }
  }
  __syncthreads();
  fft_run<1, COLS, POW2, LN, ilog2c(BC), FFT_MAXV, FFT_TPB>(s, n, B, stw);
#pragma unroll
  for (int it = 0; it < FFT_MAXV; ++it)
    if (cm.j0 + it * cm.js < n) r[it] = s[threadIdx.x + it * FFT_TPB];
  __syncthreads();
#pragma unroll
  for (int it = 0; it < FFT_MAXV; ++it)
    if (cm.j0 + it * cm.js < n)
      s[threadIdx.x + it * FFT_TPB] = make_float2(kmz * h[it].x, kmz * h[it].y);
#pragma unroll
  for (int it = 0; it < FFT_MAXV; ++it)
    if (live && cm.j0 + it * cm.js < n) g1[it * gs] = r[it];
  __syncthreads();
  fft_run<1, COLS, POW2, LN, ilog2c(BC), FFT_MAXV, FFT_TPB>(s, n, B, stw);
#pragma unroll
  for (int it = 0; it < FFT_MAXV; ++it)
    if (cm.j0 + it * cm.js < n) r[it] = s[threadIdx.x + it * FFT_TPB];
#pragma unroll
  for (int it = 0; it < FFT_MAXV; ++it)
    if (live && cm.j0 + it * cm.js < n) g2[it * gs] = r[it];
}

template <bool POW2, int LN>
__global__ void __launch_bounds__(FFT_TPB, 2) y_fwd_seq_k(Fields3 fl, int n_, long long es,
                                                        long long ostride, int ncol, int B_,
                                                        KTab kt, const float2* __restrict__ tw) {
  extern __shared__ float2 s[];
  constexpr int BC = LN ? fft_pencils(1 << LN) : 1;
  const int n = LN ? 1 << LN : n_;
  const int B = LN ? BC : B_;
  const int lB = LN ? ilog2c(BC) : __ffs(B_) - 1;
  float2* stw = s + n * B;
  int bx, by, bz;
  xcd_block3(bx, by, bz, fl.rev);
  const int c0 = bx * B;
  const int nb = ncol - c0 < B ? ncol - c0 : B;
  const long long off = (long long)by * ostride + c0;
  stage_twiddles<FFT_TPB>(stw, tw, n);
  const ColMap cm = col_map<FFT_TPB>(lB);
  const bool live = cm.b < nb;
  const float kmz = live ? kt.km[2][c0 + cm.b] : 0.f;
  float2* g1 = fl.f[1] + off + cm.b + (long long)cm.j0 * es;
  const float2* g2 = fl.f[2] + off + cm.b + (long long)cm.j0 * es;
  const long long gs = (long long)cm.js * es;
  float2 a[FFT_MAXV];
#pragma unroll
  for (int it = 0; it < FFT_MAXV; ++it) {
    a[it] = make_float2(0.f, 0.f);
    if (live && cm.j0 + it * cm.js < n) a[it] = g1[it * gs];
  }
#pragma unroll
  for (int it = 0; it < FFT_MAXV; ++it)
    if (cm.j0 + it * cm.js < n) s[threadIdx.x + it * FFT_TPB] = a[it];
  __syncthreads();
  fft_run<-1, COLS, POW2, LN, ilog2c(BC), FFT_MAXV, FFT_TPB>(s, n, B, stw);
  {
    float2 c[FFT_MAXV];
#pragma unroll
    for (int it = 0; it < FFT_MAXV; ++it) {
      c[it] = make_float2(0.f, 0.f);
      if (live && cm.j0 + it * cm.js < n) c[it] = g2[it * gs];
    }
#pragma unroll
    for (int it = 0; it < FFT_MAXV; ++it)
      if (cm.j0 + it * cm.js < n) a[it] = s[threadIdx.x + it * FFT_TPB];
    __syncthreads();
#pragma unroll
    for (int it = 0; it < FFT_MAXV; ++it)
      if (cm.j0 + it * cm.js < n) s[threadIdx.x + it * FFT_TPB] = c[it];
  }
  __syncthreads();
  fft_run<-1, COLS, POW2, LN, ilog2c(BC), FFT_MAXV, FFT_TPB>(s, n, B, stw);
#pragma unroll
  for (int it = 0; it < FFT_MAXV; ++it) {
    const int j = cm.j0 + it * cm.js;
    if (j < n) {
      const float2 c = s[threadIdx.x + it * FFT_TPB];
      const float kmy = kt.km[1][j];
      a[it] = make_float2(kmy * a[it].x + kmz * c.x, kmy * a[it].y + kmz * c.y);
    }
  }
#pragma unroll
  for (int it = 0; it < FFT_MAXV; ++it)
    if (live && cm.j0 + it * cm.js < n) g1[it * gs] = a[it];
}

// ------------------------------------------------------------- copies --
// dst rows of dst_rs floats <- src rows of src_rs floats, n2 valid columns
__global__ void copy_rows_k(float* __restrict__ dst, int dst_rs, const float* __restrict__ src,
                            int src_rs, long long nrows, int n2) {
  long long total = nrows * n2;
  for (long long i = (long long)blockIdx.x * PMX_TPB + threadIdx.x; i < total;
       i += (long long)gridDim.x * PMX_TPB) {
    long long r = i / n2;
    int j = (int)(i - r * n2);
    dst[r * dst_rs + j] = src[r * src_rs + j];
  }
}

// ------------------------------------------------------------ drivers --
#define PMX_LDS_MAX (160 * 1024)     // LDS a gfx950 workgroup may allocate

// A pencil block must fit the per-thread batches (total <= cap) and LDS.
static int fft_fits(const char* who, int n, int total, int cap, size_t lds) {
  PMX_REQUIRE_CODE(total <= cap && lds <= PMX_LDS_MAX, PMX_EUNSUP,
                   "%s: axis of length %d needs %d values and %zu bytes of LDS per workgroup "
                   "(limits %d and %d)", who, n, total, lds, cap, PMX_LDS_MAX);
  return PMX_OK;
}

static size_t lds_bytes(int n, int P, int lay = COLS) {
  return (size_t)lds_elems(lay, n, P) * sizeof(float2);
}

// the twiddle table staged after the data block
static size_t tw_bytes(int n) { return (size_t)n * sizeof(float2); }


static bool half_ok(int n2) { return n2 >= 4 && (n2 & 1) == 0 && is_pow2(n2 / 2); }

// rows per workgroup, LDS bytes and pencil length of the z pass of n2
struct RowPlan {
  bool half;
  int P, len;
  size_t lds;
};

static RowPlan row_plan(int n2) {
  RowPlan r;
  r.half = half_ok(n2);
  r.len = r.half ? n2 / 2 : n2;
  r.P = row_pencils(r.len);
  r.lds = lds_bytes(r.len, r.P, ROWS) + tw_bytes(n2);
  return r;
}

// log2 of a length that has compile-time kernel variants (64..1024), else 0
static int ct_log2(int n) {
  if (!is_pow2(n) || n < 64 || n > 1024) return 0;
  int l = 0;
  while ((1 << l) < n) ++l;
  return l;
}

// CALL(L) with L the compile-time constant for ln (0: generic kernels)
#define PMX_CT_SWITCH(ln, CALL) \
  do {                          \
    switch (ln) {               \
      case 6: CALL(6); break;   \
      case 7: CALL(7); break;   \
      case 8: CALL(8); break;   \
      case 9: CALL(9); break;   \
      case 10: CALL(10); break; \
      default: CALL(0);         \
    }                           \
  } while (0)

int pmx_fold_deferrable(const pmx_ctx* c) {
  if (!c->fold_on_load) return 0;
  const TileGeom& t = c->tg;
  if (t.tx != 16 || t.ty != 16 || t.tz != 16) return 0;
  if (c->n[0] % 16 || c->n[1] % 16 || c->n[2] % 16) return 0;
  const RowPlan r = row_plan(c->n[2]);
  return r.half && ct_log2(r.len) > 0 && c->n[1] % r.P == 0;
}

// In-place r2c of a padded real mesh; pre applies (x - 1) * coef.  With
// halo (one channel's shells; pmx_fold_deferrable must hold) the deposit's
// fold is added on load.
int pmx_fft_r2c_inplace(pmx_ctx* c, float* buf, int pre, float coef, const float* halo, int rev) {
  const long long nrows = (long long)c->n[0] * c->n[1];
  const int n2 = c->n[2];
  const RowPlan r = row_plan(n2);
  if (int rc = fft_fits("z r2c", n2, r.len * r.P, RFFT_CAP, r.lds)) return rc;
  const unsigned nblk = (unsigned)((nrows + r.P - 1) / r.P);
  const int P = r.P;
  const size_t lds = r.lds;
  const FoldIn fi = {halo, c->tg.shell, c->tg.ntx, c->tg.nty, c->tg.ntz, c->n[1]};
  const float* fc = pre ? c->dfac : nullptr;
  PMX_REQUIRE(!halo || pmx_fold_deferrable(c), "z r2c: no fold on load for this mesh");
  pmx_tic(c, PMX_T_ZR2C);
  if (r.half) {
    const int lh = nrows % P == 0 ? ct_log2(r.len) : 0;
#define R2C_HALF(L)                                                                                \
  if (halo && pre)                                                                                 \
    z_r2c_half_k<1, L, (L > 0)><<<nblk, FFT_TPB, lds, c->stream>>>(buf, nrows, n2, c->rs, P, coef,  \
                                                                  c->tw[2], fi, fc, rev);              \
  else if (halo)                                                                                   \
    z_r2c_half_k<0, L, (L > 0)><<<nblk, FFT_TPB, lds, c->stream>>>(buf, nrows, n2, c->rs, P, coef,  \
                                                                  c->tw[2], fi, fc, rev);              \
  else if (pre)                                                                                    \
    z_r2c_half_k<1, L><<<nblk, FFT_TPB, lds, c->stream>>>(buf, nrows, n2, c->rs, P, coef, c->tw[2], fi, fc, rev); \
  else                                                                                             \
    z_r2c_half_k<0, L><<<nblk, FFT_TPB, lds, c->stream>>>(buf, nrows, n2, c->rs, P, coef, c->tw[2], fi, fc, rev)
    PMX_CT_SWITCH(lh, R2C_HALF);
#undef R2C_HALF
  } else {
    const bool p2 = is_pow2(n2);
    if (pre && p2)
      z_r2c_k<1, true><<<nblk, FFT_TPB, lds, c->stream>>>(buf, nrows, n2, c->rs, P, coef, c->tw[2], fc);
    else if (pre)
      z_r2c_k<1, false><<<nblk, FFT_TPB, lds, c->stream>>>(buf, nrows, n2, c->rs, P, coef, c->tw[2], fc);
    else if (p2)
      z_r2c_k<0, true><<<nblk, FFT_TPB, lds, c->stream>>>(buf, nrows, n2, c->rs, P, coef, c->tw[2], fc);
    else
      z_r2c_k<0, false><<<nblk, FFT_TPB, lds, c->stream>>>(buf, nrows, n2, c->rs, P, coef, c->tw[2], fc);
  }
  pmx_toc(c, PMX_T_ZR2C);
  PMX_LAUNCH_CHECK();
  return PMX_OK;
}

int pmx_fft_c2r_inplace(pmx_ctx* c, float* buf, float scale, int rev) {
  const long long nrows = (long long)c->n[0] * c->n[1];
  const int n2 = c->n[2];
  const RowPlan r = row_plan(n2);
  if (int rc = fft_fits("z c2r", n2, r.len * r.P, RFFT_CAP, r.lds)) return rc;
  const unsigned nblk = (unsigned)((nrows + r.P - 1) / r.P);
  const int P = r.P;
  const size_t lds = r.lds;
  pmx_tic(c, PMX_T_ZC2R);
  if (r.half) {
    const int lh = nrows % P == 0 ? ct_log2(r.len) : 0;
#define C2R_HALF(L) \
  z_c2r_half_k<L><<<nblk, FFT_TPB, lds, c->stream>>>(buf, nrows, n2, c->rs, P, scale, c->tw[2], rev)
    PMX_CT_SWITCH(lh, C2R_HALF);
#undef C2R_HALF
  } else if (is_pow2(n2)) {
    z_c2r_k<true><<<nblk, FFT_TPB, lds, c->stream>>>(buf, nrows, n2, c->rs, P, scale, c->tw[2]);
  } else {
    z_c2r_k<false><<<nblk, FFT_TPB, lds, c->stream>>>(buf, nrows, n2, c->rs, P, scale, c->tw[2]);
  }
  pmx_toc(c, PMX_T_ZC2R);
  PMX_LAUNCH_CHECK();
  return PMX_OK;
}

// axis 1 (y) or axis 0 (x) complex FFT of nf fields in place
int pmx_fft_strided(pmx_ctx* c, float2* const* f, int nf, int axis, int dir, int rev) {
  Fields3 fl;
  for (int i = 0; i < 3; ++i) fl.f[i] = f[i < nf ? i : 0];
  fl.rev = rev;
  int n = c->n[axis];
  int B = fft_pencils(n);
  int ncol = c->n2h;
  long long es, ostride;
  unsigned nouter;
  if (axis == 1) {
    es = c->n2p;
    ostride = (long long)c->n[1] * c->n2p;
    nouter = c->n[0];
  } else {
    es = (long long)c->n[1] * c->n2p;
    ostride = c->n2p;
    nouter = c->n[1];
  }
  dim3 grid((ncol + B - 1) / B, nouter, nf);
  size_t lds = lds_bytes(n, B) + tw_bytes(n);
  if (int rc = fft_fits("strided FFT", n, n * B, FFT_CAP, lds)) return rc;
  const bool p2 = is_pow2(n);
  const int ln = ct_log2(n);
  const int cls = axis == 1 ? (dir < 0 ? PMX_T_YFWD : PMX_T_YINV) : PMX_T_OTHER;
  pmx_tic(c, cls);
#define STRIDED(L)                                                                                  \
  if (dir < 0 && (L || p2))                                                                         \
    strided_fft_k<-1, true, L><<<grid, FFT_TPB, lds, c->stream>>>(fl, n, es, ostride, ncol, B, c->tw[axis]); \
  else if (dir < 0)                                                                                 \
    strided_fft_k<-1, false, 0><<<grid, FFT_TPB, lds, c->stream>>>(fl, n, es, ostride, ncol, B, c->tw[axis]); \
  else if (L || p2)                                                                                 \
    strided_fft_k<1, true, L><<<grid, FFT_TPB, lds, c->stream>>>(fl, n, es, ostride, ncol, B, c->tw[axis]);  \
  else                                                                                              \
    strided_fft_k<1, false, 0><<<grid, FFT_TPB, lds, c->stream>>>(fl, n, es, ostride, ncol, B, c->tw[axis])
  PMX_CT_SWITCH(ln, STRIDED);
#undef STRIDED
  pmx_toc(c, cls);
  PMX_LAUNCH_CHECK();
  return PMX_OK;
}

static KTab ktab(const pmx_ctx* c) {
  KTab kt;
  for (int a = 0; a < 3; ++a) {
    kt.k[a] = c->k[a];
    kt.km[a] = c->km[a];
  }
  return kt;
}

// Paired y pass of the gradient's y and z components: dir > 0 turns h in
// f[1] into the two force components in f[1] and f[2] (y_inv_seq_k);
// dir < 0 turns the cotangents in f[1] and f[2] into km_1 FFT_y(f[1]) +
// km_2 FFT_y(f[2]) in f[1] (y_fwd_seq_k).
int pmx_fft_ypair(pmx_ctx* c, float2* const* f, int dir, int rev) {
  Fields3 fl = {{f[0], f[1], f[2]}, rev};
  const int n = c->n[1];
  const int B = fft_pencils(n);
  const int ncol = c->n2h;
  const long long es = c->n2p, ostride = (long long)n * c->n2p;
  dim3 grid((ncol + B - 1) / B, c->n[0], 1);
  const size_t lds = lds_bytes(n, B) + tw_bytes(n);
  if (int rc = fft_fits("paired y FFT", n, n * B, FFT_CAP, lds)) return rc;
  const KTab kt = ktab(c);
  const bool p2 = is_pow2(n);
  const int ln = ct_log2(n);
  const int cls = dir < 0 ? PMX_T_YFWD : PMX_T_YINV;
  pmx_tic(c, cls);
#define YPAIR(L)                                                                                  \
  if (dir < 0 && (L || p2))                                                                       \
    y_fwd_seq_k<true, L><<<grid, FFT_TPB, lds, c->stream>>>(fl, n, es, ostride, ncol, B, kt, c->tw[1]); \
  else if (dir < 0)                                                                               \
    y_fwd_seq_k<false, 0><<<grid, FFT_TPB, lds, c->stream>>>(fl, n, es, ostride, ncol, B, kt, c->tw[1]); \
  else if (L || p2)                                                                               \
    y_inv_seq_k<true, L><<<grid, FFT_TPB, lds, c->stream>>>(fl, n, es, ostride, ncol, B, kt, c->tw[1]); \
  else                                                                                            \
    y_inv_seq_k<false, 0><<<grid, FFT_TPB, lds, c->stream>>>(fl, n, es, ostride, ncol, B, kt, c->tw[1])
  PMX_CT_SWITCH(ln, YPAIR);
#undef YPAIR
  pmx_toc(c, cls);
  PMX_LAUNCH_CHECK();
  return PMX_OK;
}

int pmx_poisson_x(pmx_ctx* c, float2* const* f, int adjoint, int kin, int rev) {
  Fields3 fl = {{f[0], f[1], f[2]}, rev};
  const int n0 = c->n[0];
  const int B = fft_pencils(n0, PFFT_CAP);
  dim3 grid((c->n2h + B - 1) / B, c->n[1], 1);
  // the data block, the twiddles, then the k_0 and km_0 tables
  size_t lds = lds_bytes(n0, B) + tw_bytes(n0) + 2 * (size_t)n0 * sizeof(float);
  if (int rc = fft_fits("Poisson x pass", n0, n0 * B, PFFT_CAP, lds)) return rc;
  const bool p2 = is_pow2(n0);
  const int ln = ct_log2(n0);
  const KTab kt = ktab(c);
  pmx_tic(c, PMX_T_POISSON);
#define PX(L)                                                                                             \
  if (adjoint && (L || p2))                                                                               \
    poisson_x_adj_k<true, L><<<grid, PFFT_TPB, lds, c->stream>>>(fl, n0, c->n[1], c->n2h, c->n2p, B, kt, c->tw[0]); \
  else if (adjoint)                                                                                       \
    poisson_x_adj_k<false, 0><<<grid, PFFT_TPB, lds, c->stream>>>(fl, n0, c->n[1], c->n2h, c->n2p, B, kt, c->tw[0]); \
  else if ((L || p2) && kin)                                                                              \
    poisson_x_k<true, L, true><<<grid, PFFT_TPB, lds, c->stream>>>(fl, n0, c->n[1], c->n2h, c->n2p, B, kt, c->tw[0]); \
  else if (kin)                                                                                           \
    poisson_x_k<false, 0, true><<<grid, PFFT_TPB, lds, c->stream>>>(fl, n0, c->n[1], c->n2h, c->n2p, B, kt, c->tw[0]); \
  else if (L || p2)                                                                                       \
    poisson_x_k<true, L><<<grid, PFFT_TPB, lds, c->stream>>>(fl, n0, c->n[1], c->n2h, c->n2p, B, kt, c->tw[0]); \
  else                                                                                                    \
    poisson_x_k<false, 0><<<grid, PFFT_TPB, lds, c->stream>>>(fl, n0, c->n[1], c->n2h, c->n2p, B, kt, c->tw[0])
  PMX_CT_SWITCH(ln, PX);
#undef PX
  pmx_toc(c, PMX_T_POISSON);
  PMX_LAUNCH_CHECK();
  return PMX_OK;
}

// Full forward Poisson on the context buffers: cbuf[0] holds the density in
// the padded real layout (deferred: its tile interiors, the fold of c->halo
// still to add); on return cbuf[0..2] hold the three force meshes (padded
// real layout), scaled 1/M.
int pmx_poisson_ctx(pmx_ctx* c, float coef, int deferred) {
  int rc;
  float2* f[3] = {c->cbuf[0], c->cbuf[1], c->cbuf[2]};
  const float* halo = deferred && pmx_fold_deferrable(c) ? c->halo : nullptr;
  // block directions alternate from pass to pass (pass_block), the first
  // against the deposit's and the last against the gather's; the z passes
  // run y's components first, the x component last
  if ((rc = pmx_fft_r2c_inplace(c, (float*)f[0], 1, coef, halo, 1))) return rc;
  if ((rc = pmx_fft_strided(c, f, 1, 1, -1, 0))) return rc;
  if ((rc = pmx_poisson_x(c, f, 0, 0, 1))) return rc;
  if ((rc = pmx_fft_strided(c, f, 1, 1, 1, 0))) return rc;
  if ((rc = pmx_fft_ypair(c, f, 1, 1))) return rc;
  float scale = (float)(1.0 / (double)c->M);
  if ((rc = pmx_fft_c2r_inplace(c, (float*)f[1], scale, 0))) return rc;
  if ((rc = pmx_fft_c2r_inplace(c, (float*)f[2], scale, 1))) return rc;
  if ((rc = pmx_fft_c2r_inplace(c, (float*)f[0], scale, 1))) return rc;
  return PMX_OK;
}

// Adjoint Poisson on the context buffers: cbuf[0..2] hold the three
// cotangent meshes (padded real; deferred: with the folds of c->halo's three
// channels still to add); on return cbuf[0] holds the density cotangent
// (padded real) times `scale_out`.
int pmx_poisson_adj_ctx(pmx_ctx* c, float scale_out, int deferred) {
  float2* f[3] = {c->cbuf[0], c->cbuf[1], c->cbuf[2]};
  return pmx_poisson_adj_bufs(c, f, scale_out, deferred);
}

// The same on three other mesh buffers of the context's layout.
int pmx_poisson_adj_bufs(pmx_ctx* c, float2* const* f, float scale_out, int deferred) {
  int rc;
  const bool fold = deferred && pmx_fold_deferrable(c);
  const long long hstride = c->ntiles * c->tg.shell;
  // block directions alternate as in pmx_poisson_ctx
  for (int d = 0; d < 3; ++d)
    if ((rc = pmx_fft_r2c_inplace(c, (float*)f[d], 0, 0.f, fold ? c->halo + d * hstride : nullptr,
                                  1)))
      return rc;
  if ((rc = pmx_fft_strided(c, f, 1, 1, -1, 0))) return rc;
  if ((rc = pmx_fft_ypair(c, f, -1, 0))) return rc;
  if ((rc = pmx_poisson_x(c, f, 1, 0, 1))) return rc;
  if ((rc = pmx_fft_strided(c, f, 1, 1, 1, 0))) return rc;
  if ((rc = pmx_fft_c2r_inplace(c, (float*)f[0], scale_out, 1))) return rc;
  return PMX_OK;
}

// copies n0*n1 rows of ncols floats between row strides (ncols <= 0: n2)
int pmx_copy_rows(pmx_ctx* c, float* dst, int dst_rs, const float* src, int src_rs, int ncols) {
  if (ncols <= 0) ncols = c->n[2];
  long long nrows = (long long)c->n[0] * c->n[1];
  long long total = nrows * ncols;
  unsigned nblk = (unsigned)((total + PMX_TPB - 1) / PMX_TPB);
  if (nblk > 65536u * 4) nblk = 65536u * 4;
  copy_rows_k<<<nblk, PMX_TPB, 0, c->stream>>>(dst, dst_rs, src, src_rs, nrows, ncols);
  PMX_LAUNCH_CHECK();
  return PMX_OK;
}

extern "C" int pmx_rfftn(pmx_ctx* c, const float* in, float* out_c) {
  PMX_ENTRY(c);
  PMX_REQUIRE(c && in && out_c, "pmx_rfftn: NULL argument");
  int rc;
  float2* f[3] = {c->cbuf[0], c->cbuf[1], c->cbuf[2]};
  if ((rc = pmx_copy_rows(c, (float*)f[0], c->rs, in, c->n[2]))) return rc;
  if ((rc = pmx_fft_r2c_inplace(c, (float*)f[0], 0, 0.f))) return rc;
  if ((rc = pmx_fft_strided(c, f, 1, 1, -1))) return rc;
  if ((rc = pmx_fft_strided(c, f, 1, 0, -1))) return rc;
  return pmx_copy_rows(c, out_c, 2 * c->n2h, (const float*)f[0], c->rs, 2 * c->n2h);
}

extern "C" int pmx_irfftn(pmx_ctx* c, const float* in_c, float* out) {
  PMX_ENTRY(c);
  PMX_REQUIRE(c && in_c && out, "pmx_irfftn: NULL argument");
  int rc;
  float2* f[3] = {c->cbuf[0], c->cbuf[1], c->cbuf[2]};
  if ((rc = pmx_copy_rows(c, (float*)f[0], c->rs, in_c, 2 * c->n2h, 2 * c->n2h))) return rc;
  if ((rc = pmx_fft_strided(c, f, 1, 0, 1))) return rc;
  if ((rc = pmx_fft_strided(c, f, 1, 1, 1))) return rc;
  if ((rc = pmx_fft_c2r_inplace(c, (float*)f[0], (float)(1.0 / (double)c->M)))) return rc;
  return pmx_copy_rows(c, out, c->n[2], (const float*)f[0], c->rs);
}

extern "C" int pmx_poisson_grad(pmx_ctx* c, const float* dens, float coef, float* grad) {
  PMX_ENTRY(c);
  PMX_REQUIRE(c && dens && grad, "pmx_poisson_grad: NULL argument");
  int rc;
  if ((rc = pmx_copy_rows(c, (float*)c->cbuf[0], c->rs, dens, c->n[2]))) return rc;
  if ((rc = pmx_poisson_ctx(c, coef))) return rc;
  for (int d = 0; d < 3; ++d)
    if ((rc = pmx_copy_rows(c, grad + (long long)d * c->M, c->n[2], (const float*)c->cbuf[d], c->rs)))
      return rc;
  return PMX_OK;
}

extern "C" int pmx_poisson_grad_adj(pmx_ctx* c, const float* grad_cot, float coef, float* dens_cot) {
  PMX_ENTRY(c);
  PMX_REQUIRE(c && grad_cot && dens_cot, "pmx_poisson_grad_adj: NULL argument");
  int rc;
  for (int d = 0; d < 3; ++d)
    if ((rc = pmx_copy_rows(c, (float*)c->cbuf[d], c->rs, grad_cot + (long long)d * c->M, c->n[2])))
      return rc;
  // dens_cot = -coef * sum_d L_d(grad_cot_d), L_d = irfftn (-ik_d) laplace rfftn
  if ((rc = pmx_poisson_adj_ctx(c, (float)(-(double)coef / (double)c->M), 0))) return rc;
  return pmx_copy_rows(c, dens_cot, c->n[2], (const float*)c->cbuf[0], c->rs);
}

// ------------------------------------------------------- power spectrum --
// pmwd/spec_util.py:47-147: per complex cell of the r2c layout, |k| from
// fftfreq(shape, None) in float32 (the integer frequency divided by the
// extent in float64, then rounded, as jnp.fft.fftfreq), P = |F|^2 or
// F conj(G) in float32 (optionally deconvolved by sinc^-deconv per axis),
// the weight N = 2 (1 on the kz = 0 and, for an even n2, Nyquist planes),
// and the bin of |k| among the float64 edges (right: edge < k counts; else
// edge <= k).  Each workgroup sums k N, Re P N, Im P N and N per bin in
// float64 in LDS (atomic adds: their order within a workgroup is not
// fixed) and writes its partials; a second kernel adds the partials in
// workgroup order.
#define PS_TPB 256
#define PS_BLOCKS 1024

__device__ __forceinline__ float ps_freq(int j, int n, bool rfft) {
  const int jj = (!rfft && j >= (n + 1) / 2) ? j - n : j;
  return (float)((double)jj / (double)n);
}

__device__ __forceinline__ float ps_sinc(float x) {
  if (x == 0.f) return 1.f;
  const float px = 3.14159265358979323846f * x;
  return sinf(px) / px;
}

__global__ void __launch_bounds__(PS_TPB)
    powspec_bin_k(const float2* __restrict__ F, const float2* __restrict__ G, int n0, int n1,
                  int n2, int n2h, int n2p, const double* __restrict__ edges, int nedges, int right,
                  int deconv, double* __restrict__ part) {
  extern __shared__ double ps_sh[];
  double* acc = ps_sh;                 // 4 * nedges
  double* e = ps_sh + 4 * nedges;      // nedges
  for (int i = threadIdx.x; i < 4 * nedges; i += PS_TPB) acc[i] = 0.0;
  for (int i = threadIdx.x; i < nedges; i += PS_TPB) e[i] = edges[i];
  __syncthreads();
  const long long total = (long long)n0 * n1 * n2h;
  for (long long idx = (long long)blockIdx.x * PS_TPB + threadIdx.x; idx < total;
       idx += (long long)gridDim.x * PS_TPB) {
    const int kz = (int)(idx % n2h);
    const long long r = idx / n2h;
    const int ky = (int)(r % n1);
    const int kx = (int)(r / n1);
    const float2 f = F[((long long)kx * n1 + ky) * n2p + kz];
    float pr, pi = 0.f;
    if (G) {
      const float2 gg = G[((long long)kx * n1 + ky) * n2p + kz];
      pr = f.x * gg.x + f.y * gg.y;
      pi = f.y * gg.x - f.x * gg.y;
    } else {
      pr = f.x * f.x + f.y * f.y;
    }
    const float fx = ps_freq(kx, n0, false), fy = ps_freq(ky, n1, false),
                fz = ps_freq(kz, n2, true);
    const float k = sqrtf((fx * fx + fy * fy) + fz * fz);
    if (deconv) {
      const float sx = powf(ps_sinc(fx), (float)-deconv);
      const float sy = powf(ps_sinc(fy), (float)-deconv);
      const float sz = powf(ps_sinc(fz), (float)-deconv);
      pr = ((pr * sx) * sy) * sz;
      pi = ((pi * sx) * sy) * sz;
    }
    // number of edges below k (right) or at most k (left-inclusive bins)
    const double kd = (double)k;
    int lo = 0, hi = nedges;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      const bool below = right ? (e[mid] < kd) : (e[mid] <= kd);
      if (below) lo = mid + 1;
      else hi = mid;
    }
    if (lo >= nedges) continue;
    const float w = (kz == 0 || ((n2 & 1) == 0 && kz == n2 / 2)) ? 1.f : 2.f;
    atomicAdd(&acc[lo], (double)(k * w));
    atomicAdd(&acc[nedges + lo], (double)(pr * w));
    if (G) atomicAdd(&acc[2 * nedges + lo], (double)(pi * w));
    atomicAdd(&acc[3 * nedges + lo], (double)w);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < 4 * nedges; i += PS_TPB)
    part[(long long)blockIdx.x * 4 * nedges + i] = acc[i];
}

__global__ void powspec_reduce_k(const double* __restrict__ part, int nblocks, int m,
                                 double* __restrict__ out) {
  const int i = blockIdx.x * PS_TPB + threadIdx.x;
  if (i >= m) return;
  double s = 0.0;
  for (int b = 0; b < nblocks; ++b) s += part[(long long)b * m + i];
  out[i] = s;
}

// r2c of a caller mesh into context buffer cb (padded layout, all axes)
static int ps_r2c(pmx_ctx* c, const float* in, int cb) {
  int rc;
  float2* f[3] = {c->cbuf[cb], c->cbuf[0], c->cbuf[0]};
  if ((rc = pmx_copy_rows(c, (float*)f[0], c->rs, in, c->n[2]))) return rc;
  if ((rc = pmx_fft_r2c_inplace(c, (float*)f[0], 0, 0.f))) return rc;
  if ((rc = pmx_fft_strided(c, f, 1, 1, -1))) return rc;
  return pmx_fft_strided(c, f, 1, 0, -1);
}

extern "C" int pmx_powspec(pmx_ctx* c, const float* f, const float* g, int nedges,
                           const double* edges, int right, int deconv, double* out) {
  PMX_ENTRY(c);
  PMX_REQUIRE(c && f && edges && out, "pmx_powspec: NULL argument");
  PMX_REQUIRE(nedges >= 1 && nedges <= 2048, "pmx_powspec: 1..2048 bin edges");
  PMX_REQUIRE(deconv >= 0 && deconv <= 16, "pmx_powspec: deconv must be in 0..16");
  int rc;
  // the binned sums are read back on the host
  if ((rc = pmx_no_capture(c, "pmx_powspec"))) return rc;
  const int m = 4 * nedges;
  // partial sums of the workgroups, then the reduced sums and the edges
  const long long need = (long long)PS_BLOCKS * m + m + nedges;
  if ((rc = pmx_ensure_red(c, need))) return rc;
  double* part = c->red;
  double* sums = c->red + (long long)PS_BLOCKS * m;
  double* dedges = sums + m;
  PMX_CHECK_HIP(hipMemcpyAsync(dedges, edges, (size_t)nedges * sizeof(double),
                               hipMemcpyHostToDevice, c->stream));
  if ((rc = ps_r2c(c, f, 1))) return rc;
  if (g && (rc = ps_r2c(c, g, 2))) return rc;
  const size_t lds = (size_t)(m + nedges) * sizeof(double);
  powspec_bin_k<<<PS_BLOCKS, PS_TPB, lds, c->stream>>>(
      c->cbuf[1], g ? c->cbuf[2] : nullptr, c->n[0], c->n[1], c->n[2], c->n2h, c->n2p, dedges,
      nedges, right, deconv, part);
  PMX_LAUNCH_CHECK();
  powspec_reduce_k<<<(m + PS_TPB - 1) / PS_TPB, PS_TPB, 0, c->stream>>>(part, PS_BLOCKS, m, sums);
  PMX_LAUNCH_CHECK();
  PMX_CHECK_HIP(hipMemcpyAsync(out, sums, (size_t)m * sizeof(double), hipMemcpyDeviceToHost,
                               c->stream));
  PMX_CHECK_HIP(hipStreamSynchronize(c->stream));
  return PMX_OK;
}
