// Internal definitions shared by the pmx HIP translation units (gfx950 only).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>
#include <string>

#include "../../include/pmx.h"

#define PMX_TPB 256

// ---------------------------------------------------------------- errors --
void pmx_set_error(const char* fmt, ...);

#define PMX_CHECK_HIP(expr)                                                      \
  do {                                                                           \
    hipError_t _e = (expr);                                                      \
    if (_e != hipSuccess) {                                                      \
      pmx_set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e),       \
                    __FILE__, __LINE__);                                         \
      return PMX_EHIP;                                                           \
    }                                                                            \
  } while (0)

#define PMX_REQUIRE(cond, ...)                                                   \
  do {                                                                           \
    if (!(cond)) {                                                               \
      pmx_set_error(__VA_ARGS__);                                                \
      return PMX_EINVAL;                                                         \
    }                                                                            \
  } while (0)

#define PMX_LAUNCH_CHECK()                                                       \
  do {                                                                           \
    hipError_t _e = hipGetLastError();                                           \
    if (_e != hipSuccess) {                                                      \
      pmx_set_error("kernel launch failed: %s (%s:%d)", hipGetErrorString(_e),   \
                    __FILE__, __LINE__);                                         \
      return PMX_EHIP;                                                           \
    }                                                                            \
  } while (0)

// ------------------------------------------------------------ geometry --
// Mesh geometry handed to kernels by value.
struct MeshGeom {
  int n0, n1, n2;      // mesh shape (periodic wrap shape)
  int rs;              // row stride in floats along the last axis (>= n2)
  float cell;          // float32 cell size (enmesh a1, pm_util.py:83)
  float inv_cell;      // unused by bit-exact paths; only for gradients' /cell
};

// Scatter binning: the mesh is cut into tiles of TX x TY x TZ cells.
struct TileGeom {
  int tx, ty, tz;      // tile extent (clipped to the mesh)
  int ntx, nty, ntz;   // tiles per axis
  int shell;           // floats reserved per tile for the +1 halo shell
};

// ---------------------------------------------------------- kernel timing --
// Optional per-kernel-class timing with HIP events on the context stream.
enum {
  PMX_T_BIN = 0,     // tile count (+ fused kick/drift)
  PMX_T_SCAN,        // exclusive scan of tile counts
  PMX_T_FILL,        // bin fill
  PMX_T_DEPOSIT,     // LDS tile accumulation
  PMX_T_FOLD,        // halo fold
  PMX_T_ZR2C,        // z r2c row pass
  PMX_T_YFWD,        // y forward pass
  PMX_T_POISSON,     // fused x pass (Green, -ik, inverse x)
  PMX_T_YINV,        // y inverse pass
  PMX_T_ZC2R,        // z c2r row pass
  PMX_T_GATHER,      // 3-component gather (+ fused kick)
  PMX_T_VJPG,        // gather with the gather-adjoint disp cotangent
  PMX_T_GRADG,       // scatter-adjoint disp cotangent (weight gradients)
  PMX_T_KDADJ,       // reverse-time kick/drift with cotangent updates
  PMX_T_OTHER,
  PMX_NTIMER
};
static_assert(PMX_NTIMER == PMX_NTIMER_CLASSES, "timer classes out of step with pmx.h");

struct PmxTimer;
void pmx_tic(pmx_ctx* c, int cls);
void pmx_toc(pmx_ctx* c, int cls);

// ------------------------------------------------------------- context --
struct pmx_ctx {
  PmxTimer* timer;     // NULL unless timing is enabled
  int n[3];
  int n2h;             // n2/2+1 complex columns of the r2c layout
  int n2p;             // complex row stride: n2h rounded up to 16 (128 B) above 16
  int rs;              // padded real row stride = 2*n2p floats
  long long M;         // real cells
  long long Mc;        // complex cells held, n0*n1*n2p
  double cell;         // float64 cell size (fftfreq, pm_util.py:187-195)
  float cellf;         // float32 cell size
  hipStream_t stream;
  int device;

  // k tables (float32 as pmwd: fftfreq(...).astype(f32)); *_m has the Nyquist
  // bin zeroed (gravity.py:38-42)
  float* k[3];
  float* km[3];
  // twiddle tables exp(-2 pi i j/n) j<n per axis (forward); inverse uses conj
  float2* tw[3];

  // three complex (or padded real) mesh buffers of Mc float2 each
  float2* cbuf[3];

  // scatter binning scratch
  TileGeom tg;
  long long ntiles;
  uint32_t* tile_count;
  uint32_t* tile_start;
  float* halo;
  long long ptcl_cap;
  float4* rec;         // binned particle records (d per axis, packed local cell)
  uint32_t* pkey;      // tile of each particle
  uint32_t* prank;     // rank of each particle within its tile
  uint32_t* pidx;      // particle index of each record
  float* vrec;         // up to 3 channels of per-record values (adjoint)
  float* vmax;         // max |value| per channel of vrec (fixed-point scale)
  long long vrec_cap;

  // reductions (float64 partial sums)
  double* red;
  int red_cap;
  double* red_host;    // pinned

  // radix sort scratch (lazily sized)
  long long sort_cap;
  uint32_t* sort_k[2];
  uint32_t* sort_v[2];
  uint32_t* sort_hist;
  long long sort_hist_cap;
  uint32_t* scan_tmp;
  long long scan_tmp_cap;

  size_t bytes;        // device bytes held
};

int pmx_ensure_ptcl(pmx_ctx* c, long long n);
int pmx_ensure_vrec(pmx_ctx* c, long long n);
int pmx_ensure_red(pmx_ctx* c, int nblk);
int pmx_scan_exclusive(pmx_ctx* c, const uint32_t* in, uint32_t* out, long long n);
MeshGeom pmx_geom_padded(const pmx_ctx* c);

// grid helpers
static inline unsigned pmx_blocks(long long n, int tpb = PMX_TPB) {
  long long b = (n + tpb - 1) / tpb;
  if (b < 1) b = 1;
  return (unsigned)b;
}

// ------------------------------------------------------- device helpers --
// Workgroups are dispatched round-robin over the 8 XCDs (WG b runs on XCD
// b mod 8), each with its own L2.  This maps the hardware block index to a
// logical one so that each XCD works on one contiguous 1/8 of the logical
// range, keeping neighbouring work (and its shared cache lines) in one L2.
#define PMX_NXCD 8
__device__ __forceinline__ unsigned pmx_xcd_block(unsigned b, unsigned nb) {
  unsigned q = nb / PMX_NXCD, r = nb % PMX_NXCD;
  unsigned k = b % PMX_NXCD, i = b / PMX_NXCD;
  return (k < r ? k * (q + 1) : r * (q + 1) + (k - r) * q) + i;
}

__device__ __forceinline__ int pmx_wrap(int i, int n) {
  int r = i % n;
  return r < 0 ? r + n : r;
}

// Bit-exact CIC setup of enmesh's default branch (pm_util.py:119-137):
//   d  = disp / cell                (IEEE float32 division)
//   i  = floor(d)                   (int16 in pmwd; int here, same values)
//   d2 = d - (i + nb)               (float32)
//   f  = 1 - |d2|
//   lo = (i + pmid) mod n           (floor mod), hi = (lo + 1) mod n
// w0/w1 are the per-axis fractions of the nb=0/1 neighbours, and s0/s1 the
// signs sign(-d2) used by the fraction gradients (pm_util.py:143-152).
struct CicAxis {
  int lo, hi;
  float w0, w1;
  float s0, s1;
};

__device__ __forceinline__ float pmx_sign(float x) {
  return x > 0.f ? 1.f : (x < 0.f ? -1.f : 0.f);
}

__device__ __forceinline__ CicAxis pmx_cic_axis(int pm, float dp, float cell, int n) {
  CicAxis a;
  float d = dp / cell;
  float fl = floorf(d);
  int i = (int)fl;
  float d20 = d - (float)i;
  float d21 = d - (float)(i + 1);
  a.w0 = 1.f - fabsf(d20);
  a.w1 = 1.f - fabsf(d21);
  a.s0 = pmx_sign(-d20);
  a.s1 = pmx_sign(-d21);
  a.lo = pmx_wrap(i + pm, n);
  a.hi = (a.lo + 1 == n) ? 0 : a.lo + 1;
  return a;
}

// neighbour j: bit0 -> axis 0, bit1 -> axis 1, bit2 -> axis 2 (pm_util.py:95-97)
// weight (f0*f1)*f2 as the reference's prod over the last axis.
__device__ __forceinline__ float pmx_w(const CicAxis* ax, int j) {
  float f0 = (j & 1) ? ax[0].w1 : ax[0].w0;
  float f1 = (j & 2) ? ax[1].w1 : ax[1].w0;
  float f2 = (j & 4) ? ax[2].w1 : ax[2].w0;
  return (f0 * f1) * f2;
}

// Gradient of the weight of neighbour j w.r.t. d (pm_util.py:143-152):
// axis 0: s0 * (f1*f2); axis 1: s1 * (f2*f0); axis 2: s2 * (f0*f1)
__device__ __forceinline__ void pmx_wgrad(const CicAxis* ax, int j, float g[3]) {
  float f0 = (j & 1) ? ax[0].w1 : ax[0].w0;
  float f1 = (j & 2) ? ax[1].w1 : ax[1].w0;
  float f2 = (j & 4) ? ax[2].w1 : ax[2].w0;
  float s0 = (j & 1) ? ax[0].s1 : ax[0].s0;
  float s1 = (j & 2) ? ax[1].s1 : ax[1].s0;
  float s2 = (j & 4) ? ax[2].s1 : ax[2].s0;
  g[0] = s0 * (f1 * f2);
  g[1] = s1 * (f2 * f0);
  g[2] = s2 * (f0 * f1);
}

__device__ __forceinline__ long long pmx_cell_index(const CicAxis* ax, int j, int n1, int rs) {
  int i0 = (j & 1) ? ax[0].hi : ax[0].lo;
  int i1 = (j & 2) ? ax[1].hi : ax[1].lo;
  int i2 = (j & 4) ? ax[2].hi : ax[2].lo;
  return ((long long)i0 * n1 + i1) * rs + i2;
}

// fixed tree order used for every 8-neighbour sum
__device__ __forceinline__ float pmx_tree8(const float v[8]) {
  return ((v[0] + v[1]) + (v[2] + v[3])) + ((v[4] + v[5]) + (v[6] + v[7]));
}
