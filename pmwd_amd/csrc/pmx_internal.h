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

#define PMX_REQUIRE_CODE(cond, code, ...)                                        \
  do {                                                                           \
    if (!(cond)) {                                                               \
      pmx_set_error(__VA_ARGS__);                                                \
      return code;                                                               \
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
  // 1/cell when cell is a power of two (then x * inv_cell == x / cell for
  // every float x: both round the same exact value), else 0
  float inv_cell;
};

// Scatter binning: the mesh is cut into tiles of TX x TY x TZ cells.
struct TileGeom {
  int tx, ty, tz;      // tile extent (clipped to the mesh)
  int ntx, nty, ntz;   // tiles per axis
  int shell;           // floats reserved per tile for the +1 halo shell
  // ceil(2^20 / extent): lo / t == (lo * m) >> 20 exactly for lo < 4096 and
  // t <= 16 (error below lo / 2^20 <= 1/256 < 1/t), a 24-bit multiply
  unsigned mx, my, mz;
};

__device__ __forceinline__ int pmx_tdiv(int lo, unsigned m) {
  return (int)(__umul24((unsigned)lo, m) >> 20);
}

// A particle held in tile order by the multi-step pipeline: s0 = (disp,
// pm_x | pm_y << 10 | pm_z << 20), s1 = (vel, index in the caller's arrays),
// pm taken mod the mesh extent per axis (the CIC corner wraps, so only pm
// mod n enters; the pipelines take meshes of at most PMX_PL_MAXN cells per
// axis, so each fits 10 bits).
#define PMX_PL_MAXN 1024
struct PState {
  float4* s0;
  float4* s1;
};

// The reverse-time pipeline's particle: a PState plus c0 = (disp_cot,
// vel_cot.x) and c1 = (vel_cot.y, vel_cot.z).
struct AState {
  PState p;
  float4* c0;
  float2* c1;
};

// Per-channel value sources of a deposit: channel ch of slot i is
// p[ch][i * step[ch]].
struct ValSrc {
  const float* p[3];
  int step[3];
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
  // device allocator of pmx_set_allocator at creation (NULL: hipMalloc/hipFree)
  pmx_alloc_fn dalloc;
  pmx_free_fn dfree;
  void* dalloc_user;
  // recorded on the context's stream as each entry point returns (not under
  // graph capture); a stream the context is rebound to waits for it, so the
  // previous stream's handle is never touched again (moved_valid: the last
  // record follows every call so far)
  hipEvent_t moved;
  int moved_valid;

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
  // Capacity binning: each tile gets room for its count of the last binning
  // plus slack (tile_cap), tile_start the scan of the capacities, and the
  // binning kernel writes the records itself; flags[0] is set on the device
  // when a tile overflows, which runs the exact count/scan/fill after it.
  uint32_t* tile_cap;
  uint32_t* flags;
  int caps_valid;      // host: tile_count holds the exact counts of caps_n particles
  long long caps_n;
  long long rec_cap;   // record slots of rec/pidx (and the vrec channel stride)
  // multi-step pipeline: two tile-ordered particle copies of rec_cap slots
  // (the second aliases rec/vrec/pidx), their tile counts and starts, and
  // the per-slot tile/rank of an exact re-bin (aliases pkey/prank)
  PState pl[2];
  uint32_t* pl_count[2];
  uint32_t* pl_start[2];
  uint32_t* pl_tile;
  uint32_t* pl_rank;
  long long pl_cap;    // slots of the pipeline's own copy
  // reverse-time pipeline: its own particle copy (the second aliases rec,
  // vrec, pidx and the forward pipeline's copy) and three mesh buffers
  AState al[2];
  long long al_cap;
  float2* abuf[3];     // the reverse pipeline's cotangent meshes (F stays in cbuf)
  double* al_dots;     // per-step dot products of one reverse run (device)
  long long al_dots_cap;
  float* halo;         // nch shells per tile, channel-major
  int halo_nch;        // channels the halo buffer holds
  int fold_on_load;    // the Poisson row pass may take the deposit's fold
                       // (environment PMX_FOLD_ON_LOAD=0 at creation: never)
  int deposit_split;   // several-channel deposits: a workgroup per tile and
                       // channel (environment PMX_DEPOSIT_SPLIT=0 at
                       // creation: the channels in turn in one workgroup)
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
  long long red_cap;
  double* red_out;     // PMX_RED_SLOTS reduced sums (device)
  double* red_host;    // PMX_RED_SLOTS pinned host doubles

  // radix sort scratch (lazily sized)
  long long sort_cap;
  uint32_t* sort_k[2];
  uint32_t* sort_v[2];
  uint32_t* sort_hist;
  long long sort_hist_cap;
  uint32_t* scan_tmp;
  long long scan_tmp_cap;

  size_t bytes;        // device bytes held
  int no_alloc;        // 1: scratch carved from static device memory; never allocate

  // Device factor table of the call in progress (pmx_nbody_steps_dev,
  // pmx_graph_capture), NULL otherwise: {1.5 Omega_m, k1_0, d_0, k2_0, k1_1,
  // d_1, k2_1, ...} in float32.  The kernels that take a factor read it from
  // here at run time, so a captured graph replays for any cosmology.
  const float* dfac;
  // the 2LPT source field of pmx_lpt (one more padded mesh, lazily allocated)
  float2* lbuf;
  // float64 tables of pmx_linear_modes (transfer k and T)
  double* tab;
  size_t tab_cap;
  // the graph of the last pmx_graph_capture and its executable
  hipGraph_t graph;
  hipGraphExec_t gexec;
};

// Geometry fields of a context (extents, padded strides, cell size, stream,
// device, scatter tiles); no device memory.
void pmx_ctx_geometry(pmx_ctx* c, int n0, int n1, int n2, double cell_size, hipStream_t stream);

int pmx_ensure_ptcl(pmx_ctx* c, long long n);
int pmx_ensure_vrec(pmx_ctx* c);
int pmx_ensure_red(pmx_ctx* c, long long nblk);

// Whether the context's stream is under graph capture (1), not (0).
int pmx_capturing(pmx_ctx* c);
// PMX_EINVAL with a message when the context's stream is under graph
// capture: device allocations and host waits synchronise the stream, which a
// capture forbids.  Callers that capture run each call once beforehand, so
// that the context has grown to its size, and read host values outside.
int pmx_no_capture(pmx_ctx* c, const char* what);
// Records c->moved on the context's stream unless it is capturing.
void pmx_mark_queued(pmx_ctx* c);
struct PmxQueued {
  pmx_ctx* c;
  explicit PmxQueued(pmx_ctx* c_) : c(c_) {}
  ~PmxQueued() { pmx_mark_queued(c); }
};
// First statement of every entry point that enqueues work on the context.
#define PMX_ENTRY(ctx) PmxQueued pmx_queued_guard_(ctx)
int pmx_ensure_scan(pmx_ctx* c, long long n);
// Blocks (scratch words) of a scan of n entries.
long long pmx_scan_blocks(long long n);
// Device memory of a context through its allocator (pmx_set_allocator or
// hipMalloc), counted in pmx_ctx_bytes; the free takes the allocated size.
int pmx_dev_alloc(pmx_ctx* c, void** p, size_t bytes);
void pmx_dev_free(pmx_ctx* c, void** p, size_t bytes);
int pmx_scan_exclusive(pmx_ctx* c, const uint32_t* in, uint32_t* out, long long n,
                       const uint32_t* cond = nullptr);

// Slack of a tile's capacity over its last count, and the record slots the
// capacities of n particles can take (sum over tiles of count + slack).
__host__ __device__ inline uint32_t pmx_tile_slack(uint32_t count) { return (count >> 3) + 32; }
inline long long pmx_record_cap(const pmx_ctx* c, long long n) {
  return n + n / 8 + 33 * (c->ntiles + 1);
}
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

// i mod n in [0, n); one period either side without a division
__device__ __forceinline__ int pmx_wrap(int i, int n) {
  if ((unsigned)i < (unsigned)n) return i;
  if (i < 0 && i >= -n) return i + n;
  if (i >= n && i - n < n) return i - n;
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
  float d;             // disp / cell
};

// x / cell, as a product when inv_cell = 1/cell is exact (inv_cell != 0)
__device__ __forceinline__ float pmx_div_cell(float x, float cell, float inv_cell) {
  return inv_cell != 0.f ? x * inv_cell : x / cell;
}

__device__ __forceinline__ float pmx_sign(float x) {
  return x > 0.f ? 1.f : (x < 0.f ? -1.f : 0.f);
}

__device__ __forceinline__ CicAxis pmx_cic_axis(int pm, float dp, float cell, float inv_cell,
                                                int n) {
  CicAxis a;
  float d = pmx_div_cell(dp, cell, inv_cell);
  a.d = d;
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

// --------------------------------------------------- particle helpers --
__device__ __forceinline__ void load_ptcl(const int16_t* __restrict__ pmid,
                                          const float* __restrict__ disp, long long i,
                                          int pm[3], float dp[3]) {
  pm[0] = pmid[3 * i + 0];
  pm[1] = pmid[3 * i + 1];
  pm[2] = pmid[3 * i + 2];
  dp[0] = disp[3 * i + 0];
  dp[1] = disp[3 * i + 1];
  dp[2] = disp[3 * i + 2];
}

__device__ __forceinline__ void cic_setup(const int pm[3], const float dp[3], const MeshGeom& g,
                                          CicAxis ax[3]) {
  ax[0] = pmx_cic_axis(pm[0], dp[0], g.cell, g.inv_cell, g.n0);
  ax[1] = pmx_cic_axis(pm[1], dp[1], g.cell, g.inv_cell, g.n1);
  ax[2] = pmx_cic_axis(pm[2], dp[2], g.cell, g.inv_cell, g.n2);
}

__device__ __forceinline__ uint32_t tile_of(const CicAxis ax[3], const TileGeom& tg) {
  return ((uint32_t)pmx_tdiv(ax[0].lo, tg.mx) * tg.nty + (uint32_t)pmx_tdiv(ax[1].lo, tg.my)) *
             tg.ntz +
         (uint32_t)pmx_tdiv(ax[2].lo, tg.mz);
}

// Record of a binned particle: d = disp/cell per axis (float32, the value
// enmesh divides out) and the lower CIC corner relative to the tile origin
// packed as lx | ly << 10 | lz << 20 (bit pattern stored in .w).
__device__ __forceinline__ float4 pmx_make_record(const CicAxis ax[3], const TileGeom& tg) {
  const int lx = ax[0].lo - pmx_tdiv(ax[0].lo, tg.mx) * tg.tx;
  const int ly = ax[1].lo - pmx_tdiv(ax[1].lo, tg.my) * tg.ty;
  const int lz = ax[2].lo - pmx_tdiv(ax[2].lo, tg.mz) * tg.tz;
  const uint32_t packed = (uint32_t)lx | ((uint32_t)ly << 10) | ((uint32_t)lz << 20);
  return make_float4(ax[0].d, ax[1].d, ax[2].d, __uint_as_float(packed));
}

// Where a binning kernel puts particle i of tile t with rank r.  Exact mode
// (CAP false): the tile and rank, for the fill after the scan of the counts.
// Capacity mode: the record itself at start[t] + r when r is below the
// tile's capacity start[t+1] - start[t], else the overflow flag.
struct BinOut {
  uint32_t* ptile;
  uint32_t* prank;
  const uint32_t* start;
  float4* rec;
  uint32_t* pidx;
  uint32_t* flag;
};

template <bool CAP>
__device__ __forceinline__ void pmx_bin_emit(const BinOut& o, long long i, uint32_t t, uint32_t r,
                                             const CicAxis ax[3], const TileGeom& tg) {
  if (!CAP) {
    o.ptile[i] = t;
    o.prank[i] = r;
    return;
  }
  const uint32_t b = o.start[t], e = o.start[t + 1];
  if (r < e - b) {
    o.rec[b + r] = pmx_make_record(ax, tg);
    o.pidx[b + r] = (uint32_t)i;
  } else {
    o.flag[0] = 1u;
  }
}

// Wave-aggregated atomicAdd(&ctr[key], 1): returns this lane's slot.
// Every lane of the wave must call it (inactive lanes pass active=false).
// The lanes are grouped by key with ballots first; then each group's leader
// adds the group size, all leaders in one instruction, so the wave waits for
// a single atomic round trip however many keys it holds.
__device__ __forceinline__ uint32_t agg_slot(uint32_t* __restrict__ ctr, uint32_t key, bool active) {
  unsigned long long pending = __ballot(active);
  const int lane = __lane_id();
  int my_leader = lane;
  uint32_t my_rank = 0, group = 0;
  while (pending) {
    const int leader = __ffsll((long long)pending) - 1;  // wave-uniform
    const uint32_t lkey = (uint32_t)__builtin_amdgcn_readlane((int)key, leader);
    const unsigned long long peers = __ballot(active && key == lkey) & pending;
    if (active && key == lkey) {
      my_leader = leader;
      my_rank = (uint32_t)__popcll(peers & ((1ull << lane) - 1ull));
    }
    if (lane == leader) group = (uint32_t)__popcll(peers);
    pending &= ~peers;
  }
  uint32_t base = 0;
  if (group) base = atomicAdd(&ctr[key], group);
  base = __shfl(base, my_leader);
  return base + my_rank;
}

// ------------------------------------------------ AoS staging through LDS --
// A block of TPB consecutive particles holds 3*TPB floats of each (n,3)
// array.  These move whole blocks between HBM and LDS with 16-byte accesses
// (3*TPB/4 lanes) when the block is full and its rows are 16-byte aligned,
// else element by element; each thread then works on its particle's three
// values in LDS.
template <int TPB>
__device__ __forceinline__ bool pmx_block_aligned(const void* p) {
  return ((uintptr_t)p & 15) == 0;
}

// NS row pointers handed over by value (kept in registers)
template <int NS, class T>
struct PmxPtrs {
  T* p[NS];
};

template <int NS, int TPB>
__device__ __forceinline__ void pmx_stage_in3(float* sh, PmxPtrs<NS, const float> g, long long b0,
                                              int cnt) {
  bool vec = cnt == TPB;
#pragma unroll
  for (int k = 0; k < NS; ++k) vec = vec && pmx_block_aligned<TPB>(g.p[k] + 3 * b0);
  if (vec) {
    if (threadIdx.x < 3 * TPB / 4) {
      float4 r[NS];
#pragma unroll
      for (int k = 0; k < NS; ++k) r[k] = reinterpret_cast<const float4*>(g.p[k] + 3 * b0)[threadIdx.x];
#pragma unroll
      for (int k = 0; k < NS; ++k) reinterpret_cast<float4*>(sh + k * 3 * TPB)[threadIdx.x] = r[k];
    }
  } else {
    for (int i = threadIdx.x; i < 3 * cnt; i += TPB) {
#pragma unroll
      for (int k = 0; k < NS; ++k) sh[k * 3 * TPB + i] = g.p[k][3 * b0 + i];
    }
  }
}

template <int NS, int TPB>
__device__ __forceinline__ void pmx_stage_out3(PmxPtrs<NS, float> g, const float* sh, long long b0,
                                               int cnt) {
  bool vec = cnt == TPB;
#pragma unroll
  for (int k = 0; k < NS; ++k) vec = vec && pmx_block_aligned<TPB>(g.p[k] + 3 * b0);
  if (vec) {
    if (threadIdx.x < 3 * TPB / 4) {
#pragma unroll
      for (int k = 0; k < NS; ++k)
        reinterpret_cast<float4*>(g.p[k] + 3 * b0)[threadIdx.x] =
            reinterpret_cast<const float4*>(sh + k * 3 * TPB)[threadIdx.x];
    }
  } else {
    for (int i = threadIdx.x; i < 3 * cnt; i += TPB) {
#pragma unroll
      for (int k = 0; k < NS; ++k) g.p[k][3 * b0 + i] = sh[k * 3 * TPB + i];
    }
  }
}

// pmid (n,3) int16: 6*TPB bytes per block, 16-byte accesses by 3*TPB/8 lanes
template <int TPB>
__device__ __forceinline__ void pmx_stage_pmid(int16_t* sp, const int16_t* pmid, long long b0, int cnt) {
  const int16_t* src = pmid + 3 * b0;
  if (cnt == TPB && pmx_block_aligned<TPB>(src)) {
    if (threadIdx.x < 3 * TPB / 8)
      reinterpret_cast<float4*>(sp)[threadIdx.x] = reinterpret_cast<const float4*>(src)[threadIdx.x];
  } else {
    for (int i = threadIdx.x; i < 3 * cnt; i += TPB) sp[i] = src[i];
  }
}

__device__ __forceinline__ void tile_origin(int t, const TileGeom& tg, const MeshGeom& g, int o[3],
                                            int s[3], int ti[3]) {
  ti[2] = t % tg.ntz;
  ti[1] = (t / tg.ntz) % tg.nty;
  ti[0] = t / (tg.ntz * tg.nty);
  o[0] = ti[0] * tg.tx;
  o[1] = ti[1] * tg.ty;
  o[2] = ti[2] * tg.tz;
  s[0] = min(tg.tx, g.n0 - o[0]);
  s[1] = min(tg.ty, g.n1 - o[1]);
  s[2] = min(tg.tz, g.n2 - o[2]);
}

// The fold of 16^3 tiles at tile-local cell (lx, ly, lz) of tile (tx, ty,
// tz): the +1 shells of the lower neighbours that cover the cell (S = 1..7,
// bits 0/1/2 a step down in x/y/z, periodic), summed in that order from 0
// as fold_k sums them.  Shell layout of a 16^3 tile: the x-high face (17 x
// 17), the y-high face (16 x 17), then the z-high face (16 x 16).
__device__ __forceinline__ float pmx_fold_sum16(const float* __restrict__ halo, int shell, int ntx,
                                                int nty, int ntz, int tx, int ty, int tz, int lx,
                                                int ly, int lz) {
  float sum = 0.f;
#pragma unroll
  for (int S = 1; S < 8; ++S) {
    if ((S & 1) && lx != 0) continue;
    if ((S & 2) && ly != 0) continue;
    if ((S & 4) && lz != 0) continue;
    const int nx = (S & 1) ? (tx == 0 ? ntx - 1 : tx - 1) : tx;
    const int ny = (S & 2) ? (ty == 0 ? nty - 1 : ty - 1) : ty;
    const int nz = (S & 4) ? (tz == 0 ? ntz - 1 : tz - 1) : tz;
    const int qz = (S & 4) ? 16 : lz;
    const int q = (S & 1) ? ((S & 2) ? 16 : ly) * 17 + qz
                          : ((S & 2) ? 289 + lx * 17 + qz : 561 + lx * 16 + ly);
    sum += halo[(long long)((nx * nty + ny) * ntz + nz) * shell + q];
  }
  return sum;
}

// fractions and the sign(-d2) factors of their gradients from d = disp/cell,
// exactly as pmx_cic_axis forms them
__device__ __forceinline__ void frac_sign_of(float d, float& w0, float& w1, float& s0, float& s1) {
  int i = (int)floorf(d);
  float d20 = d - (float)i;
  float d21 = d - (float)(i + 1);
  w0 = 1.f - fabsf(d20);
  w1 = 1.f - fabsf(d21);
  s0 = pmx_sign(-d20);
  s1 = pmx_sign(-d21);
}

// Per-axis fractions of the nb = 0 / 1 neighbours, w[a][0] and w[a][1], and
// the sign(-d2) factors of their gradients, sg[a][*] (pmx_cic_axis).  Built
// from d (records) or taken from the axes cic_setup already formed, with
// the same values; fields a caller does not read are not computed.
struct CicFrac {
  float w[3][2];
  float sg[3][2];
};

__device__ __forceinline__ CicFrac pmx_frac_of_d(const float d[3]) {
  CicFrac f;
#pragma unroll
  for (int a = 0; a < 3; ++a) frac_sign_of(d[a], f.w[a][0], f.w[a][1], f.sg[a][0], f.sg[a][1]);
  return f;
}

__device__ __forceinline__ CicFrac pmx_frac_of_axes(const CicAxis ax[3]) {
  CicFrac f;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    f.w[a][0] = ax[a].w0;
    f.w[a][1] = ax[a].w1;
    f.sg[a][0] = ax[a].s0;
    f.sg[a][1] = ax[a].s1;
  }
  return f;
}

// Largest tile edge of the scatter binning, and the cells of a tile plus
// its +1 layer.
#define PMX_TILE_MAX 16
#define PMX_TILE_CELLS ((PMX_TILE_MAX + 1) * (PMX_TILE_MAX + 1) * (PMX_TILE_MAX + 1))

// Stages NC padded meshes' values over tile t's cells plus the +1 layer
// (periodic wrap) into LDS, component-major: sh[c*ltot + (qx*LY + qy)*LZ + qz].
// All of a thread's loads are issued before its LDS stores.
template <int NC, int TPB>
__device__ __forceinline__ void pmx_stage_tile(float* sh, const int o[3], int LY, int LZ, int ltot,
                                               const MeshGeom& g, const float* const* F) {
  constexpr int NI = (PMX_TILE_CELLS + TPB - 1) / TPB;
  // i / LZ and (i / LZ) / LY as 24-bit multiplies by ceil(2^20 / L): exact
  // for i < 2^16 and L <= 17 (the error i (M L - 2^20) / 2^20 / L stays
  // below 1 / L)
  const unsigned mz = ((1u << 20) + LZ - 1) / LZ, my = ((1u << 20) + LY - 1) / LY;
  float r[NI][NC];
#pragma unroll
  for (int it = 0; it < NI; ++it) {
    const int i = threadIdx.x + it * TPB;
    if (i < ltot) {
      const int iz = (int)(__umul24((unsigned)i, mz) >> 20);
      const int qx = (int)(__umul24((unsigned)iz, my) >> 20);
      const int qz = i - iz * LZ;
      const int qy = iz - qx * LY;
      int x = o[0] + qx, y = o[1] + qy, z = o[2] + qz;
      if (x >= g.n0) x -= g.n0;
      if (y >= g.n1) y -= g.n1;
      if (z >= g.n2) z -= g.n2;
      const long long m = (long long)(__umul24(x, g.n1) + y) * g.rs + z;
#pragma unroll
      for (int c = 0; c < NC; ++c) r[it][c] = F[c][m];
    }
  }
#pragma unroll
  for (int it = 0; it < NI; ++it) {
    const int i = threadIdx.x + it * TPB;
    if (i < ltot) {
#pragma unroll
      for (int c = 0; c < NC; ++c) sh[c * ltot + i] = r[it][c];
    }
  }
}

// acc = sum_j F[c_j] w_j over the eight corners of a binned particle (d =
// disp/cell per axis, lower corner (lx, ly, lz) in the staged tile of three
// channels ltot floats apart, rows LY x LZ), summed in the tree order of
// pmx_tree8 -- the arithmetic of a per-particle gather.
__device__ __forceinline__ void pmx_tile_acc(const float* sh, int ltot, int LY, int LZ,
                                             const CicFrac& fr, int lx, int ly, int lz,
                                             float out[3]) {
  const float a0 = fr.w[0][0], a1 = fr.w[0][1], b0 = fr.w[1][0], b1 = fr.w[1][1];
  const float c0 = fr.w[2][0], c1 = fr.w[2][1];
  const int b = __umul24(__umul24(lx, LY) + ly, LZ) + lz;
  float v0[8], v1[8], v2[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    float f0 = (j & 1) ? a1 : a0;
    float f1 = (j & 2) ? b1 : b0;
    float f2 = (j & 4) ? c1 : c0;
    float w = (f0 * f1) * f2;
    int li = b + ((j & 1) ? LY * LZ : 0) + ((j & 2) ? LZ : 0) + ((j & 4) ? 1 : 0);
    v0[j] = sh[li] * w;
    v1[j] = sh[ltot + li] * w;
    v2[j] = sh[2 * ltot + li] * w;
  }
  out[0] = pmx_tree8(v0);
  out[1] = pmx_tree8(v1);
  out[2] = pmx_tree8(v2);
}

// Force VJP of one binned particle over a staged tile of the three force
// meshes (gather.py:80-116 for three components): acc = sum_j F[c_j] w_j and
// acot = sum_j (v . F[c_j]) grad w_j / cell, v = vel_cot, each in the tree
// order of pmx_tree8.
__device__ __forceinline__ void pmx_vjp_one(const float* sh, int ltot, int LY, int LZ,
                                            const CicFrac& fr, const int l[3], const float v[3],
                                            float cell, float inv_cell, float acc[3],
                                            float acot[3]) {
  const float* wa = fr.w[0];
  const float* wb = fr.w[1];
  const float* wc = fr.w[2];
  const float* sa = fr.sg[0];
  const float* sb = fr.sg[1];
  const float* sc = fr.sg[2];
  const int b = __umul24(__umul24(l[0], LY) + l[1], LZ) + l[2];
  float r0[8], r1[8], r2[8], gx[8], gy[8], gz[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int ja = j & 1, jb = (j >> 1) & 1, jc = (j >> 2) & 1;
    const float f0 = wa[ja], f1 = wb[jb], f2 = wc[jc];
    const float w = (f0 * f1) * f2;
    const int li = b + (ja ? LY * LZ : 0) + (jb ? LZ : 0) + jc;
    const float a0 = sh[li], a1 = sh[ltot + li], a2 = sh[2 * ltot + li];
    r0[j] = a0 * w;
    r1[j] = a1 * w;
    r2[j] = a2 * w;
    const float dc = v[0] * a0 + v[1] * a1 + v[2] * a2;
    gx[j] = dc * (sa[ja] * (f1 * f2));
    gy[j] = dc * (sb[jb] * (f2 * f0));
    gz[j] = dc * (sc[jc] * (f0 * f1));
  }
  acc[0] = pmx_tree8(r0);
  acc[1] = pmx_tree8(r1);
  acc[2] = pmx_tree8(r2);
  acot[0] = pmx_div_cell(pmx_tree8(gx), cell, inv_cell);
  acot[1] = pmx_div_cell(pmx_tree8(gy), cell, inv_cell);
  acot[2] = pmx_div_cell(pmx_tree8(gz), cell, inv_cell);
}

// Scatter-adjoint disp cotangent of one binned particle over a staged tile
// of the density cotangent R (scatter.py:86-121, scalar val):
// sum_j (R[c_j] val) grad w_j / cell per axis, tree order.
__device__ __forceinline__ void pmx_grad_one(const float* sh, int LY, int LZ, const CicFrac& fr,
                                             const int l[3], float val, float cell,
                                             float inv_cell, float out[3]) {
  const float* wa = fr.w[0];
  const float* wb = fr.w[1];
  const float* wc = fr.w[2];
  const float* sa = fr.sg[0];
  const float* sb = fr.sg[1];
  const float* sc = fr.sg[2];
  const int b = __umul24(__umul24(l[0], LY) + l[1], LZ) + l[2];
  float gx[8], gy[8], gz[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int ja = j & 1, jb = (j >> 1) & 1, jc = (j >> 2) & 1;
    const float f0 = wa[ja], f1 = wb[jb], f2 = wc[jc];
    const int li = b + (ja ? LY * LZ : 0) + (jb ? LZ : 0) + jc;
    const float dc = sh[li] * val;
    gx[j] = dc * (sa[ja] * (f1 * f2));
    gy[j] = dc * (sb[jb] * (f2 * f0));
    gz[j] = dc * (sc[jc] * (f0 * f1));
  }
  out[0] = pmx_div_cell(pmx_tree8(gx), cell, inv_cell);
  out[1] = pmx_div_cell(pmx_tree8(gy), cell, inv_cell);
  out[2] = pmx_div_cell(pmx_tree8(gz), cell, inv_cell);
}

// Fixed-order block sum of s (wave butterflies, then wave sums in order)
// stored to part[blockIdx.x].
template <int TPB>
__device__ __forceinline__ void block_sum_store(double s, double* __restrict__ part) {
  __shared__ double wsum[TPB / 64];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  if (__lane_id() == 0) wsum[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0;
    for (int w = 0; w < TPB / 64; ++w) t += wsum[w];
    part[blockIdx.x] = t;
  }
}

// Two fixed-order block sums at once (one barrier), stored to
// pa[blockIdx.x] and pb[blockIdx.x].
template <int TPB>
__device__ __forceinline__ void block_sum2_store(double a, double b, double* __restrict__ pa,
                                                 double* __restrict__ pb) {
  __shared__ double wsum2[2][TPB / 64];
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_xor(a, off);
    b += __shfl_xor(b, off);
  }
  if (__lane_id() == 0) {
    wsum2[0][threadIdx.x >> 6] = a;
    wsum2[1][threadIdx.x >> 6] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double ta = 0, tb = 0;
    for (int w = 0; w < TPB / 64; ++w) {
      ta += wsum2[0][w];
      tb += wsum2[1][w];
    }
    pa[blockIdx.x] = ta;
    pb[blockIdx.x] = tb;
  }
}

// Running max of |v| per channel: wave max, then one atomic per wave only
// when it raises the stored maximum (non-negative floats order like their
// bit patterns).
__device__ __forceinline__ void wave_max_store(float m, float* vmax) {
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
  if (__lane_id() == 0) {
    unsigned* vm = reinterpret_cast<unsigned*>(vmax);
    if (__float_as_uint(m) > __atomic_load_n(vm, __ATOMIC_RELAXED)) atomicMax(vm, __float_as_uint(m));
  }
}

// ------------------------------------------------ tile-ordered particles --
__device__ __forceinline__ uint32_t pmx_state_pack(uint32_t px, uint32_t py, uint32_t pz) {
  return px | (py << 10) | (pz << 20);
}

__device__ __forceinline__ void pmx_state_pm(float4 s0, int pm[3]) {
  const uint32_t b = __float_as_uint(s0.w);
  pm[0] = (int)(b & 0x3ffu);
  pm[1] = (int)((b >> 10) & 0x3ffu);
  pm[2] = (int)(b >> 20);
}

// CIC fractions and the lower corner relative to tile origin o of a state
// entry (from the d = disp/cell and corner pmx_make_record packs); false
// when the corner is not in the tile of extent s.
__device__ __forceinline__ bool pmx_state_local(float4 s0, const MeshGeom& g, const int o[3],
                                                const int s[3], CicFrac& fr, int l[3]) {
  int pm[3];
  pmx_state_pm(s0, pm);
  const float dp[3] = {s0.x, s0.y, s0.z};
  CicAxis ax[3];
  cic_setup(pm, dp, g, ax);
  fr = pmx_frac_of_axes(ax);
#pragma unroll
  for (int a = 0; a < 3; ++a) l[a] = ax[a].lo - o[a];
  return (unsigned)l[0] < (unsigned)s[0] && (unsigned)l[1] < (unsigned)s[1] &&
         (unsigned)l[2] < (unsigned)s[2];
}

// Moves between the tile-ordered copies of the multi-step pipelines.  A
// workgroup owns one source tile; per round its particles bound for each of
// the 27 tiles around it (code (cx*3 + cy)*3 + cz, c = 0/1/2 for -1/0/+1
// with periodic wrap) are counted in LDS and take one contiguous run of the
// destination tile from a single global atomic; code 27 (farther) ranks per
// wave on the global counters.
__device__ __forceinline__ int pl_nbr_code(const CicAxis ax[3], const int ti[3],
                                           const TileGeom& tg) {
  const unsigned tm[3] = {tg.mx, tg.my, tg.mz};
  const int ntl[3] = {tg.ntx, tg.nty, tg.ntz};
  int cc = 0;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int dl = pmx_tdiv(ax[a].lo, tm[a]) - ti[a];
    int ca = 3;
    if (dl == 0) ca = 1;
    else if (dl == 1 || dl == 1 - ntl[a]) ca = 2;
    else if (dl == -1 || dl == ntl[a] - 1) ca = 0;
    cc = ca == 3 || cc == 27 ? 27 : cc * 3 + ca;
  }
  return cc;
}

// Threads 0..26: this round's LDS count of neighbour code t becomes a run
// base in count[] of that tile (one atomic), and the count is cleared.
// Call between two barriers.
__device__ __forceinline__ void pl_round_bases(uint32_t* s_cnt, uint32_t* s_base, const int ti[3],
                                               const TileGeom& tg, uint32_t* count) {
  if (threadIdx.x >= 27) return;
  const int ntl[3] = {tg.ntx, tg.nty, tg.ntz};
  const uint32_t k = s_cnt[threadIdx.x];
  uint32_t base = 0;
  if (k) {
    int c3 = threadIdx.x, nb[3];
#pragma unroll
    for (int a = 2; a >= 0; --a) {
      const int v = ti[a] + c3 % 3 - 1;
      nb[a] = v < 0 ? v + ntl[a] : (v >= ntl[a] ? v - ntl[a] : v);
      c3 /= 3;
    }
    const uint32_t tile = ((uint32_t)nb[0] * tg.nty + (uint32_t)nb[1]) * tg.ntz + (uint32_t)nb[2];
    base = atomicAdd(&count[tile], k);
  }
  s_base[threadIdx.x] = base;
  s_cnt[threadIdx.x] = 0u;
}

// Unpacks a binned record's tile-local lower corner (lx | ly << 10 | lz << 20).
__device__ __forceinline__ void pmx_unpack_local(float w, int& lx, int& ly, int& lz) {
  uint32_t pk = __float_as_uint(w);
  lx = pk & 1023;
  ly = (pk >> 10) & 1023;
  lz = (pk >> 20) & 1023;
}

// ------------------------------------------------- shared host drivers --
// with cond, the fill, scan and binning kernels do nothing unless *cond != 0
int pmx_fill_u32(pmx_ctx* c, uint32_t* p, long long n, uint32_t v, const uint32_t* cond = nullptr);
int pmx_scan_tiles(pmx_ctx* c, const uint32_t* cond = nullptr);
int pmx_bin_fill(pmx_ctx* c, const int16_t* pmid, const float* disp, long long n,
                 const uint32_t* cond = nullptr);
int pmx_bin_particles(pmx_ctx* c, const int16_t* pmid, const float* disp, long long n);
// A binning pass in three parts: pmx_bin_begin zeroes the counts and, when
// the counts of the last pass can size the tiles, sets up the capacity
// layout (returns *cap = 1); the caller's binning kernel then counts the
// particles and emits them (pmx_bin_out gives its BinOut); pmx_bin_end
// finishes the exact path, or re-bins exactly on the device when a tile
// overflowed, and records the counts for the next pass.
int pmx_bin_begin(pmx_ctx* c, long long n, int* cap);
BinOut pmx_bin_out(pmx_ctx* c);
int pmx_bin_end(pmx_ctx* c, const int16_t* pmid, const float* disp, long long n, int cap);
// The r2c row pass of the Poisson solve can add the deposit's halo shells as
// it loads the density (16^3 tiles covering the mesh, a fixed-length
// half-size row pass whose row blocks stay in one x plane).  Deposits made
// with defer = 1 then leave the fold to it, and the solve is told so with
// deferred = 1; otherwise both run as before.
int pmx_fold_deferrable(const pmx_ctx* c);
int pmx_deposit_binned(pmx_ctx* c, float val, long long n, float* mesh, int mrs, int accumulate,
                       int defer = 0);
int pmx_deposit_state(pmx_ctx* c, const PState& st, const uint32_t* start, const uint32_t* count,
                      float val, float* mesh, int mrs, int defer = 0);
// Output meshes of a deposit (one per channel).
struct DepOut {
  float* mesh[3];
};
int pmx_deposit_state_vals(pmx_ctx* c, const PState& st, const uint32_t* start,
                           const uint32_t* count, float val, const ValSrc* vals, int nch,
                           DepOut out, int mrs, int defer = 0);
// kick (k1) and drift (d) of the caller's arrays, binned into count/o (cap:
// capacity mode)
int pmx_kick_drift_bin(pmx_ctx* c, const int16_t* pmid, float* disp, float* vel, const float* acc,
                       long long n, float k1, float d, uint32_t* count, const BinOut& o, int cap);
int pmx_ensure_state(pmx_ctx* c);
int pmx_ensure_adj_state(pmx_ctx* c, int nsteps);
// reverse kick(s) and drift of the caller's arrays (pending closing kick kp
// first when nonzero), binned into count/o; per-block partials of the drift
// dot product at part[0, pmx_blocks(n))
int pmx_kd_adj_bin(pmx_ctx* c, const int16_t* pmid, float* disp, float* vel, const float* acc,
                   float* disp_cot, float* vel_cot, const float* acc_cot, const float* disp_snap,
                   long long n, float kp, float ka, float d, uint32_t* count, const BinOut& o,
                   int cap, double* part);
int pmx_deposit_binned3(pmx_ctx* c, long long n, float* const* mesh, int mrs, int defer = 0);
int pmx_ensure_halo(pmx_ctx* c, int nch);
int pmx_poisson_ctx(pmx_ctx* c, float coef, int deferred = 0);
// FFT passes on meshes in the context's padded layout (pmx_fft.hip):
// in-place z r2c (pre: (x - 1) * coef on load; halo: the deposit's fold
// added on load), z c2r with a final scale, a strided complex FFT along
// axis 1 (y) or 0 (x) of nf fields (dir -1 forward, +1 inverse), the paired
// y pass of the gradient's y and z components, and the Poisson x pass (kin:
// f[0] holds a full 3D spectrum, no forward x FFT).
// Every tile a full 16^3 tile: kernels built for that case (their FULL)
// take the tile's LDS strides as constants.
inline bool pmx_full16(const pmx_ctx* c) {
  const TileGeom& t = c->tg;
  return t.tx == 16 && t.ty == 16 && t.tz == 16 && c->n[0] % 16 == 0 && c->n[1] % 16 == 0 &&
         c->n[2] % 16 == 0;
}

// rev: the pass runs its blocks in reverse order (bits unchanged)
int pmx_fft_r2c_inplace(pmx_ctx* c, float* buf, int pre, float coef, const float* halo = nullptr,
                        int rev = 0);
int pmx_fft_c2r_inplace(pmx_ctx* c, float* buf, float scale, int rev = 0);
int pmx_fft_strided(pmx_ctx* c, float2* const* f, int nf, int axis, int dir, int rev = 0);
int pmx_fft_ypair(pmx_ctx* c, float2* const* f, int dir, int rev = 0);
int pmx_poisson_x(pmx_ctx* c, float2* const* f, int adjoint, int kin = 0, int rev = 0);
// copies n0*n1 rows of ncols floats between row strides (ncols <= 0: n2)
int pmx_copy_rows(pmx_ctx* c, float* dst, int dst_rs, const float* src, int src_rs, int ncols = 0);
int pmx_poisson_adj_ctx(pmx_ctx* c, float scale_out, int deferred = 0);
int pmx_poisson_adj_bufs(pmx_ctx* c, float2* const* f, float scale_out, int deferred = 0);
// bins the particles, deposits the density and solves for the three force
// meshes in c->cbuf (padded); then the tile-staged gather of those meshes
int pmx_gravity_fwd_meshes(pmx_ctx* c, const int16_t* pmid, const float* disp, long long n,
                           float val, float coef);
int pmx_tile_gather(pmx_ctx* c, float* acc, float* vel, float k2);
// float64 partial sums: reduce red[off, off+nblk) in a fixed order into
// device slot `slot` (< PMX_RED_SLOTS), then fetch slots [0, k) to the host
#define PMX_RED_SLOTS 8
int pmx_red_reduce(pmx_ctx* c, long long off, long long nblk, int slot);
// the same into a device double
int pmx_red_reduce_to(pmx_ctx* c, long long off, long long nblk, double* dst);
int pmx_red_fetch(pmx_ctx* c, int k, double* out);
int pmx_red_finish(pmx_ctx* c, long long nblk, double* out);
// grid of the exact re-binning launched after a capacity pass
#define PMX_REBIN_GRID 4096u
// blocks of a capped grid-stride launch over n items
static inline unsigned pmx_blocks_capped(long long n, unsigned cap, int tpb = PMX_TPB) {
  unsigned b = pmx_blocks(n, tpb);
  return b < cap ? b : cap;
}
