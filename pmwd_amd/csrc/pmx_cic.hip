// Cloud-in-cell scatter/gather, gravity and time stepping (gfx950).
//
// Hot scatter (pmx_scatter_mass): particles are binned by the mesh tile that
// holds their lower CIC corner (count with wave-aggregated integer atomics,
// exclusive scan, fill), then one workgroup per tile accumulates its
// particles into an LDS copy of the tile plus a +1 halo shell and writes the
// tile interior with plain stores and the shell to a halo buffer; a fold
// pass adds each tile's incoming shells.  No float atomics touch global
// memory on this path.  The tile-staged gather and the forward KDK step
// live here too; the reverse-time adjoint is in pmx_adjoint.hip.
#include "pmx_internal.h"

#include <math.h>

#include <type_traits>

__global__ void fill_u32_k(uint32_t* __restrict__ p, long long n, uint32_t v,
                           const uint32_t* __restrict__ cond) {
  if (cond && *cond == 0) return;
  for (long long i = (long long)blockIdx.x * PMX_TPB + threadIdx.x; i < n;
       i += (long long)gridDim.x * PMX_TPB)
    p[i] = v;
}

int pmx_fill_u32(pmx_ctx* c, uint32_t* p, long long n, uint32_t v, const uint32_t* cond) {
  unsigned nb = pmx_blocks(n);
  if (nb > 4096) nb = 4096;
  pmx_tic(c, PMX_T_OTHER);
  fill_u32_k<<<nb, PMX_TPB, 0, c->stream>>>(p, n, v, cond);
  pmx_toc(c, PMX_T_OTHER);
  PMX_LAUNCH_CHECK();
  return PMX_OK;
}

// ------------------------------------------------------------- keys --
__global__ void cell_keys_k(const int16_t* __restrict__ pmid, const float* __restrict__ disp,
                            long long n, MeshGeom g, uint32_t* __restrict__ key) {
  long long i = (long long)blockIdx.x * PMX_TPB + threadIdx.x;
  if (i >= n) return;
  int pm[3];
  float dp[3];
  load_ptcl(pmid, disp, i, pm, dp);
  CicAxis ax[3];
  cic_setup(pm, dp, g, ax);
  key[i] = ((uint32_t)ax[0].lo * g.n1 + ax[1].lo) * g.n2 + ax[2].lo;
}

__global__ void tile_keys_k(const int16_t* __restrict__ pmid, const float* __restrict__ disp,
                            long long n, MeshGeom g, TileGeom tg, uint32_t* __restrict__ tile) {
  long long i = (long long)blockIdx.x * PMX_TPB + threadIdx.x;
  if (i >= n) return;
  int pm[3];
  float dp[3];
  load_ptcl(pmid, disp, i, pm, dp);
  CicAxis ax[3];
  cic_setup(pm, dp, g, ax);
  tile[i] = tile_of(ax, tg);
}

extern "C" int pmx_cell_keys(pmx_ctx* c, const int16_t* pmid, const float* disp, long long n,
                             uint32_t* key) {
  PMX_ENTRY(c);
  PMX_REQUIRE(c && n >= 0, "pmx_cell_keys: bad argument");
  if (n == 0) return PMX_OK;
  PMX_REQUIRE(pmid && disp && key, "pmx_cell_keys: NULL pointer");
  MeshGeom g = pmx_geom_padded(c);
  g.rs = c->n[2];
  cell_keys_k<<<pmx_blocks(n), PMX_TPB, 0, c->stream>>>(pmid, disp, n, g, key);
  PMX_LAUNCH_CHECK();
  return PMX_OK;
}

extern "C" int pmx_tile_keys(pmx_ctx* c, const int16_t* pmid, const float* disp, long long n,
                             uint32_t* tile) {
  PMX_ENTRY(c);
  PMX_REQUIRE(c && n >= 0, "pmx_tile_keys: bad argument");
  if (n == 0) return PMX_OK;
  PMX_REQUIRE(pmid && disp && tile, "pmx_tile_keys: NULL pointer");
  tile_keys_k<<<pmx_blocks(n), PMX_TPB, 0, c->stream>>>(pmid, disp, n, pmx_geom_padded(c), c->tg,
                                                        tile);
  PMX_LAUNCH_CHECK();
  return PMX_OK;
}

// ---------------------------------------------------- binned scatter --
// The binning kernels below run over blocks of PMX_TPB particles, grid-
// striding when the grid is capped (the exact re-binning after a capacity
// overflow, launched before the device knows whether it is needed, exits at
// once unless *cond != 0).
template <bool CAP>
__global__ void __launch_bounds__(PMX_TPB)
    bin_count_k(const int16_t* __restrict__ pmid, const float* __restrict__ disp, long long n,
                MeshGeom g, TileGeom tg, uint32_t* __restrict__ tile_count, BinOut o,
                const uint32_t* __restrict__ cond) {
  if (cond && *cond == 0) return;
  __shared__ __align__(16) float sd[3 * PMX_TPB];
  __shared__ __align__(16) int16_t sp[3 * PMX_TPB];
  for (long long b0 = (long long)blockIdx.x * PMX_TPB; b0 < n;
       b0 += (long long)gridDim.x * PMX_TPB) {
    const int cnt = (int)(n - b0 < PMX_TPB ? n - b0 : PMX_TPB);
    pmx_stage_in3<1, PMX_TPB>(sd, {{disp}}, b0, cnt);
    pmx_stage_pmid<PMX_TPB>(sp, pmid, b0, cnt);
    __syncthreads();
    const int q = threadIdx.x;
    const bool active = q < cnt;
    uint32_t t = 0;
    CicAxis ax[3];
    float dp[3];
    if (active) {
      const int pm[3] = {sp[3 * q], sp[3 * q + 1], sp[3 * q + 2]};
      dp[0] = sd[3 * q];
      dp[1] = sd[3 * q + 1];
      dp[2] = sd[3 * q + 2];
      cic_setup(pm, dp, g, ax);
      t = tile_of(ax, tg);
    }
    const uint32_t r = agg_slot(tile_count, t, active);
    if (active) pmx_bin_emit<CAP>(o, b0 + q, t, r, ax, tg);
    __syncthreads();
  }
}

// Writes the record (pmx_make_record) and index of each particle at
// tile_start[tile] + rank once the tile counts are scanned.
__global__ void __launch_bounds__(PMX_TPB)
    bin_fill_k(const int16_t* __restrict__ pmid, const float* __restrict__ disp, long long n,
               MeshGeom g, TileGeom tg, const uint32_t* __restrict__ ptile,
               const uint32_t* __restrict__ prank, const uint32_t* __restrict__ tile_start,
               float4* __restrict__ rec, uint32_t* __restrict__ pidx,
               const uint32_t* __restrict__ cond) {
  if (cond && *cond == 0) return;
  __shared__ __align__(16) float sd[3 * PMX_TPB];
  __shared__ __align__(16) int16_t sp[3 * PMX_TPB];
  for (long long b0 = (long long)blockIdx.x * PMX_TPB; b0 < n;
       b0 += (long long)gridDim.x * PMX_TPB) {
    const int cnt = (int)(n - b0 < PMX_TPB ? n - b0 : PMX_TPB);
    pmx_stage_in3<1, PMX_TPB>(sd, {{disp}}, b0, cnt);
    pmx_stage_pmid<PMX_TPB>(sp, pmid, b0, cnt);
    __syncthreads();
    const int q = threadIdx.x;
    const long long i = b0 + q;
    if (q < cnt) {
      const int pm[3] = {sp[3 * q], sp[3 * q + 1], sp[3 * q + 2]};
      const float dp[3] = {sd[3 * q], sd[3 * q + 1], sd[3 * q + 2]};
      CicAxis ax[3];
      cic_setup(pm, dp, g, ax);
      const uint32_t slot = tile_start[ptile[i]] + prank[i];
      rec[slot] = pmx_make_record(ax, tg);
      pidx[slot] = (uint32_t)i;
    }
    __syncthreads();
  }
}

__device__ __forceinline__ int shell_index(int qx, int qy, int qz, int sx, int sy, int sz) {
  if (qx == sx) return qy * (sz + 1) + qz;
  int offA = (sy + 1) * (sz + 1);
  if (qy == sy) return offA + qx * (sz + 1) + qz;
  return offA + sx * (sz + 1) + qx * sy + qy;
}

// Fixed-point scale for tile accumulation: 2^k with |v| * 2^k < 2^30, so a
// sum of up to 2^33 contributions fits a signed 64-bit integer; resolution
// is 2^-30 of the largest |v|.  LDS float atomics serialise on gfx950 while
// 64-bit integer LDS atomics run at full rate, and integer sums make the
// accumulation order-independent (bitwise deterministic).
__device__ __forceinline__ double fixed_scale(float vmax) {
  int e = 0;
  if (vmax > 0.f) frexpf(vmax, &e);  // vmax < 2^e
  int k = 30 - e;
  if (k > 300) k = 300;
  return ldexp(1.0, k);
}


// One workgroup per tile: LDS accumulation of val*w (or vrec[ch][slot]*w)
// of the tile's binned records over the (sx+1)(sy+1)(sz+1) cells they touch;
// each thread streams DEP_U coalesced records per round.  The interior is
// stored with plain stores, the +1 shell to the tile's halo slot.  With
// nch > 1 the channels run one after the other in the same workgroup, the
// records of the tile coming back from L2.  STATE reads tile-ordered
// particles (rec = their s0, pm packed in s0.w) and forms d and the corner
// itself.
#define DEP_TPB 512
#ifndef DEP_U
#define DEP_U 1
#endif
// eight waves per SIMD (at most 64 VGPRs): four 39 KB workgroups per CU
#ifndef DEP_WPE
#define DEP_WPE 8
#endif

// Adds v*w_j for the eight corners of a particle (d = disp/cell, lower corner
// at LDS cell b) to the fixed-point tile accumulator.  The scale is a power
// of two; up to 2^126 (sf = scale as a float, fq) x * sf is exact in float
// and |x * sf| < 2^30, so rintf and an int32 conversion give the value
// rint((double)x * scale) gives, in fewer instructions.
template <bool FQ>
__device__ __forceinline__ void dep_one(unsigned long long* acc, const CicFrac& fr, int b, int LY,
                                        int LZ, float vv, double scale, float sf) {
  const float a0 = fr.w[0][0], a1 = fr.w[0][1], b0 = fr.w[1][0], b1 = fr.w[1][1];
  const float c0 = fr.w[2][0], c1 = fr.w[2][1];
  const int sxy = __umul24(LY, LZ);
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    float f0 = (j & 1) ? a1 : a0;
    float f1 = (j & 2) ? b1 : b0;
    float f2 = (j & 4) ? c1 : c0;
    int li = b + ((j & 1) ? sxy : 0) + ((j & 2) ? LZ : 0) + ((j & 4) ? 1 : 0);
    float x = vv * ((f0 * f1) * f2);
    long long qx = FQ ? (long long)(int)rintf(x * sf) : (long long)rint((double)x * scale);
    atomicAdd(&acc[li], (unsigned long long)qx);
  }
}

// The same for non-negative contributions (a non-negative constant value,
// the density): the tile's 64-bit sums as two 32-bit words, lo[] taking
// every contribution (< 2^30) and hi[] one per carry out of lo, so each add
// touches one LDS bank instead of two.  hi * 2^32 + lo is the 64-bit sum.
// PRE: vv is the value times sf already (a constant value of at least 2^-50,
// so that vv * w stays a normal float and the product with the power of
// two sf is the same float either way).
template <bool PRE = false>
__device__ __forceinline__ void dep_one_u32(unsigned* lo, unsigned* hi, const CicFrac& fr, int b,
                                            int LY, int LZ, float vv, float sf) {
  const float a0 = fr.w[0][0], a1 = fr.w[0][1], b0 = fr.w[1][0], b1 = fr.w[1][1];
  const float c0 = fr.w[2][0], c1 = fr.w[2][1];
  const int sxy = __umul24(LY, LZ);
  unsigned q[8], old[8];
  int li[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    float f0 = (j & 1) ? a1 : a0;
    float f1 = (j & 2) ? b1 : b0;
    float f2 = (j & 4) ? c1 : c0;
    li[j] = b + ((j & 1) ? sxy : 0) + ((j & 2) ? LZ : 0) + ((j & 4) ? 1 : 0);
    float x = vv * ((f0 * f1) * f2);
    q[j] = (unsigned)(int)rintf(PRE ? x : x * sf);
    old[j] = atomicAdd(&lo[li[j]], q[j]);
  }
#pragma unroll
  for (int j = 0; j < 8; ++j)
    if (old[j] + q[j] < old[j]) atomicAdd(&hi[li[j]], 1u);
}

// Converts the accumulated tile back to float: the interior to the mesh
// (row stride mrs), the +1 shell to the tile's halo slot hs.
// sum(i): the 64-bit fixed-point sum of LDS cell i
template <int TPB, class Sum>
__device__ __forceinline__ void dep_store(Sum sum, float* mesh, float* hs,
                                          const int o[3], const int s[3], int LY, int LZ,
                                          const MeshGeom& g, int mrs, int accumulate,
                                          double inv) {
  const int itot = s[0] * s[1] * s[2];
  if (s[0] == 16 && s[1] == 16 && s[2] == 16) {
    // a full tile: shifts in place of the divisions
    float* base = mesh + ((long long)o[0] * g.n1 + o[1]) * mrs + o[2];
    for (int i = threadIdx.x; i < 4096; i += TPB) {
      const int lz = i & 15, ly = (i >> 4) & 15, lx = i >> 8;
      float* p = base + ((long long)lx * g.n1 + ly) * mrs + lz;
      float a = (float)((double)sum((lx * 17 + ly) * 17 + lz) * inv);
      *p = accumulate ? *p + a : a;
    }
  } else {
    for (int i = threadIdx.x; i < itot; i += TPB) {
      int lz = i % s[2];
      int ly = (i / s[2]) % s[1];
      int lx = i / (s[2] * s[1]);
      long long m = ((long long)(o[0] + lx) * g.n1 + (o[1] + ly)) * mrs + (o[2] + lz);
      float a = (float)((double)sum((lx * LY + ly) * LZ + lz) * inv);
      mesh[m] = accumulate ? mesh[m] + a : a;
    }
  }
  const int offA = LY * LZ;
  const int offB = s[0] * LZ;
  const int stot = offA + offB + s[0] * s[1];
  for (int i = threadIdx.x; i < stot; i += TPB) {
    int qx, qy, qz;
    if (i < offA) {
      qx = s[0]; qy = i / LZ; qz = i % LZ;
    } else if (i < offA + offB) {
      int k = i - offA;
      qx = k / LZ; qy = s[1]; qz = k % LZ;
    } else {
      int k = i - offA - offB;
      qx = k / s[1]; qy = k % s[1]; qz = s[2];
    }
    hs[i] = (float)((double)sum((qx * LY + qy) * LZ + qz) * inv);
  }
}

// FULL: every tile is a full 16^3 tile, so the LDS strides are constants
// (corner offsets become immediates).  SPLIT: one channel per workgroup
// (nch * ntiles workgroups, ntiles a multiple of 8): the nch workgroups of a
// tile take consecutive slots of one XCD, so they run together and read the
// tile's records through the same L2, instead of one workgroup reading them
// nch times.
template <int PERVAL, int STATE = 0, int FULL = 0, int SPLIT = 0>
__global__ void __launch_bounds__(DEP_TPB) __attribute__((amdgpu_waves_per_eu(DEP_WPE)))
    deposit_k(const float4* __restrict__ rec, ValSrc vals, const float* __restrict__ vmaxp, const uint32_t* __restrict__ tile_start,
              const uint32_t* __restrict__ tile_count, MeshGeom g, TileGeom tg, float val,
              DepOut out, int nch, int mrs, int accumulate, float* __restrict__ halo,
              long long hstride) {
  extern __shared__ unsigned long long acc[];
  int o[3], s[3], ti[3];
  int t, ch0 = 0, ch1 = nch;
  if (SPLIT) {
    const unsigned x = blockIdx.x % PMX_NXCD, sl = blockIdx.x / PMX_NXCD;
    const unsigned j = sl / (unsigned)nch;
    ch0 = (int)(sl - j * (unsigned)nch);
    ch1 = ch0 + 1;
    t = (int)pmx_xcd_block(j * PMX_NXCD + x, gridDim.x / (unsigned)nch);
  } else {
    t = (int)pmx_xcd_block(blockIdx.x, gridDim.x);
  }
  tile_origin(t, tg, g, o, s, ti);
  if (FULL) s[0] = s[1] = s[2] = 16;
  const int LY = s[1] + 1, LZ = s[2] + 1;
  const int ltot = (s[0] + 1) * LY * LZ;
  const uint32_t start = tile_start[t];
  const uint32_t cnt = tile_count[t];
  for (int ch = ch0; ch < ch1; ++ch) {
    for (int i = threadIdx.x; i < ltot; i += DEP_TPB) acc[i] = 0ull;
    const double scale = fixed_scale(PERVAL ? vmaxp[ch] : fabsf(val));
    const double inv = 1.0 / scale;
    const bool fq = scale <= 0x1p126;
    const float sf = fq ? (float)scale : 0.f;
    const float* vr = PERVAL ? vals.p[ch] : nullptr;
    const int vstep = PERVAL ? vals.step[ch] : 0;
    __syncthreads();
    // the particle rounds, with the float or the double fixed-point conversion
    // (the next round's records are loaded before this round's arithmetic,
    // so they are in flight meanwhile)
    auto rounds = [&](auto fqc, auto u32c, auto prec) {
    constexpr bool FQ = decltype(fqc)::value;
    constexpr bool U32 = decltype(u32c)::value;
    constexpr bool PRE = decltype(prec)::value;
    float4 rn[DEP_U];
    float vn[DEP_U];
    auto fetch = [&](uint32_t q0, float4* r, float* v) {
#pragma unroll
      for (int u = 0; u < DEP_U; ++u) {
        uint32_t q = q0 + u * DEP_TPB + threadIdx.x;
        if (q < cnt) {
          r[u] = rec[start + q];
          v[u] = PERVAL ? vr[(long long)(start + q) * vstep] : val;
        }
      }
    };
    fetch(0, rn, vn);
    for (uint32_t q0 = 0; q0 < cnt; q0 += DEP_TPB * DEP_U) {
      float4 r[DEP_U];
      float v[DEP_U];
#pragma unroll
      for (int u = 0; u < DEP_U; ++u) {
        r[u] = rn[u];
        v[u] = vn[u];
      }
      if (q0 + DEP_TPB * DEP_U < cnt) fetch(q0 + DEP_TPB * DEP_U, rn, vn);
#pragma unroll
      for (int u = 0; u < DEP_U; ++u) {
        uint32_t q = q0 + u * DEP_TPB + threadIdx.x;
        if (q >= cnt) continue;
        int lx, ly, lz;
        CicFrac fr;
        if (STATE) {
          int l[3];
          // binned by this tile, so always inside; the guard keeps LDS in bounds
          if (!pmx_state_local(r[u], g, o, s, fr, l)) continue;
          lx = l[0];
          ly = l[1];
          lz = l[2];
        } else {
          pmx_unpack_local(r[u].w, lx, ly, lz);
          if (lx >= s[0] || ly >= s[1] || lz >= s[2]) continue;
          const float d[3] = {r[u].x, r[u].y, r[u].z};
          fr = pmx_frac_of_d(d);
        }
        const int b = __umul24(__umul24(lx, LY) + ly, LZ) + lz;
        if constexpr (U32)
          dep_one_u32<PRE>(reinterpret_cast<unsigned*>(acc), reinterpret_cast<unsigned*>(acc) + ltot,
                           fr, b, LY, LZ, PRE ? v[u] * sf : v[u], sf);
        else
          dep_one<FQ>(acc, fr, b, LY, LZ, v[u], scale, sf);
      }
    }
    };
    // split words for a non-negative constant value (the density)
    const bool u32 = !PERVAL && fq && val >= 0.f;
    if (u32 && val >= 0x1p-50f)
      rounds(std::true_type{}, std::true_type{}, std::true_type{});
    else if (u32)
      rounds(std::true_type{}, std::true_type{}, std::false_type{});
    else if (fq)
      rounds(std::true_type{}, std::false_type{}, std::false_type{});
    else
      rounds(std::false_type{}, std::false_type{}, std::false_type{});
    __syncthreads();
    float* mesh = ch == 0 ? out.mesh[0] : (ch == 1 ? out.mesh[1] : out.mesh[2]);
    float* hs = halo + ch * hstride + (long long)t * tg.shell;
    if (u32) {
      const unsigned* lo = reinterpret_cast<const unsigned*>(acc);
      const unsigned* hi = lo + ltot;
      dep_store<DEP_TPB>([&](int i) { return (long long)(((unsigned long long)hi[i] << 32) | lo[i]); },
                         mesh, hs, o, s, LY, LZ, g, mrs, accumulate, inv);
    } else {
      dep_store<DEP_TPB>([&](int i) { return (long long)acc[i]; }, mesh, hs, o, s, LY, LZ, g,
                         mrs, accumulate, inv);
    }
    __syncthreads();  // acc is cleared for the next channel
  }
}

// One workgroup per tile: the tile's three force meshes plus the +1 layer
// (periodic) are staged in LDS, then each binned record gathers its eight
// neighbours from LDS: acc = sum_j F[c_j] w_j in the tree order of
// pmx_tree8, written to the particle's index; KICK also closes the step,
// vel += acc * k2.  Same arithmetic as a per-particle gather.
#define TG_TPB 512
#define TG_U 4
template <int KICK>
__global__ void __launch_bounds__(TG_TPB)
    tile_gather_k(const float4* __restrict__ rec, const uint32_t* __restrict__ pidx,
                  const uint32_t* __restrict__ tile_start, const uint32_t* __restrict__ tile_count,
                  MeshGeom g, TileGeom tg, const float* __restrict__ F0,
                  const float* __restrict__ F1, const float* __restrict__ F2,
                  float* __restrict__ acc, float* __restrict__ vel, float k2) {
  extern __shared__ float sh[];
  int o[3], s[3], ti[3];
  const int t = (int)pmx_xcd_block(blockIdx.x, gridDim.x);
  tile_origin(t, tg, g, o, s, ti);
  const int LY = s[1] + 1, LZ = s[2] + 1;
  const int ltot = (s[0] + 1) * LY * LZ;
  const float* F[3] = {F0, F1, F2};
  pmx_stage_tile<3, TG_TPB>(sh, o, LY, LZ, ltot, g, F);
  __syncthreads();
  const uint32_t start = tile_start[t];
  const uint32_t cnt = tile_count[t];
  // TG_U records per thread per round: their loads (record, index, then
  // velocity) are all issued before the arithmetic
  for (uint32_t q0 = 0; q0 < cnt; q0 += TG_TPB * TG_U) {
    float4 r[TG_U];
    long long pp[TG_U];
    float vv[TG_U][3];
#pragma unroll
    for (int u = 0; u < TG_U; ++u) {
      const uint32_t q = q0 + u * TG_TPB + threadIdx.x;
      pp[u] = -1;
      if (q < cnt) {
        r[u] = rec[start + q];
        pp[u] = pidx[start + q];
      }
    }
    if (KICK) {
#pragma unroll
      for (int u = 0; u < TG_U; ++u)
        if (pp[u] >= 0) {
#pragma unroll
          for (int a = 0; a < 3; ++a) vv[u][a] = vel[3 * pp[u] + a];
        }
    }
#pragma unroll
    for (int u = 0; u < TG_U; ++u) {
      if (pp[u] < 0) continue;
      const long long p = pp[u];
      int lx, ly, lz;
      pmx_unpack_local(r[u].w, lx, ly, lz);
      if (lx >= s[0] || ly >= s[1] || lz >= s[2]) continue;
      float o3[3];
      const float d[3] = {r[u].x, r[u].y, r[u].z};
      pmx_tile_acc(sh, ltot, LY, LZ, pmx_frac_of_d(d), lx, ly, lz, o3);
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        acc[3 * p + a] = o3[a];
        if (KICK) vel[3 * p + a] = vv[u][a] + o3[a] * k2;
      }
    }
  }
}

// Adds to every low-face cell of tile t the shells of the (up to seven)
// lower neighbour tiles that cover it, in a fixed order; each of nch
// channels has its mesh and its halo slice (hstride floats apart).
__global__ void __launch_bounds__(PMX_TPB)
    fold_k(DepOut out, int nch, int mrs, const float* __restrict__ halo_all, long long hstride,
           MeshGeom g, TileGeom tg) {
  int o[3], s[3], ti[3];
  const int t = blockIdx.x;
  tile_origin(t, tg, g, o, s, ti);
  const int nA = s[1] * s[2];
  const int nB = (s[0] - 1) * s[2];
  const int nC = (s[0] - 1) * (s[1] - 1);
  const int ntot = nA + nB + nC;
  for (int ch = 0; ch < nch; ++ch)
  for (int i = threadIdx.x; i < ntot; i += PMX_TPB) {
    float* mesh = ch == 0 ? out.mesh[0] : (ch == 1 ? out.mesh[1] : out.mesh[2]);
    const float* halo = halo_all + ch * hstride;
    int l[3];
    if (i < nA) {
      l[0] = 0; l[1] = i / s[2]; l[2] = i % s[2];
    } else if (i < nA + nB) {
      int k = i - nA;
      l[0] = 1 + k / s[2]; l[1] = 0; l[2] = k % s[2];
    } else {
      int k = i - nA - nB;
      l[0] = 1 + k / (s[1] - 1); l[1] = 1 + k % (s[1] - 1); l[2] = 0;
    }
    float sum = 0.f;
#pragma unroll
    for (int S = 1; S < 8; ++S) {
      if ((S & 1) && l[0] != 0) continue;
      if ((S & 2) && l[1] != 0) continue;
      if ((S & 4) && l[2] != 0) continue;
      int nt[3], ns[3], q[3];
      nt[0] = (S & 1) ? (ti[0] == 0 ? tg.ntx - 1 : ti[0] - 1) : ti[0];
      nt[1] = (S & 2) ? (ti[1] == 0 ? tg.nty - 1 : ti[1] - 1) : ti[1];
      nt[2] = (S & 4) ? (ti[2] == 0 ? tg.ntz - 1 : ti[2] - 1) : ti[2];
      ns[0] = min(tg.tx, g.n0 - nt[0] * tg.tx);
      ns[1] = min(tg.ty, g.n1 - nt[1] * tg.ty);
      ns[2] = min(tg.tz, g.n2 - nt[2] * tg.tz);
      q[0] = (S & 1) ? ns[0] : l[0];
      q[1] = (S & 2) ? ns[1] : l[1];
      q[2] = (S & 4) ? ns[2] : l[2];
      long long nti = ((long long)nt[0] * tg.nty + nt[1]) * tg.ntz + nt[2];
      sum += halo[nti * tg.shell + shell_index(q[0], q[1], q[2], ns[0], ns[1], ns[2])];
    }
    long long m = ((long long)(o[0] + l[0]) * g.n1 + (o[1] + l[1])) * mrs + (o[2] + l[2]);
    mesh[m] += sum;
  }
}

// Writes the binned records once the tile counts are scanned.
int pmx_bin_fill(pmx_ctx* c, const int16_t* pmid, const float* disp, long long n,
                 const uint32_t* cond) {
  if (n == 0) return PMX_OK;
  const MeshGeom g = pmx_geom_padded(c);
  const unsigned nb = cond ? pmx_blocks_capped(n, PMX_REBIN_GRID) : pmx_blocks(n);
  pmx_tic(c, PMX_T_FILL);
  bin_fill_k<<<nb, PMX_TPB, 0, c->stream>>>(pmid, disp, n, g, c->tg, c->pkey, c->prank,
                                            c->tile_start, c->rec, c->pidx, cond);
  pmx_toc(c, PMX_T_FILL);
  PMX_LAUNCH_CHECK();
  return PMX_OK;
}

int pmx_scan_tiles(pmx_ctx* c, const uint32_t* cond) {
  pmx_tic(c, PMX_T_SCAN);
  int rc = pmx_scan_exclusive(c, c->tile_count, c->tile_start, c->ntiles + 1, cond);
  pmx_toc(c, PMX_T_SCAN);
  return rc;
}

// Capacities from the counts of the last binning: count + slack per tile
// (0 for the end slot), then the counts and the overflow flag are cleared.
__global__ void tile_caps_k(uint32_t* __restrict__ count, uint32_t* __restrict__ cap, long long nt,
                            uint32_t* __restrict__ flag) {
  const long long i = (long long)blockIdx.x * PMX_TPB + threadIdx.x;
  if (i <= nt) {
    const uint32_t k = count[i];
    cap[i] = i < nt ? k + pmx_tile_slack(k) : 0u;
    count[i] = 0u;
  }
  if (i == 0) flag[0] = 0u;
}

int pmx_bin_begin(pmx_ctx* c, long long n, int* cap) {
  int rc;
  if ((rc = pmx_ensure_ptcl(c, n))) return rc;
  // the capacities of caps_n particles fit the record slots and 32-bit slots
  *cap = c->caps_valid && c->caps_n <= c->ptcl_cap && c->rec_cap < (1LL << 32) - 1;
  c->caps_valid = 0;  // until pmx_bin_end has exact counts again
  if (!*cap) return pmx_fill_u32(c, c->tile_count, c->ntiles + 1, 0u);
  pmx_tic(c, PMX_T_SCAN);
  tile_caps_k<<<pmx_blocks(c->ntiles + 1), PMX_TPB, 0, c->stream>>>(c->tile_count, c->tile_cap,
                                                                    c->ntiles, c->flags);
  PMX_LAUNCH_CHECK();
  rc = pmx_scan_exclusive(c, c->tile_cap, c->tile_start, c->ntiles + 1);
  pmx_toc(c, PMX_T_SCAN);
  return rc;
}

BinOut pmx_bin_out(pmx_ctx* c) {
  return BinOut{c->pkey, c->prank, c->tile_start, c->rec, c->pidx, c->flags};
}

int pmx_bin_end(pmx_ctx* c, const int16_t* pmid, const float* disp, long long n, int cap) {
  int rc;
  if (!cap) {
    if ((rc = pmx_scan_tiles(c))) return rc;
    if ((rc = pmx_bin_fill(c, pmid, disp, n))) return rc;
  } else if (n > 0) {
    // a tile overflowed its capacity: count, scan and fill exactly, all
    // conditional on the device-side flag
    const uint32_t* f = c->flags;
    const MeshGeom g = pmx_geom_padded(c);
    if ((rc = pmx_fill_u32(c, c->tile_count, c->ntiles + 1, 0u, f))) return rc;
    pmx_tic(c, PMX_T_BIN);
    bin_count_k<false><<<pmx_blocks_capped(n, PMX_REBIN_GRID), PMX_TPB, 0, c->stream>>>(
        pmid, disp, n, g, c->tg, c->tile_count, pmx_bin_out(c), f);
    pmx_toc(c, PMX_T_BIN);
    PMX_LAUNCH_CHECK();
    if ((rc = pmx_scan_tiles(c, f))) return rc;
    if ((rc = pmx_bin_fill(c, pmid, disp, n, f))) return rc;
  }
  c->caps_valid = 1;
  c->caps_n = n;
  return PMX_OK;
}

// Bins the particles by tile (c->rec, c->tile_start, c->tile_count).
int pmx_bin_particles(pmx_ctx* c, const int16_t* pmid, const float* disp, long long n) {
  int rc, cap;
  if ((rc = pmx_bin_begin(c, n, &cap))) return rc;
  if (n > 0) {
    const MeshGeom g = pmx_geom_padded(c);
    const BinOut o = pmx_bin_out(c);
    pmx_tic(c, PMX_T_BIN);
    if (cap)
      bin_count_k<true><<<pmx_blocks(n), PMX_TPB, 0, c->stream>>>(pmid, disp, n, g, c->tg,
                                                                  c->tile_count, o, nullptr);
    else
      bin_count_k<false><<<pmx_blocks(n), PMX_TPB, 0, c->stream>>>(pmid, disp, n, g, c->tg,
                                                                   c->tile_count, o, nullptr);
    pmx_toc(c, PMX_T_BIN);
    PMX_LAUNCH_CHECK();
  }
  return pmx_bin_end(c, pmid, disp, n, cap);
}

// Deposits the binned records into mesh (row stride mrs), with the scalar
// value val or the per-record values of channel ch of c->vrec (ch >= 0).
// defer: leave the fold to the Poisson solve's row pass when it can take it.
static int deposit_run(pmx_ctx* c, float val, int perval, int nch, long long n, DepOut out,
                       int mrs, int accumulate, int defer) {
  const MeshGeom g = pmx_geom_padded(c);
  const TileGeom& tg = c->tg;
  const size_t lds = (size_t)(tg.tx + 1) * (tg.ty + 1) * (tg.tz + 1) * sizeof(unsigned long long);
  const unsigned nt = (unsigned)c->ntiles;
  const long long hstride = c->ntiles * tg.shell;
  int rc;
  if ((rc = pmx_ensure_halo(c, nch))) return rc;
  // value channel ch of record slot i: vrec[ch * rec_cap + i]
  ValSrc vs = {};
  for (int ch = 0; ch < 3; ++ch) {
    vs.p[ch] = c->vrec ? c->vrec + ch * c->rec_cap : nullptr;
    vs.step[ch] = 1;
  }
  pmx_tic(c, PMX_T_DEPOSIT);
  const bool full = pmx_full16(c);
  // several channels: one workgroup per tile and channel
  const bool split = c->deposit_split && perval && nch > 1 && nt % PMX_NXCD == 0;
#define DEPREC(PV, F, SP)                                                                        \
  deposit_k<PV, 0, F, SP><<<SP ? nt * (unsigned)nch : nt, DEP_TPB, lds, c->stream>>>(            \
      c->rec, vs, PV ? c->vmax : nullptr, c->tile_start, c->tile_count, g, tg, val, out, nch, mrs, \
      accumulate, c->halo, hstride)
  if (split && full)
    DEPREC(1, 1, 1);
  else if (split)
    DEPREC(1, 0, 1);
  else if (perval && full)
    DEPREC(1, 1, 0);
  else if (perval)
    DEPREC(1, 0, 0);
  else if (full)
    DEPREC(0, 1, 0);
  else
    DEPREC(0, 0, 0);
#undef DEPREC
  pmx_toc(c, PMX_T_DEPOSIT);
  PMX_LAUNCH_CHECK();
  if (defer && !accumulate && pmx_fold_deferrable(c)) return PMX_OK;
  pmx_tic(c, PMX_T_FOLD);
  fold_k<<<nt, PMX_TPB, 0, c->stream>>>(out, nch, mrs, c->halo, hstride, g, tg);
  pmx_toc(c, PMX_T_FOLD);
  PMX_LAUNCH_CHECK();
  return PMX_OK;
}

// Deposits the binned records into mesh (row stride mrs) with the scalar
// value val.
int pmx_deposit_binned(pmx_ctx* c, float val, long long n, float* mesh, int mrs, int accumulate,
                       int defer) {
  DepOut out = {{mesh, mesh, mesh}};
  return deposit_run(c, val, 0, 1, n, out, mrs, accumulate, defer);
}

// Deposits the three per-record value channels of c->vrec into three meshes
// (row stride mrs), overwriting them.
int pmx_deposit_binned3(pmx_ctx* c, long long n, float* const* mesh, int mrs, int defer) {
  DepOut out = {{mesh[0], mesh[1], mesh[2]}};
  return deposit_run(c, 0.f, 1, 3, n, out, mrs, 0, defer);
}

// Deposits val per tile-ordered particle (st, tile ranges start/count) into
// mesh (row stride mrs), overwriting it.
int pmx_deposit_state(pmx_ctx* c, const PState& st, const uint32_t* start, const uint32_t* count,
                      float val, float* mesh, int mrs, int defer) {
  DepOut out = {{mesh, mesh, mesh}};
  return pmx_deposit_state_vals(c, st, start, count, val, nullptr, 1, out, mrs, defer);
}

// Deposits tile-ordered particles into nch meshes (row stride mrs),
// overwriting them: val each, or (vals) channel ch of slot i at
// vals->p[ch][i * vals->step[ch]] with the fixed-point scale from c->vmax.
int pmx_deposit_state_vals(pmx_ctx* c, const PState& st, const uint32_t* start,
                           const uint32_t* count, float val, const ValSrc* vals, int nch,
                           DepOut out, int mrs, int defer) {
  const MeshGeom g = pmx_geom_padded(c);
  const TileGeom& tg = c->tg;
  const size_t lds = (size_t)(tg.tx + 1) * (tg.ty + 1) * (tg.tz + 1) * sizeof(unsigned long long);
  const unsigned nt = (unsigned)c->ntiles;
  const long long hstride = c->ntiles * tg.shell;
  int rc;
  if ((rc = pmx_ensure_halo(c, nch))) return rc;
  pmx_tic(c, PMX_T_DEPOSIT);
  const bool full = pmx_full16(c);
  // several channels: one workgroup per tile and channel
  const bool split = c->deposit_split && vals && nch > 1 && nt % PMX_NXCD == 0;
#define DEPST(PV, F, SP)                                                                           \
  deposit_k<PV, 1, F, SP><<<SP ? nt * (unsigned)nch : nt, DEP_TPB, lds, c->stream>>>(              \
      st.s0, PV ? *vals : ValSrc{}, PV ? c->vmax : nullptr, start, count, g, tg, PV ? 0.f : val, \
      out, nch, mrs, 0, c->halo, hstride)
  if (split && full)
    DEPST(1, 1, 1);
  else if (split)
    DEPST(1, 0, 1);
  else if (vals && full)
    DEPST(1, 1, 0);
  else if (vals)
    DEPST(1, 0, 0);
  else if (full)
    DEPST(0, 1, 0);
  else
    DEPST(0, 0, 0);
#undef DEPST
  pmx_toc(c, PMX_T_DEPOSIT);
  PMX_LAUNCH_CHECK();
  if (defer && pmx_fold_deferrable(c)) return PMX_OK;
  pmx_tic(c, PMX_T_FOLD);
  fold_k<<<nt, PMX_TPB, 0, c->stream>>>(out, nch, mrs, c->halo, hstride, g, tg);
  pmx_toc(c, PMX_T_FOLD);
  PMX_LAUNCH_CHECK();
  return PMX_OK;
}

extern "C" int pmx_scatter_mass(pmx_ctx* c, const int16_t* pmid, const float* disp, long long n,
                                float val, float* mesh, int accumulate) {
  PMX_ENTRY(c);
  PMX_REQUIRE(c && mesh && n >= 0, "pmx_scatter_mass: bad argument");
  PMX_REQUIRE(n == 0 || (pmid && disp), "pmx_scatter_mass: NULL particle arrays");
  PMX_REQUIRE(n < (1LL << 32) - 2048, "pmx_scatter_mass: too many particles");
  int rc;
  if ((rc = pmx_bin_particles(c, pmid, disp, n))) return rc;
  return pmx_deposit_binned(c, val, n, mesh, c->n[2], accumulate);
}

// Tile-staged gather of the force meshes in c->cbuf (padded) for the
// binned particles; with vel also the closing kick.
int pmx_tile_gather(pmx_ctx* c, float* acc, float* vel, float k2) {
  MeshGeom g = pmx_geom_padded(c);
  const TileGeom& tg = c->tg;
  size_t lds = (size_t)3 * (tg.tx + 1) * (tg.ty + 1) * (tg.tz + 1) * sizeof(float);
  unsigned nt = (unsigned)c->ntiles;
  const float* F0 = (const float*)c->cbuf[0];
  const float* F1 = (const float*)c->cbuf[1];
  const float* F2 = (const float*)c->cbuf[2];
  pmx_tic(c, PMX_T_GATHER);
  if (vel)
    tile_gather_k<1><<<nt, TG_TPB, lds, c->stream>>>(c->rec, c->pidx, c->tile_start, c->tile_count,
                                                     g, tg, F0, F1, F2, acc, vel, k2);
  else
    tile_gather_k<0><<<nt, TG_TPB, lds, c->stream>>>(c->rec, c->pidx, c->tile_start, c->tile_count,
                                                     g, tg, F0, F1, F2, acc, nullptr, 0.f);
  pmx_toc(c, PMX_T_GATHER);
  PMX_LAUNCH_CHECK();
  return PMX_OK;
}

// ----------------------------------------------------------- gravity --
int pmx_gravity_fwd_meshes(pmx_ctx* c, const int16_t* pmid, const float* disp, long long n,
                              float val, float coef) {
  int rc;
  if ((rc = pmx_bin_particles(c, pmid, disp, n))) return rc;
  if ((rc = pmx_deposit_binned(c, val, n, (float*)c->cbuf[0], c->rs, 0, 1))) return rc;
  return pmx_poisson_ctx(c, coef, 1);
}

extern "C" int pmx_gravity(pmx_ctx* c, const int16_t* pmid, const float* disp, long long n,
                           float val, float coef, float* acc) {
  PMX_ENTRY(c);
  PMX_REQUIRE(c && n >= 0, "pmx_gravity: bad argument");
  PMX_REQUIRE(n == 0 || (pmid && disp && acc), "pmx_gravity: NULL pointer");
  PMX_REQUIRE(n < (1LL << 32) - 2048, "pmx_gravity: too many particles");
  int rc;
  if ((rc = pmx_gravity_fwd_meshes(c, pmid, disp, n, val, coef))) return rc;
  if (n == 0) return PMX_OK;
  return pmx_tile_gather(c, acc, nullptr, 0.f);
}

// ------------------------------------------------------ time stepping --
__global__ void axpy3_k(float* __restrict__ y, const float* __restrict__ x, long long n3, float f) {
  long long i = (long long)blockIdx.x * PMX_TPB + threadIdx.x;
  if (i < n3) y[i] = y[i] + x[i] * f;
}

extern "C" int pmx_kick(pmx_ctx* c, long long n, float* vel, const float* acc, float k) {
  PMX_ENTRY(c);
  PMX_REQUIRE(c && n >= 0 && (n == 0 || (vel && acc)), "pmx_kick: bad argument");
  if (n == 0) return PMX_OK;
  axpy3_k<<<pmx_blocks(3 * n), PMX_TPB, 0, c->stream>>>(vel, acc, 3 * n, k);
  PMX_LAUNCH_CHECK();
  return PMX_OK;
}

extern "C" int pmx_drift(pmx_ctx* c, long long n, float* disp, const float* vel, float d) {
  PMX_ENTRY(c);
  PMX_REQUIRE(c && n >= 0 && (n == 0 || (disp && vel)), "pmx_drift: bad argument");
  if (n == 0) return PMX_OK;
  axpy3_k<<<pmx_blocks(3 * n), PMX_TPB, 0, c->stream>>>(disp, vel, 3 * n, d);
  PMX_LAUNCH_CHECK();
  return PMX_OK;
}

// kick (k1) + drift (d) fused with the binning of the drifted particles;
// the block's vel, acc and disp rows are staged through LDS
template <bool CAP>
__global__ void __launch_bounds__(PMX_TPB)
    kick_drift_bin_k(const int16_t* __restrict__ pmid, float* __restrict__ disp,
                     float* __restrict__ vel, const float* __restrict__ acc, long long n, float k1,
                     float dd, MeshGeom g, TileGeom tg, uint32_t* __restrict__ tile_count,
                     BinOut o, const float* __restrict__ fdev) {
  // factors from the device table when the call has one: {k1, d}
  if (fdev) {
    k1 = fdev[0];
    dd = fdev[1];
  }
  constexpr int T3 = 3 * PMX_TPB;
  __shared__ __align__(16) float sh[3 * T3];  // vel | acc | disp
  __shared__ __align__(16) int16_t sp[T3];
  const long long b0 = (long long)blockIdx.x * PMX_TPB;
  const int cnt = (int)(n - b0 < PMX_TPB ? n - b0 : PMX_TPB);
  pmx_stage_in3<3, PMX_TPB>(sh, {{vel, acc, disp}}, b0, cnt);
  pmx_stage_pmid<PMX_TPB>(sp, pmid, b0, cnt);
  __syncthreads();
  const int q = threadIdx.x;
  const bool active = q < cnt;
  uint32_t t = 0;
  CicAxis ax[3];
  float dp[3];
  if (active) {
    int pm[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      pm[a] = sp[3 * q + a];
      float v = sh[3 * q + a];
      v = v + sh[T3 + 3 * q + a] * k1;
      float x = sh[2 * T3 + 3 * q + a];
      x = x + v * dd;
      sh[3 * q + a] = v;
      sh[2 * T3 + 3 * q + a] = x;
      dp[a] = x;
    }
    cic_setup(pm, dp, g, ax);
    t = tile_of(ax, tg);
  }
  const uint32_t r = agg_slot(tile_count, t, active);
  if (active) pmx_bin_emit<CAP>(o, b0 + q, t, r, ax, tg);
  __syncthreads();
  pmx_stage_out3<1, PMX_TPB>({{vel}}, sh, b0, cnt);
  pmx_stage_out3<1, PMX_TPB>({{disp}}, sh + 2 * T3, b0, cnt);
}

int pmx_kick_drift_bin(pmx_ctx* c, const int16_t* pmid, float* disp, float* vel, const float* acc,
                       long long n, float k1, float d, uint32_t* count, const BinOut& o, int cap) {
  const MeshGeom g = pmx_geom_padded(c);
  // the first kick and drift of a call: entries 1 and 2 of a device table
  const float* fdev = c->dfac ? c->dfac + 1 : nullptr;
  pmx_tic(c, PMX_T_BIN);
  if (cap)
    kick_drift_bin_k<true><<<pmx_blocks(n), PMX_TPB, 0, c->stream>>>(pmid, disp, vel, acc, n, k1, d,
                                                                     g, c->tg, count, o, fdev);
  else
    kick_drift_bin_k<false><<<pmx_blocks(n), PMX_TPB, 0, c->stream>>>(pmid, disp, vel, acc, n, k1,
                                                                      d, g, c->tg, count, o, fdev);
  pmx_toc(c, PMX_T_BIN);
  PMX_LAUNCH_CHECK();
  return PMX_OK;
}

extern "C" int pmx_nbody_step(pmx_ctx* c, const int16_t* pmid, float* disp, float* vel, float* acc,
                              long long n, float val, float coef, float k1, float d, float k2) {
  PMX_ENTRY(c);
  PMX_REQUIRE(c && n >= 0, "pmx_nbody_step: bad argument");
  PMX_REQUIRE(n == 0 || (pmid && disp && vel && acc), "pmx_nbody_step: NULL pointer");
  PMX_REQUIRE(n < (1LL << 32) - 2048, "pmx_nbody_step: too many particles");
  int rc, cap;
  // every allocation before the first kernel writes disp and vel, so that
  // PMX_ENOMEM leaves the caller's arrays as they were
  if ((rc = pmx_ensure_scan(c, c->ntiles + 1))) return rc;
  if ((rc = pmx_bin_begin(c, n, &cap))) return rc;
  if (n > 0 && (rc = pmx_kick_drift_bin(c, pmid, disp, vel, acc, n, k1, d, c->tile_count,
                                         pmx_bin_out(c), cap)))
    return rc;
  if ((rc = pmx_bin_end(c, pmid, disp, n, cap))) return rc;
  if ((rc = pmx_deposit_binned(c, val, n, (float*)c->cbuf[0], c->rs, 0, 1))) return rc;
  if ((rc = pmx_poisson_ctx(c, coef, 1))) return rc;
  if (n > 0 && (rc = pmx_tile_gather(c, acc, vel, k2))) return rc;
  return PMX_OK;
}
