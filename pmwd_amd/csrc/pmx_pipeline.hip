// Multi-step KDK integration with the particles kept in tile order
// (pmx_nbody_steps).
//
// pmx_nbody_step reads each particle twice per step in the caller's order:
// once to kick, drift and bin it, once (through the binned index) to gather
// its force and close the step.  Here the particles live, between steps, in
// one of two tile-ordered copies (PState): the step's workgroup for a tile
// stages the tile's force meshes in LDS, reads the tile's particles
// contiguously, gathers the force, applies the closing kick of this step and
// the opening kick and drift of the next, and writes each particle to its
// new tile's slot in the other copy (capacity layout from this step's
// counts).  The deposit of the next step reads the new copy in tile order.
// Only the first kick/drift and the last closing kick touch the caller's
// arrays.  A tile that overflows its capacity makes the device redo the step
// exactly (count, scan, write), all launched unconditionally and skipped on
// the device unless the overflow flag is set.
//
// Every floating-point operation is the one pmx_nbody_step performs, in the
// same order, so the results are bitwise those of nsteps pmx_nbody_step calls
// (nbody.py:205-212 for each step).
#include "pmx_internal.h"

#ifndef PL_TPB
#define PL_TPB 512
#endif
#ifndef PL_U
#define PL_U 4
#endif
#ifndef PL_MINB
#define PL_MINB 1
#endif

enum { PL_CAP = 0, PL_COUNT = 1, PL_WRITE = 2, PL_LAST = 3 };

// Destination of a pipeline step.
struct PlOut {
  PState st;
  uint32_t* count;        // per-tile counts of the destination (atomics)
  const uint32_t* start;  // per-tile starts of the destination
  uint32_t* ptile;        // exact re-bin: tile and rank per source slot
  uint32_t* prank;
  uint32_t* flag;         // overflow flag
};

// The caller's arrays (last step).
struct PlUser {
  float* disp;
  float* vel;
  float* acc;
};

__device__ __forceinline__ void pl_put(const PState& st, uint32_t slot, float4 s0, float4 s1,
                                       uint16_t pz) {
  st.s0[slot] = s0;
  st.s1[slot] = s1;
  st.pz[slot] = pz;
}

// Writes particle i (after the first kick and drift) to the slot
// start[ptile[i]] + prank[i] of st, pm reduced mod the mesh extent.
__global__ void __launch_bounds__(PMX_TPB)
    pl_fill_k(const int16_t* __restrict__ pmid, const float* __restrict__ disp,
              const float* __restrict__ vel, long long n, MeshGeom g,
              const uint32_t* __restrict__ ptile, const uint32_t* __restrict__ prank,
              const uint32_t* __restrict__ start, PState st) {
  constexpr int T3 = 3 * PMX_TPB;
  __shared__ __align__(16) float sh[2 * T3];  // disp | vel
  __shared__ __align__(16) int16_t sp[T3];
  const long long b0 = (long long)blockIdx.x * PMX_TPB;
  const int cnt = (int)(n - b0 < PMX_TPB ? n - b0 : PMX_TPB);
  pmx_stage_in3<2, PMX_TPB>(sh, {{disp, vel}}, b0, cnt);
  pmx_stage_pmid<PMX_TPB>(sp, pmid, b0, cnt);
  __syncthreads();
  const int q = threadIdx.x;
  if (q >= cnt) return;
  const long long i = b0 + q;
  const uint32_t px = (uint32_t)pmx_wrap(sp[3 * q], g.n0);
  const uint32_t py = (uint32_t)pmx_wrap(sp[3 * q + 1], g.n1);
  const uint16_t pz = (uint16_t)pmx_wrap(sp[3 * q + 2], g.n2);
  const float4 s0 = make_float4(sh[3 * q], sh[3 * q + 1], sh[3 * q + 2],
                                __uint_as_float(px | (py << 16)));
  const float4 s1 = make_float4(sh[T3 + 3 * q], sh[T3 + 3 * q + 1], sh[T3 + 3 * q + 2],
                                __uint_as_float((uint32_t)i));
  pl_put(st, start[ptile[i]] + prank[i], s0, s1, pz);
}

// Capacities of the next copy from the counts of the current one; clears
// the next copy's counts and the overflow flag.
__global__ void pl_caps_k(const uint32_t* __restrict__ count_in, uint32_t* __restrict__ cap,
                          uint32_t* __restrict__ count_out, long long nt,
                          uint32_t* __restrict__ flag) {
  const long long i = (long long)blockIdx.x * PMX_TPB + threadIdx.x;
  if (i <= nt) {
    const uint32_t k = count_in[i];
    cap[i] = i < nt ? k + pmx_tile_slack(k) : 0u;
    count_out[i] = 0u;
  }
  if (i == 0) flag[0] = 0u;
}

// One workgroup per tile of the source copy: gather, closing kick k2; then
// (not LAST) opening kick k1n and drift dn of the next step and the move to
// the new tile:
//   PL_CAP    capacity slot in the destination, or the overflow flag;
//   PL_COUNT  exact re-bin, pass 1: tile and rank per source slot;
//   PL_WRITE  exact re-bin, pass 2: the particle to its exact slot;
//   PL_LAST   acc, vel and disp to the caller's arrays at the particle index.
// PL_COUNT and PL_WRITE return at once unless the overflow flag is set.
template <int MODE>
__global__ void __launch_bounds__(PL_TPB, PL_MINB)
    pl_step_k(PState in, const uint32_t* __restrict__ in_start,
              const uint32_t* __restrict__ in_count, MeshGeom g, TileGeom tg,
              const float* __restrict__ F0, const float* __restrict__ F1,
              const float* __restrict__ F2, float k2, float k1n, float dn, PlOut out,
              PlUser user) {
  if ((MODE == PL_COUNT || MODE == PL_WRITE) && out.flag[0] == 0u) return;
  extern __shared__ float sh[];
  // PL_CAP: per round, particles moving to each of the 27 tiles around this
  // one are counted in LDS (wave-aggregated) and take one contiguous run per
  // tile from a single global atomic, so a round's stores into a tile are
  // one run instead of a short run per wave; farther moves rank per wave
  __shared__ uint32_t s_cnt[27], s_base[27];
  int o[3], s[3], ti[3];
  const int t = (int)pmx_xcd_block(blockIdx.x, gridDim.x);
  tile_origin(t, tg, g, o, s, ti);
  const int LY = s[1] + 1, LZ = s[2] + 1;
  const int ltot = (s[0] + 1) * LY * LZ;
  const float* F[3] = {F0, F1, F2};
  if (MODE == PL_CAP && threadIdx.x < 27) s_cnt[threadIdx.x] = 0u;
  pmx_stage_tile<3, PL_TPB>(sh, o, LY, LZ, ltot, g, F);
  __syncthreads();
  const uint32_t start = in_start[t];
  const uint32_t cnt = in_count[t];
  const int ntl[3] = {tg.ntx, tg.nty, tg.ntz};
  for (uint32_t q0 = 0; q0 < cnt; q0 += PL_TPB * PL_U) {
    float4 a0[PL_U], a1[PL_U], n0[PL_U], n1[PL_U];
    uint16_t az[PL_U];
    uint32_t nt[PL_U];
    int code[PL_U];
#pragma unroll
    for (int u = 0; u < PL_U; ++u) {
      const uint32_t q = q0 + u * PL_TPB + threadIdx.x;
      if (q < cnt) {
        a0[u] = in.s0[start + q];
        a1[u] = in.s1[start + q];
        az[u] = in.pz[start + q];
      }
    }
#pragma unroll
    for (int u = 0; u < PL_U; ++u) {
      const uint32_t q = q0 + u * PL_TPB + threadIdx.x;
      n0[u] = make_float4(0.f, 0.f, 0.f, 0.f);
      n1[u] = n0[u];
      nt[u] = 0;
      code[u] = 27;
      if (q < cnt) {
        float acc[3] = {0.f, 0.f, 0.f};
        float d[3];
        int l[3];
        // binned by this tile, so always inside; the guard keeps LDS in bounds
        if (pmx_state_local(a0[u], az[u], g, o, s, d, l))
          pmx_tile_acc(sh, ltot, LY, LZ, d[0], d[1], d[2], l[0], l[1], l[2], acc);
        float v[3] = {a1[u].x, a1[u].y, a1[u].z};
        float x[3] = {a0[u].x, a0[u].y, a0[u].z};
#pragma unroll
        for (int a = 0; a < 3; ++a) v[a] = v[a] + acc[a] * k2;
        if (MODE == PL_LAST) {
          const long long i = __float_as_uint(a1[u].w);
#pragma unroll
          for (int a = 0; a < 3; ++a) {
            user.acc[3 * i + a] = acc[a];
            user.vel[3 * i + a] = v[a];
            user.disp[3 * i + a] = x[a];
          }
        } else {
#pragma unroll
          for (int a = 0; a < 3; ++a) {
            v[a] = v[a] + acc[a] * k1n;
            x[a] = x[a] + v[a] * dn;
          }
          int pm[3];
          pmx_state_pm(a0[u], az[u], pm);
          CicAxis ax[3];
          cic_setup(pm, x, g, ax);
          nt[u] = tile_of(ax, tg);
          n0[u] = make_float4(x[0], x[1], x[2], a0[u].w);
          n1[u] = make_float4(v[0], v[1], v[2], a1[u].w);
          if (MODE == PL_CAP) {
            // neighbour code 0..26 of the new tile, 27 when farther
            const unsigned tm[3] = {tg.mx, tg.my, tg.mz};
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
            code[u] = cc;
          }
        }
      }
    }
    if (MODE == PL_LAST) continue;
    if (MODE == PL_WRITE) {
#pragma unroll
      for (int u = 0; u < PL_U; ++u) {
        const uint32_t q = q0 + u * PL_TPB + threadIdx.x;
        if (q < cnt)
          pl_put(out.st, out.start[out.ptile[start + q]] + out.prank[start + q], n0[u], n1[u],
                 az[u]);
      }
      continue;
    }
    if (MODE == PL_COUNT) {
#pragma unroll
      for (int u = 0; u < PL_U; ++u) {
        const uint32_t q = q0 + u * PL_TPB + threadIdx.x;
        const uint32_t r = agg_slot(out.count, nt[u], q < cnt);
        if (q < cnt) {
          out.ptile[start + q] = nt[u];
          out.prank[start + q] = r;
        }
      }
      continue;
    }
    // PL_CAP: ranks within this round's run per neighbour tile, then the runs
    uint32_t rk[PL_U];
#pragma unroll
    for (int u = 0; u < PL_U; ++u) {
      const uint32_t q = q0 + u * PL_TPB + threadIdx.x;
      const bool near = q < cnt && code[u] < 27;
      const uint32_t rl = agg_slot(s_cnt, near ? (uint32_t)code[u] : 0u, near);
      const uint32_t rf = agg_slot(out.count, nt[u], q < cnt && code[u] == 27);
      rk[u] = near ? rl : rf;
    }
    __syncthreads();
    if (threadIdx.x < 27) {
      const uint32_t k = s_cnt[threadIdx.x];
      uint32_t base = 0;
      if (k) {
        int c3 = threadIdx.x, nb[3];
#pragma unroll
        for (int a = 2; a >= 0; --a) {
          int v = ti[a] + c3 % 3 - 1;
          nb[a] = v < 0 ? v + ntl[a] : (v >= ntl[a] ? v - ntl[a] : v);
          c3 /= 3;
        }
        const uint32_t tile = ((uint32_t)nb[0] * tg.nty + (uint32_t)nb[1]) * tg.ntz + (uint32_t)nb[2];
        base = atomicAdd(&out.count[tile], k);
      }
      s_base[threadIdx.x] = base;
      s_cnt[threadIdx.x] = 0u;
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < PL_U; ++u) {
      const uint32_t q = q0 + u * PL_TPB + threadIdx.x;
      if (q >= cnt) continue;
      const uint32_t r = code[u] < 27 ? s_base[code[u]] + rk[u] : rk[u];
      const uint32_t b = out.start[nt[u]], e = out.start[nt[u] + 1];
      if (r < e - b)
        pl_put(out.st, b + r, n0[u], n1[u], az[u]);
      else
        out.flag[0] = 1u;
    }
  }
}

template <int MODE>
static void pl_launch(pmx_ctx* c, int src, PlOut out, PlUser user, float k2, float k1n, float dn) {
  const MeshGeom g = pmx_geom_padded(c);
  const TileGeom& tg = c->tg;
  const size_t lds = (size_t)3 * (tg.tx + 1) * (tg.ty + 1) * (tg.tz + 1) * sizeof(float);
  pl_step_k<MODE><<<(unsigned)c->ntiles, PL_TPB, lds, c->stream>>>(
      c->pl[src], c->pl_start[src], c->pl_count[src], g, tg, (const float*)c->cbuf[0],
      (const float*)c->cbuf[1], (const float*)c->cbuf[2], k2, k1n, dn, out, user);
}

extern "C" int pmx_nbody_steps(pmx_ctx* c, const int16_t* pmid, float* disp, float* vel, float* acc,
                               long long n, float val, float coef, int nsteps, const float* k1,
                               const float* d, const float* k2) {
  PMX_REQUIRE(c && n >= 0 && nsteps >= 0, "pmx_nbody_steps: bad argument");
  PMX_REQUIRE(n == 0 || (pmid && disp && vel && acc), "pmx_nbody_steps: NULL pointer");
  PMX_REQUIRE(nsteps == 0 || (k1 && d && k2), "pmx_nbody_steps: NULL factor array");
  PMX_REQUIRE(n < (1LL << 32) - 2048, "pmx_nbody_steps: too many particles");
  int rc;
  if (nsteps == 0) return PMX_OK;
  if (n == 0) {
    // no particles: the steps solve an empty mesh, as pmx_nbody_step does
    for (int s = 0; s < nsteps; ++s)
      if ((rc = pmx_nbody_step(c, pmid, disp, vel, acc, 0, val, coef, k1[s], d[s], k2[s])))
        return rc;
    return PMX_OK;
  }
  if ((rc = pmx_ensure_ptcl(c, n))) return rc;
  if ((rc = pmx_ensure_state(c))) return rc;
  PMX_REQUIRE(c->rec_cap < (1LL << 32) - 1, "pmx_nbody_steps: too many particles");
  const MeshGeom g = pmx_geom_padded(c);
  const long long nt = c->ntiles;
  int A = 0;

  // first kick and drift on the caller's arrays, exact binning, fill copy A
  if ((rc = pmx_fill_u32(c, c->pl_count[A], nt + 1, 0u))) return rc;
  const BinOut bo = {c->pkey, c->prank, nullptr, nullptr, nullptr, c->flags};
  if ((rc = pmx_kick_drift_bin(c, pmid, disp, vel, acc, n, k1[0], d[0], c->pl_count[A], bo, 0)))
    return rc;
  pmx_tic(c, PMX_T_SCAN);
  rc = pmx_scan_exclusive(c, c->pl_count[A], c->pl_start[A], nt + 1);
  pmx_toc(c, PMX_T_SCAN);
  if (rc) return rc;
  pmx_tic(c, PMX_T_FILL);
  pl_fill_k<<<pmx_blocks(n), PMX_TPB, 0, c->stream>>>(pmid, disp, vel, n, g, c->pkey, c->prank,
                                                      c->pl_start[A], c->pl[A]);
  pmx_toc(c, PMX_T_FILL);
  PMX_LAUNCH_CHECK();

  const PlUser user = {disp, vel, acc};
  for (int s = 0; s < nsteps; ++s) {
    if ((rc = pmx_deposit_state(c, c->pl[A], c->pl_start[A], c->pl_count[A], val,
                                (float*)c->cbuf[0], c->rs)))
      return rc;
    if ((rc = pmx_poisson_ctx(c, coef))) return rc;
    if (s + 1 == nsteps) {
      pmx_tic(c, PMX_T_GATHER);
      pl_launch<PL_LAST>(c, A, PlOut{}, user, k2[s], 0.f, 0.f);
      pmx_toc(c, PMX_T_GATHER);
      PMX_LAUNCH_CHECK();
      break;
    }
    const int B = 1 - A;
    pmx_tic(c, PMX_T_SCAN);
    pl_caps_k<<<pmx_blocks(nt + 1), PMX_TPB, 0, c->stream>>>(c->pl_count[A], c->tile_cap,
                                                             c->pl_count[B], nt, c->flags);
    PMX_LAUNCH_CHECK();
    rc = pmx_scan_exclusive(c, c->tile_cap, c->pl_start[B], nt + 1);
    pmx_toc(c, PMX_T_SCAN);
    if (rc) return rc;
    const PlOut out = {c->pl[B], c->pl_count[B], c->pl_start[B], c->pl_tile, c->pl_rank, c->flags};
    pmx_tic(c, PMX_T_GATHER);
    pl_launch<PL_CAP>(c, A, out, user, k2[s], k1[s + 1], d[s + 1]);
    pmx_toc(c, PMX_T_GATHER);
    PMX_LAUNCH_CHECK();
    // a tile overflowed: redo the step exactly (skipped on the device otherwise)
    if ((rc = pmx_fill_u32(c, c->pl_count[B], nt + 1, 0u, c->flags))) return rc;
    pmx_tic(c, PMX_T_BIN);
    pl_launch<PL_COUNT>(c, A, out, user, k2[s], k1[s + 1], d[s + 1]);
    PMX_LAUNCH_CHECK();
    if ((rc = pmx_scan_exclusive(c, c->pl_count[B], c->pl_start[B], nt + 1, c->flags))) return rc;
    pl_launch<PL_WRITE>(c, A, out, user, k2[s], k1[s + 1], d[s + 1]);
    pmx_toc(c, PMX_T_BIN);
    PMX_LAUNCH_CHECK();
    A = B;
  }
  return PMX_OK;
}
