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
// workgroups of the exact re-bin launches, which return at once unless a
// tile overflowed
#define PL_REBIN_GRID 2048u

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

__device__ __forceinline__ void pl_put(const PState& st, uint32_t slot, float4 s0, float4 s1) {
  st.s0[slot] = s0;
  st.s1[slot] = s1;
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
  const uint32_t pz = (uint32_t)pmx_wrap(sp[3 * q + 2], g.n2);
  const float4 s0 = make_float4(sh[3 * q], sh[3 * q + 1], sh[3 * q + 2],
                                __uint_as_float(pmx_state_pack(px, py, pz)));
  const float4 s1 = make_float4(sh[T3 + 3 * q], sh[T3 + 3 * q + 1], sh[T3 + 3 * q + 2],
                                __uint_as_float((uint32_t)i));
  pl_put(st, start[ptile[i]] + prank[i], s0, s1);
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
template <int MODE, int FULL = 0>
__global__ void __launch_bounds__(PL_TPB, PL_MINB)
    pl_step_k(PState in, const uint32_t* __restrict__ in_start,
              const uint32_t* __restrict__ in_count, MeshGeom g, TileGeom tg,
              const float* __restrict__ F0, const float* __restrict__ F1,
              const float* __restrict__ F2, float k2, float k1n, float dn, PlOut out,
              PlUser user, const float* __restrict__ fdev) {
  if ((MODE == PL_COUNT || MODE == PL_WRITE) && out.flag[0] == 0u) return;
  // factors from the device table when the call has one: this step's
  // {k1, d, k2} then the next step's {k1, d}
  if (fdev) {
    k2 = fdev[2];
    if (MODE != PL_LAST) {
      k1n = fdev[3];
      dn = fdev[4];
    }
  }
  extern __shared__ float sh[];
  // PL_CAP: runs per neighbour tile and round (pl_nbr_code, pl_round_bases),
  // so a round's stores into a tile are one run instead of one per wave
  __shared__ uint32_t s_cnt[27], s_base[27];
  // the exact re-bin (COUNT, WRITE: rare) walks the tiles with a small grid,
  // so that it costs little when the overflow flag is clear
  auto tile = [&](const int t) {
    int o[3], s[3], ti[3];
    tile_origin(t, tg, g, o, s, ti);
    if (FULL) s[0] = s[1] = s[2] = 16;
    const int LY = s[1] + 1, LZ = s[2] + 1;
    const int ltot = (s[0] + 1) * LY * LZ;
    const float* F[3] = {F0, F1, F2};
    if (MODE == PL_CAP && threadIdx.x < 27) s_cnt[threadIdx.x] = 0u;
    pmx_stage_tile<3, PL_TPB>(sh, o, LY, LZ, ltot, g, F);
    __syncthreads();
    const uint32_t start = in_start[t];
    const uint32_t cnt = in_count[t];
    for (uint32_t q0 = 0; q0 < cnt; q0 += PL_TPB * PL_U) {
      float4 a0[PL_U], a1[PL_U], n0[PL_U], n1[PL_U];
      uint32_t nt[PL_U];
      int code[PL_U];
  #pragma unroll
      for (int u = 0; u < PL_U; ++u) {
        const uint32_t q = q0 + u * PL_TPB + threadIdx.x;
        if (q < cnt) {
          a0[u] = in.s0[start + q];
          a1[u] = in.s1[start + q];
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
          CicFrac fr;
          int l[3];
          // binned by this tile, so always inside; the guard keeps LDS in bounds
          if (pmx_state_local(a0[u], g, o, s, fr, l))
            pmx_tile_acc(sh, ltot, LY, LZ, fr, l[0], l[1], l[2], acc);
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
            pmx_state_pm(a0[u], pm);
            CicAxis ax[3];
            cic_setup(pm, x, g, ax);
            nt[u] = tile_of(ax, tg);
            n0[u] = make_float4(x[0], x[1], x[2], a0[u].w);
            n1[u] = make_float4(v[0], v[1], v[2], a1[u].w);
            if (MODE == PL_CAP) code[u] = pl_nbr_code(ax, ti, tg);
          }
        }
      }
      if (MODE == PL_LAST) continue;
      if (MODE == PL_WRITE) {
  #pragma unroll
        for (int u = 0; u < PL_U; ++u) {
          const uint32_t q = q0 + u * PL_TPB + threadIdx.x;
          if (q < cnt)
            pl_put(out.st, out.start[out.ptile[start + q]] + out.prank[start + q], n0[u], n1[u]);
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
      pl_round_bases(s_cnt, s_base, ti, tg, out.count);
      __syncthreads();
  #pragma unroll
      for (int u = 0; u < PL_U; ++u) {
        const uint32_t q = q0 + u * PL_TPB + threadIdx.x;
        if (q >= cnt) continue;
        const uint32_t r = code[u] < 27 ? s_base[code[u]] + rk[u] : rk[u];
        const uint32_t b = out.start[nt[u]], e = out.start[nt[u] + 1];
        if (r < e - b)
          pl_put(out.st, b + r, n0[u], n1[u]);
        else
          out.flag[0] = 1u;
      }
    }
  };
  if (MODE == PL_COUNT || MODE == PL_WRITE) {
    const unsigned ntl = (unsigned)tg.ntx * tg.nty * tg.ntz;
    for (unsigned tb = blockIdx.x; tb < ntl; tb += gridDim.x) {
      tile((int)pmx_xcd_block(tb, ntl));
      __syncthreads();
    }
  } else {
    tile((int)pmx_xcd_block(blockIdx.x, gridDim.x));
  }
}

template <int MODE>
static void pl_launch(pmx_ctx* c, int src, PlOut out, PlUser user, float k2, float k1n, float dn,
                      int step) {
  const float* fdev = c->dfac ? c->dfac + 1 + 3 * (long long)step : nullptr;
  const MeshGeom g = pmx_geom_padded(c);
  const TileGeom& tg = c->tg;
  const size_t lds = (size_t)3 * (tg.tx + 1) * (tg.ty + 1) * (tg.tz + 1) * sizeof(float);
  const bool rebin = MODE == PL_COUNT || MODE == PL_WRITE;
  const unsigned grid = rebin && c->ntiles > PL_REBIN_GRID ? PL_REBIN_GRID : (unsigned)c->ntiles;
  // tile strides as constants when every tile is a full 16^3 tile
  if (pmx_full16(c))
    pl_step_k<MODE, 1><<<grid, PL_TPB, lds, c->stream>>>(
        c->pl[src], c->pl_start[src], c->pl_count[src], g, tg, (const float*)c->cbuf[0],
        (const float*)c->cbuf[1], (const float*)c->cbuf[2], k2, k1n, dn, out, user, fdev);
  else
    pl_step_k<MODE, 0><<<grid, PL_TPB, lds, c->stream>>>(
        c->pl[src], c->pl_start[src], c->pl_count[src], g, tg, (const float*)c->cbuf[0],
        (const float*)c->cbuf[1], (const float*)c->cbuf[2], k2, k1n, dn, out, user, fdev);
}

// The K-step pipeline; the factors come from the host arrays k1, d, k2, or
// (when c->dfac is set) from the device table, the host arrays then NULL.
static int nbody_steps_run(pmx_ctx* c, const int16_t* pmid, float* disp, float* vel, float* acc,
                           long long n, float val, float coef, int nsteps, const float* k1,
                           const float* d, const float* k2) {
  int rc;
  // host factors (zero placeholders when the kernels read the device table)
  auto K1 = [&](int s) { return k1 ? k1[s] : 0.f; };
  auto D = [&](int s) { return d ? d[s] : 0.f; };
  auto K2 = [&](int s) { return k2 ? k2[s] : 0.f; };
  // every allocation before the first kernel writes disp and vel, so that
  // PMX_ENOMEM leaves the caller's arrays as they were
  if ((rc = pmx_ensure_ptcl(c, n))) return rc;
  if ((rc = pmx_ensure_state(c))) return rc;
  if ((rc = pmx_ensure_scan(c, c->ntiles + 1))) return rc;
  PMX_REQUIRE(c->rec_cap < (1LL << 32) - 1, "pmx_nbody_steps: too many particles");
  const MeshGeom g = pmx_geom_padded(c);
  const long long nt = c->ntiles;
  int A = 0;

  // first kick and drift on the caller's arrays, exact binning, fill copy A
  if ((rc = pmx_fill_u32(c, c->pl_count[A], nt + 1, 0u))) return rc;
  const BinOut bo = {c->pkey, c->prank, nullptr, nullptr, nullptr, c->flags};
  if ((rc = pmx_kick_drift_bin(c, pmid, disp, vel, acc, n, K1(0), D(0), c->pl_count[A], bo, 0)))
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
                                (float*)c->cbuf[0], c->rs, 1)))
      return rc;
    if ((rc = pmx_poisson_ctx(c, coef, 1))) return rc;
    if (s + 1 == nsteps) {
      pmx_tic(c, PMX_T_GATHER);
      pl_launch<PL_LAST>(c, A, PlOut{}, user, K2(s), 0.f, 0.f, s);
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
    pl_launch<PL_CAP>(c, A, out, user, K2(s), K1(s + 1), D(s + 1), s);
    pmx_toc(c, PMX_T_GATHER);
    PMX_LAUNCH_CHECK();
    // a tile overflowed: redo the step exactly (skipped on the device otherwise)
    if ((rc = pmx_fill_u32(c, c->pl_count[B], nt + 1, 0u, c->flags))) return rc;
    pmx_tic(c, PMX_T_BIN);
    pl_launch<PL_COUNT>(c, A, out, user, K2(s), K1(s + 1), D(s + 1), s);
    PMX_LAUNCH_CHECK();
    if ((rc = pmx_scan_exclusive(c, c->pl_count[B], c->pl_start[B], nt + 1, c->flags))) return rc;
    pl_launch<PL_WRITE>(c, A, out, user, K2(s), K1(s + 1), D(s + 1), s);
    pmx_toc(c, PMX_T_BIN);
    PMX_LAUNCH_CHECK();
    A = B;
  }
  return PMX_OK;
}

extern "C" int pmx_nbody_steps(pmx_ctx* c, const int16_t* pmid, float* disp, float* vel, float* acc,
                               long long n, float val, float coef, int nsteps, const float* k1,
                               const float* d, const float* k2) {
  PMX_ENTRY(c);
  PMX_REQUIRE(c && n >= 0 && nsteps >= 0, "pmx_nbody_steps: bad argument");
  PMX_REQUIRE(n == 0 || (pmid && disp && vel && acc), "pmx_nbody_steps: NULL pointer");
  PMX_REQUIRE(nsteps == 0 || (k1 && d && k2), "pmx_nbody_steps: NULL factor array");
  PMX_REQUIRE(n < (1LL << 32) - 2048, "pmx_nbody_steps: too many particles");
  int rc;
  if (nsteps == 0) return PMX_OK;
  if (n == 0 || c->n[0] > PMX_PL_MAXN || c->n[1] > PMX_PL_MAXN || c->n[2] > PMX_PL_MAXN) {
    // no particles (the steps solve an empty mesh), or a mesh axis longer
    // than the tile-ordered state packs: one step at a time
    for (int s = 0; s < nsteps; ++s)
      if ((rc = pmx_nbody_step(c, pmid, disp, vel, acc, n, val, coef, k1[s], d[s], k2[s])))
        return rc;
    return PMX_OK;
  }
  return nbody_steps_run(c, pmid, disp, vel, acc, n, val, coef, nsteps, k1, d, k2);
}

// Sets c->dfac for the duration of a call.
struct DfacScope {
  pmx_ctx* c;
  DfacScope(pmx_ctx* c_, const float* f) : c(c_) { c->dfac = f; }
  ~DfacScope() { c->dfac = nullptr; }
};

static int steps_dev_check(pmx_ctx* c, const int16_t* pmid, float* disp, float* vel, float* acc,
                           long long n, int nsteps, const float* fac, const char* who) {
  PMX_REQUIRE(c && n > 0 && nsteps > 0 && fac, "%s: bad argument", who);
  PMX_REQUIRE(pmid && disp && vel && acc, "%s: NULL pointer", who);
  PMX_REQUIRE(n < (1LL << 32) - 2048, "%s: too many particles", who);
  PMX_REQUIRE(c->n[0] <= PMX_PL_MAXN && c->n[1] <= PMX_PL_MAXN && c->n[2] <= PMX_PL_MAXN,
              "%s: mesh axes above %d cells run pmx_nbody_steps", who, PMX_PL_MAXN);
  return PMX_OK;
}

extern "C" int pmx_nbody_steps_dev(pmx_ctx* c, const int16_t* pmid, float* disp, float* vel,
                                   float* acc, long long n, float val, int nsteps,
                                   const float* fac) {
  PMX_ENTRY(c);
  int rc;
  if ((rc = steps_dev_check(c, pmid, disp, vel, acc, n, nsteps, fac, "pmx_nbody_steps_dev")))
    return rc;
  DfacScope scope(c, fac);
  return nbody_steps_run(c, pmid, disp, vel, acc, n, val, 0.f, nsteps, nullptr, nullptr, nullptr);
}

// The forward model's N-body part with device factors: the initial force
// (nbody_init), nsteps KDK steps, and the CIC density of the final state.
static int forward_dev(pmx_ctx* c, const int16_t* pmid, float* disp, float* vel, float* acc,
                       long long n, float val, int nsteps, const float* fac, float* dens) {
  int rc;
  DfacScope scope(c, fac);
  if ((rc = pmx_gravity(c, pmid, disp, n, val, 0.f, acc))) return rc;
  if ((rc = nbody_steps_run(c, pmid, disp, vel, acc, n, val, 0.f, nsteps, nullptr, nullptr,
                            nullptr)))
    return rc;
  if (dens && (rc = pmx_scatter_mass(c, pmid, disp, n, val, dens, 0))) return rc;
  return PMX_OK;
}

extern "C" int pmx_graph_capture(pmx_ctx* c, const int16_t* pmid, float* disp, float* vel,
                                 float* acc, long long n, float val, int nsteps, const float* fac,
                                 float* dens) {
  int rc;
  if ((rc = steps_dev_check(c, pmid, disp, vel, acc, n, nsteps, fac, "pmx_graph_capture")))
    return rc;
  PMX_REQUIRE(!c->timer, "pmx_graph_capture: timing is enabled on the context");
  PMX_REQUIRE(!pmx_capturing(c), "pmx_graph_capture: the stream is already capturing");
  PMX_REQUIRE(c->stream != nullptr, "pmx_graph_capture: the context is on the null stream");
  PMX_CHECK_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  rc = forward_dev(c, pmid, disp, vel, acc, n, val, nsteps, fac, dens);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(c->stream, &g);
  if (rc || e != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    (void)hipGetLastError();
    if (!rc) pmx_set_error("pmx_graph_capture: %s", hipGetErrorString(e));
    return rc ? rc : PMX_EHIP;
  }
  hipGraphExec_t x = nullptr;
  const hipError_t ei = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  if (ei != hipSuccess) {
    (void)hipGraphDestroy(g);
    (void)hipGetLastError();
    pmx_set_error("pmx_graph_capture: instantiate: %s", hipGetErrorString(ei));
    return PMX_EHIP;
  }
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  c->graph = g;
  c->gexec = x;
  return PMX_OK;
}

extern "C" int pmx_graph_launch(pmx_ctx* c) {
  PMX_ENTRY(c);
  PMX_REQUIRE(c && c->gexec, "pmx_graph_launch: no captured graph");
  PMX_CHECK_HIP(hipGraphLaunch(c->gexec, c->stream));
  return PMX_OK;
}

// A pack: one executable graph whose independent branches are the captured
// graphs of several contexts, so one launch runs several simulations.
struct pmx_pack {
  hipGraph_t graph;
  hipGraphExec_t exec;
};

extern "C" int pmx_pack_create(pmx_ctx* const* cs, int k, pmx_pack** out) {
  PMX_REQUIRE(cs && k > 0 && out, "pmx_pack_create: bad argument");
  for (int i = 0; i < k; ++i)
    PMX_REQUIRE(cs[i] && cs[i]->graph, "pmx_pack_create: context %d has no captured graph", i);
  for (int i = 0; i < k; ++i)
    for (int j = 0; j < i; ++j)
      PMX_REQUIRE(cs[i] != cs[j], "pmx_pack_create: context %d appears twice", i);
  pmx_pack* p = (pmx_pack*)calloc(1, sizeof(pmx_pack));
  if (!p) {
    pmx_set_error("pmx_pack_create: out of host memory");
    return PMX_ENOMEM;
  }
  hipError_t e = hipGraphCreate(&p->graph, 0);
  for (int i = 0; i < k && e == hipSuccess; ++i) {
    hipGraphNode_t node;
    e = hipGraphAddChildGraphNode(&node, p->graph, nullptr, 0, cs[i]->graph);
  }
  if (e == hipSuccess) e = hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    if (p->graph) (void)hipGraphDestroy(p->graph);
    free(p);
    (void)hipGetLastError();
    pmx_set_error("pmx_pack_create: %s", hipGetErrorString(e));
    return PMX_EHIP;
  }
  *out = p;
  return PMX_OK;
}

extern "C" int pmx_pack_launch(pmx_pack* p, void* stream) {
  PMX_REQUIRE(p && p->exec, "pmx_pack_launch: bad pack");
  PMX_CHECK_HIP(hipGraphLaunch(p->exec, (hipStream_t)stream));
  return PMX_OK;
}

extern "C" int pmx_pack_destroy(pmx_pack* p) {
  if (!p) return PMX_OK;
  if (p->exec) (void)hipGraphExecDestroy(p->exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  free(p);
  return PMX_OK;
}

// ------------------------------------------------ reverse-time pipeline --
// pmx_nbody_steps_adj: nsteps reverse KDK steps (pmx_nbody_step_adj chained
// with its pending closing kick) with the particles and their cotangents in
// tile order between steps.  Per step, over the copy binned at this step's
// drifted positions:
//   density deposit, Poisson: the force meshes F stay in the context buffers;
//   three-channel deposit of vel_cot and the adjoint Poisson solve into three
//   more buffers: the density cotangent R;
//   al_step_k  with F and R staged: acc and the gather part of acc_cot (F),
//              the scatter part (R), sum vel_cot.acc; then this step's
//              closing kick adjoint, the next step's opening kick and drift
//              adjoints, |vel_cot| maxima for the next deposit, and the move
//              to the other copy (last step: the caller's arrays).
// Float operations are those of the per-step path in the same order, so
// every array is bitwise what the chained calls give; the float64 dot
// products sum the same terms in another grouping.

__device__ __forceinline__ void al_put(const AState& st, uint32_t slot, float4 s0, float4 s1,
                                       float4 c0, float2 c1) {
  st.p.s0[slot] = s0;
  st.p.s1[slot] = s1;
  st.c0[slot] = c0;
  st.c1[slot] = c1;
}

// Writes particle i (after the first reverse kick and drift) with its
// cotangents to slot start[ptile[i]] + prank[i] of st; |vel_cot| maxima per
// channel to vmax.
__global__ void __launch_bounds__(PMX_TPB)
    al_fill_k(const int16_t* __restrict__ pmid, const float* __restrict__ disp,
              const float* __restrict__ vel, const float* __restrict__ disp_cot,
              const float* __restrict__ vel_cot, long long n, MeshGeom g,
              const uint32_t* __restrict__ ptile, const uint32_t* __restrict__ prank,
              const uint32_t* __restrict__ start, AState st, float* __restrict__ vmax) {
  constexpr int T3 = 3 * PMX_TPB;
  __shared__ __align__(16) float sh[4 * T3];  // disp | vel | disp_cot | vel_cot
  __shared__ __align__(16) int16_t sp[T3];
  const long long b0 = (long long)blockIdx.x * PMX_TPB;
  const int cnt = (int)(n - b0 < PMX_TPB ? n - b0 : PMX_TPB);
  pmx_stage_in3<4, PMX_TPB>(sh, {{disp, vel, disp_cot, vel_cot}}, b0, cnt);
  pmx_stage_pmid<PMX_TPB>(sp, pmid, b0, cnt);
  __syncthreads();
  const int q = threadIdx.x;
  float vm[3] = {0.f, 0.f, 0.f};
  if (q < cnt) {
    const long long i = b0 + q;
    const uint32_t px = (uint32_t)pmx_wrap(sp[3 * q], g.n0);
    const uint32_t py = (uint32_t)pmx_wrap(sp[3 * q + 1], g.n1);
    const uint32_t pz = (uint32_t)pmx_wrap(sp[3 * q + 2], g.n2);
    const float* e = sh + 3 * q;
#pragma unroll
    for (int a = 0; a < 3; ++a) vm[a] = fabsf(e[3 * T3 + a]);
    al_put(st, start[ptile[i]] + prank[i],
           make_float4(e[0], e[1], e[2], __uint_as_float(pmx_state_pack(px, py, pz))),
           make_float4(e[T3], e[T3 + 1], e[T3 + 2], __uint_as_float((uint32_t)i)),
           make_float4(e[2 * T3], e[2 * T3 + 1], e[2 * T3 + 2], e[3 * T3]),
           make_float2(e[3 * T3 + 1], e[3 * T3 + 2]));
  }
#pragma unroll
  for (int a = 0; a < 3; ++a) wave_max_store(vm[a], vmax + a);
}

#define AL_TPB 512
#ifndef AL_U
#define AL_U 1
#endif

// Destination of a reverse pipeline step.
struct AlOut {
  AState st;
  uint32_t* count;
  const uint32_t* start;
  uint32_t* ptile;
  uint32_t* prank;
  uint32_t* flag;
};

// The caller's arrays (last step).
struct AlUser {
  float* disp;
  float* vel;
  float* acc;
  float* disp_cot;
  float* vel_cot;
  float* acc_cot;
};

// One workgroup per tile with the three force meshes F and the density
// cotangent R staged: acc and acc_cot of the force VJP (pmx_vjp_one, then
// + pmx_grad_one as the per-step path adds them); pf[block] = sum
// vel_cot.acc.  Then (not LAST) the closing kick adjoint kp of this step
// (skipped when 0), the next step's kick adjoint ka and drift adjoint dd
// (position replaced by snap[index] when snap is given), pd[block] = the
// drift adjoint's sum disp_cot.vel, |vel_cot| maxima to vmax, and the move
// to the other copy as pl_step_k's modes; LAST writes the caller's arrays.
// Partials and maxima come from PL_CAP and PL_LAST only.
// (two workgroups per CU: at most 128 VGPRs, as the 79 KB of LDS allows)
template <int MODE, int FULL = 0>
__global__ void __launch_bounds__(AL_TPB) __attribute__((amdgpu_waves_per_eu(4)))
    al_step_k(AState in, const uint32_t* __restrict__ in_start,
              const uint32_t* __restrict__ in_count, MeshGeom g, TileGeom tg,
              const float* __restrict__ F0, const float* __restrict__ F1,
              const float* __restrict__ F2, const float* __restrict__ R, float val, float kp,
              float ka, float dd, const float* __restrict__ snap, AlOut out, AlUser user,
              float* __restrict__ vmax, double* __restrict__ pf, double* __restrict__ pd) {
  if ((MODE == PL_COUNT || MODE == PL_WRITE) && out.flag[0] == 0u) return;
  extern __shared__ float sh[];
  __shared__ uint32_t s_cnt[27], s_base[27];
  double fsum = 0, dsum = 0;
  float vm[3] = {0.f, 0.f, 0.f};
  auto tile = [&](const int t) {
    int o[3], s[3], ti[3];
    tile_origin(t, tg, g, o, s, ti);
    if (FULL) s[0] = s[1] = s[2] = 16;
    const int LY = s[1] + 1, LZ = s[2] + 1;
    const int ltot = (s[0] + 1) * LY * LZ;
    const float* F[4] = {F0, F1, F2, R};
    if (MODE == PL_CAP && threadIdx.x < 27) s_cnt[threadIdx.x] = 0u;
    pmx_stage_tile<4, AL_TPB>(sh, o, LY, LZ, ltot, g, F);
    __syncthreads();
    const uint32_t start = in_start[t];
    const uint32_t cnt = in_count[t];
    // the next round's particles are loaded before this round's arithmetic,
    // so they are in flight meanwhile (5% on the 512^3 reverse step)
    float4 pa0[AL_U], pa1[AL_U], pc0[AL_U];
    float2 pc1[AL_U];
    auto fetch = [&](uint32_t q0, float4* a0, float4* a1, float4* c0, float2* c1) {
#pragma unroll
      for (int u = 0; u < AL_U; ++u) {
        const uint32_t q = q0 + u * AL_TPB + threadIdx.x;
        if (q < cnt) {
          const uint32_t sl = start + q;
          a0[u] = in.p.s0[sl];
          a1[u] = in.p.s1[sl];
          c0[u] = in.c0[sl];
          c1[u] = in.c1[sl];
        }
      }
    };
    fetch(0, pa0, pa1, pc0, pc1);
    for (uint32_t q0 = 0; q0 < cnt; q0 += AL_TPB * AL_U) {
      float4 a0[AL_U], a1[AL_U], c0[AL_U];
      float2 c1[AL_U];
#pragma unroll
      for (int u = 0; u < AL_U; ++u) {
        a0[u] = pa0[u];
        a1[u] = pa1[u];
        c0[u] = pc0[u];
        c1[u] = pc1[u];
      }
      if (q0 + AL_TPB * AL_U < cnt) fetch(q0 + AL_TPB * AL_U, pa0, pa1, pc0, pc1);
      float4 n0[AL_U], n1[AL_U], m0[AL_U];
      float2 m1[AL_U];
      uint32_t nt[AL_U];
      int code[AL_U];
#pragma unroll
      for (int u = 0; u < AL_U; ++u) {
        const uint32_t q = q0 + u * AL_TPB + threadIdx.x;
        nt[u] = 0;
        code[u] = 27;
        if (q >= cnt) continue;
        CicFrac fr;
        int l[3];
        float ac[3] = {0.f, 0.f, 0.f}, at[3] = {0.f, 0.f, 0.f}, gr[3] = {0.f, 0.f, 0.f};
        float x[3] = {a0[u].x, a0[u].y, a0[u].z}, v[3] = {a1[u].x, a1[u].y, a1[u].z};
        float dc[3] = {c0[u].x, c0[u].y, c0[u].z}, vc[3] = {c0[u].w, c1[u].x, c1[u].y};
        // binned by this tile, so always inside; the guard keeps LDS in bounds
        if (pmx_state_local(a0[u], g, o, s, fr, l)) {
          pmx_vjp_one(sh, ltot, LY, LZ, fr, l, vc, g.cell, g.inv_cell, ac, at);
          pmx_grad_one(sh + 3 * ltot, LY, LZ, fr, l, val, g.cell, g.inv_cell, gr);
        }
        if (MODE == PL_CAP || MODE == PL_LAST)
          fsum += (double)vc[0] * ac[0] + (double)vc[1] * ac[1] + (double)vc[2] * ac[2];
#pragma unroll
        for (int a = 0; a < 3; ++a) at[a] = at[a] + gr[a];
        if (MODE == PL_LAST) {
          const long long i = __float_as_uint(a1[u].w);
#pragma unroll
          for (int a = 0; a < 3; ++a) {
            user.disp[3 * i + a] = x[a];
            user.vel[3 * i + a] = v[a];
            user.disp_cot[3 * i + a] = dc[a];
            user.vel_cot[3 * i + a] = vc[a];
            user.acc[3 * i + a] = ac[a];
            user.acc_cot[3 * i + a] = at[a];
          }
          continue;
        }
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          if (kp != 0.f) {
            v[a] = v[a] + ac[a] * kp;
            dc[a] = dc[a] - at[a] * kp;
          }
          v[a] = v[a] + ac[a] * ka;
          dc[a] = dc[a] - at[a] * ka;
          x[a] = x[a] + v[a] * dd;
          vc[a] = vc[a] - dc[a] * dd;
          dsum += (double)dc[a] * v[a];
          vm[a] = fmaxf(vm[a], fabsf(vc[a]));
        }
        if (snap) {
          const long long i = __float_as_uint(a1[u].w);
#pragma unroll
          for (int a = 0; a < 3; ++a) x[a] = snap[3 * i + a];
        }
        int pm[3];
        pmx_state_pm(a0[u], pm);
        CicAxis ax[3];
        cic_setup(pm, x, g, ax);
        nt[u] = tile_of(ax, tg);
        if (MODE == PL_CAP) code[u] = pl_nbr_code(ax, ti, tg);
        n0[u] = make_float4(x[0], x[1], x[2], a0[u].w);
        n1[u] = make_float4(v[0], v[1], v[2], a1[u].w);
        m0[u] = make_float4(dc[0], dc[1], dc[2], vc[0]);
        m1[u] = make_float2(vc[1], vc[2]);
      }
      if (MODE == PL_LAST) continue;
      if (MODE == PL_WRITE) {
#pragma unroll
        for (int u = 0; u < AL_U; ++u) {
          const uint32_t q = q0 + u * AL_TPB + threadIdx.x;
          if (q < cnt)
            al_put(out.st, out.start[out.ptile[start + q]] + out.prank[start + q], n0[u], n1[u],
                   m0[u], m1[u]);
        }
        continue;
      }
      if (MODE == PL_COUNT) {
#pragma unroll
        for (int u = 0; u < AL_U; ++u) {
          const uint32_t q = q0 + u * AL_TPB + threadIdx.x;
          const uint32_t r = agg_slot(out.count, nt[u], q < cnt);
          if (q < cnt) {
            out.ptile[start + q] = nt[u];
            out.prank[start + q] = r;
          }
        }
        continue;
      }
      uint32_t rk[AL_U];
#pragma unroll
      for (int u = 0; u < AL_U; ++u) {
        const uint32_t q = q0 + u * AL_TPB + threadIdx.x;
        const bool near = q < cnt && code[u] < 27;
        const uint32_t rl = agg_slot(s_cnt, near ? (uint32_t)code[u] : 0u, near);
        const uint32_t rf = agg_slot(out.count, nt[u], q < cnt && code[u] == 27);
        rk[u] = near ? rl : rf;
      }
      __syncthreads();
      pl_round_bases(s_cnt, s_base, ti, tg, out.count);
      __syncthreads();
#pragma unroll
      for (int u = 0; u < AL_U; ++u) {
        const uint32_t q = q0 + u * AL_TPB + threadIdx.x;
        if (q >= cnt) continue;
        const uint32_t r = code[u] < 27 ? s_base[code[u]] + rk[u] : rk[u];
        const uint32_t b = out.start[nt[u]], e = out.start[nt[u] + 1];
        if (r < e - b)
          al_put(out.st, b + r, n0[u], n1[u], m0[u], m1[u]);
        else
          out.flag[0] = 1u;
      }
    }
  };
  if (MODE == PL_COUNT || MODE == PL_WRITE) {
    const unsigned ntl = (unsigned)tg.ntx * tg.nty * tg.ntz;
    for (unsigned tb = blockIdx.x; tb < ntl; tb += gridDim.x) {
      tile((int)pmx_xcd_block(tb, ntl));
      __syncthreads();
    }
  } else {
    tile((int)pmx_xcd_block(blockIdx.x, gridDim.x));
  }
  if (MODE == PL_CAP) {
#pragma unroll
    for (int a = 0; a < 3; ++a) wave_max_store(vm[a], vmax + a);
    block_sum2_store<AL_TPB>(fsum, dsum, pf, pd);
  }
  if (MODE == PL_LAST) block_sum_store<AL_TPB>(fsum, pf);
}

template <int MODE>
static void al_launch(pmx_ctx* c, int src, float val, float kp, float ka, float dd,
                      const float* snap, AlOut out, AlUser user) {
  const MeshGeom g = pmx_geom_padded(c);
  const TileGeom& tg = c->tg;
  const size_t lds = (size_t)4 * (tg.tx + 1) * (tg.ty + 1) * (tg.tz + 1) * sizeof(float);
  const bool rebin = MODE == PL_COUNT || MODE == PL_WRITE;
  const unsigned grid = rebin && c->ntiles > PL_REBIN_GRID ? PL_REBIN_GRID : (unsigned)c->ntiles;
#define AL_LAUNCH(F)                                                                            \
  al_step_k<MODE, F><<<grid, AL_TPB, lds, c->stream>>>(                                          \
      c->al[src], c->pl_start[src], c->pl_count[src], g, tg, (const float*)c->cbuf[0],           \
      (const float*)c->cbuf[1], (const float*)c->cbuf[2], (const float*)c->abuf[0], val, kp, ka, \
      dd, snap, out, user, c->vmax, c->red, c->red + c->ntiles)
  if (pmx_full16(c))
    AL_LAUNCH(1);
  else
    AL_LAUNCH(0);
#undef AL_LAUNCH
}

extern "C" int pmx_nbody_steps_adj(pmx_ctx* c, const int16_t* pmid, float* disp, float* vel,
                                   float* acc, float* disp_cot, float* vel_cot, float* acc_cot,
                                   long long n, float val, float coef, float kp, int nsteps,
                                   const float* ka, const float* d, const float* kb,
                                   const float* disp_snap, double* dots) {
  PMX_ENTRY(c);
  PMX_REQUIRE(c && n >= 0 && nsteps >= 0, "pmx_nbody_steps_adj: bad argument");
  PMX_REQUIRE(n == 0 || (pmid && disp && vel && acc && disp_cot && vel_cot && acc_cot),
              "pmx_nbody_steps_adj: NULL pointer");
  PMX_REQUIRE(nsteps == 0 || (ka && d && kb), "pmx_nbody_steps_adj: NULL factor array");
  PMX_REQUIRE(n < (1LL << 32) - 2048, "pmx_nbody_steps_adj: too many particles");
  PMX_REQUIRE(!disp_snap || disp_snap != disp || nsteps == 1,
              "pmx_nbody_steps_adj: disp_snap may alias disp only for a single step");
  if (nsteps == 0) return PMX_OK;
  int rc;
  // the dot products are read back on the host
  if (dots && (rc = pmx_no_capture(c, "pmx_nbody_steps_adj with dots"))) return rc;
  if (n == 0) {
    if (dots)
      for (int s = 0; s < 2 * nsteps; ++s) dots[s] = 0;
    return PMX_OK;
  }
  if (c->n[0] > PMX_PL_MAXN || c->n[1] > PMX_PL_MAXN || c->n[2] > PMX_PL_MAXN) {
    // a mesh axis longer than the tile-ordered state packs: one reverse step
    // at a time, each closing kick pending into the next
    float kpi = kp;
    for (int s = 0; s < nsteps; ++s) {
      double dd[2];
      if ((rc = pmx_nbody_step_adj(c, pmid, disp, vel, acc, disp_cot, vel_cot, acc_cot, n, val,
                                   coef, kpi, ka[s], d[s],
                                   s + 1 == nsteps ? disp_snap : nullptr, dd)))
        return rc;
      if (dots) {
        dots[2 * s] = dd[0];
        dots[2 * s + 1] = dd[1];
      }
      kpi = kb[s];
    }
    return PMX_OK;
  }
  // every allocation (the three-channel halo shells of the cotangent
  // deposits and the scan scratch included) before the first kernel writes
  // the caller's arrays, so that PMX_ENOMEM leaves them as they were
  if ((rc = pmx_ensure_ptcl(c, n))) return rc;
  if ((rc = pmx_ensure_adj_state(c, nsteps))) return rc;
  PMX_REQUIRE(c->rec_cap < (1LL << 32) - 1, "pmx_nbody_steps_adj: too many particles");
  const long long nt = c->ntiles;
  const unsigned nb1 = pmx_blocks(n);
  if ((rc = pmx_ensure_red(c, nb1 > 2 * nt ? (long long)nb1 : 2 * nt))) return rc;
  if ((rc = pmx_ensure_halo(c, 3))) return rc;
  if ((rc = pmx_ensure_scan(c, nt + 1))) return rc;
  const MeshGeom g = pmx_geom_padded(c);
  const float adj_scale = (float)(-(double)coef / (double)c->M);
  uint32_t* vmax_u = reinterpret_cast<uint32_t*>(c->vmax);
  int A = 0;

  // first step's closing kick of the step before (kp), opening kick and drift
  // adjoints on the caller's arrays, exact binning, fill copy A
  if ((rc = pmx_fill_u32(c, c->pl_count[A], nt + 1, 0u))) return rc;
  const BinOut bo = {c->pkey, c->prank, nullptr, nullptr, nullptr, c->flags};
  if ((rc = pmx_kd_adj_bin(c, pmid, disp, vel, acc, disp_cot, vel_cot, acc_cot,
                            nsteps == 1 ? disp_snap : nullptr, n, kp, ka[0], d[0],
                            c->pl_count[A], bo, 0, c->red)))
    return rc;
  if ((rc = pmx_red_reduce_to(c, 0, nb1, c->al_dots))) return rc;
  pmx_tic(c, PMX_T_SCAN);
  rc = pmx_scan_exclusive(c, c->pl_count[A], c->pl_start[A], nt + 1);
  pmx_toc(c, PMX_T_SCAN);
  if (rc) return rc;
  if ((rc = pmx_fill_u32(c, vmax_u, 3, 0u))) return rc;
  pmx_tic(c, PMX_T_FILL);
  al_fill_k<<<nb1, PMX_TPB, 0, c->stream>>>(pmid, disp, vel, disp_cot, vel_cot, n, g, c->pkey,
                                            c->prank, c->pl_start[A], c->al[A], c->vmax);
  pmx_toc(c, PMX_T_FILL);
  PMX_LAUNCH_CHECK();

  const AlUser user = {disp, vel, acc, disp_cot, vel_cot, acc_cot};
  for (int s = 0; s < nsteps; ++s) {
    // the force meshes at this step's positions (kept in cbuf)
    if ((rc = pmx_deposit_state(c, c->al[A].p, c->pl_start[A], c->pl_count[A], val,
                                (float*)c->cbuf[0], c->rs, 1)))
      return rc;
    if ((rc = pmx_poisson_ctx(c, coef, 1))) return rc;
    // the cotangents of the three force meshes, then of the density (abuf[0])
    DepOut out3 = {{(float*)c->abuf[0], (float*)c->abuf[1], (float*)c->abuf[2]}};
    const float* cv0 = reinterpret_cast<const float*>(c->al[A].c0);
    const float* cv1 = reinterpret_cast<const float*>(c->al[A].c1);
    const ValSrc vc = {{cv0 + 3, cv1, cv1 + 1}, {4, 2, 2}};  // vel_cot x, y, z
    if ((rc = pmx_deposit_state_vals(c, c->al[A].p, c->pl_start[A], c->pl_count[A], 0.f, &vc, 3,
                                     out3, c->rs, 1)))
      return rc;
    if ((rc = pmx_poisson_adj_bufs(c, c->abuf, adj_scale, 1))) return rc;
    if (s + 1 == nsteps) {
      pmx_tic(c, PMX_T_GRADG);
      al_launch<PL_LAST>(c, A, val, 0.f, 0.f, 0.f, nullptr, AlOut{}, user);
      pmx_toc(c, PMX_T_GRADG);
      PMX_LAUNCH_CHECK();
      if ((rc = pmx_red_reduce_to(c, 0, nt, c->al_dots + 2 * s + 1))) return rc;
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
    if ((rc = pmx_fill_u32(c, vmax_u, 3, 0u))) return rc;
    const AlOut out = {c->al[B], c->pl_count[B], c->pl_start[B], c->pkey, c->prank, c->flags};
    const float* snap = s + 2 == nsteps ? disp_snap : nullptr;
    pmx_tic(c, PMX_T_GRADG);
    al_launch<PL_CAP>(c, A, val, kb[s], ka[s + 1], d[s + 1], snap, out, user);
    pmx_toc(c, PMX_T_GRADG);
    PMX_LAUNCH_CHECK();
    if ((rc = pmx_red_reduce_to(c, 0, nt, c->al_dots + 2 * s + 1))) return rc;
    if ((rc = pmx_red_reduce_to(c, nt, nt, c->al_dots + 2 * (s + 1)))) return rc;
    // a tile overflowed: redo the move exactly (skipped on the device otherwise)
    if ((rc = pmx_fill_u32(c, c->pl_count[B], nt + 1, 0u, c->flags))) return rc;
    pmx_tic(c, PMX_T_BIN);
    al_launch<PL_COUNT>(c, A, val, kb[s], ka[s + 1], d[s + 1], snap, out, user);
    PMX_LAUNCH_CHECK();
    if ((rc = pmx_scan_exclusive(c, c->pl_count[B], c->pl_start[B], nt + 1, c->flags))) return rc;
    al_launch<PL_WRITE>(c, A, val, kb[s], ka[s + 1], d[s + 1], snap, out, user);
    pmx_toc(c, PMX_T_BIN);
    PMX_LAUNCH_CHECK();
    A = B;
  }
  if (!dots) return PMX_OK;
  PMX_CHECK_HIP(hipMemcpyAsync(dots, c->al_dots, (size_t)2 * nsteps * sizeof(double),
                               hipMemcpyDeviceToHost, c->stream));
  PMX_CHECK_HIP(hipStreamSynchronize(c->stream));
  return PMX_OK;
}
