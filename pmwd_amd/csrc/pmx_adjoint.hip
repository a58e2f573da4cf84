// Reverse-time adjoint of the KDK step (gfx950): the force VJP
// (pmwd/nbody.py:108-118 over gravity.py:47-72, i.e. gather.py:80-144 and
// scatter.py:86-150 around the adjoint Poisson solve), the reverse kick and
// drift (nbody.py:49-99), and one fused reverse step (nbody.py:143-162)
// whose closing kick is deferred into the next step's opening kernel, where
// it streams in particle order.
//
// Everything per particle runs over the tile-binned records: one workgroup
// per tile stages the meshes it reads (tile + 1 layer) in LDS, reads its
// records in order and writes each particle's results to its index.  Sums
// for the cosmology cotangent are float64 per-block partials reduced on the
// device in a fixed order.
#include "pmx_internal.h"

#include <math.h>

#define ADJ_TPB 512
#define ADJ_U 4

// ------------------------------------------------ tile-staged VJP gather --
// Per record (particle p) with the three force meshes F staged in LDS:
//   acc_p     = sum_j F[c_j] w_j                         (gather, tree order)
//   acc_cot_p = sum_j (sum_d v_d F_d[c_j]) grad w_j / cell, v = vel_cot_p
//               (the gather adjoint's disp cotangent, gather.py:80-116)
// and the record's three vel_cot components are staged channel-major in
// vrec (slot order) for the cotangent deposit, with their |max| in vmax.
// part[block] = sum over the tile of vel_cot_p . acc_p (float64).
__global__ void __launch_bounds__(ADJ_TPB)
    tile_vjp_gather_k(const float4* __restrict__ rec, const uint32_t* __restrict__ pidx,
                      const uint32_t* __restrict__ tile_start,
                      const uint32_t* __restrict__ tile_count, MeshGeom g, TileGeom tg,
                      const float* __restrict__ F0, const float* __restrict__ F1,
                      const float* __restrict__ F2, const float* __restrict__ vel_cot,
                      float* __restrict__ acc, float* __restrict__ acc_cot,
                      float* __restrict__ vrec, long long vstride, float* __restrict__ vmax,
                      double* __restrict__ part) {
  extern __shared__ float sh[];
  int o[3], s[3], ti[3];
  const int t = (int)pmx_xcd_block(blockIdx.x, gridDim.x);
  tile_origin(t, tg, g, o, s, ti);
  const int LY = s[1] + 1, LZ = s[2] + 1;
  const int ltot = (s[0] + 1) * LY * LZ;
  const float* F[3] = {F0, F1, F2};
  pmx_stage_tile<3, ADJ_TPB>(sh, o, LY, LZ, ltot, g, F);
  __syncthreads();
  const uint32_t start = tile_start[t];
  const uint32_t cnt = tile_count[t];
  double dotv = 0;
  float vm0 = 0.f, vm1 = 0.f, vm2 = 0.f;
  // ADJ_U records per thread per round, loads issued before the arithmetic
  for (uint32_t q0 = 0; q0 < cnt; q0 += ADJ_TPB * ADJ_U) {
    float4 rr[ADJ_U];
    long long pp[ADJ_U];
    float vc[ADJ_U][3];
#pragma unroll
    for (int u = 0; u < ADJ_U; ++u) {
      const uint32_t q = q0 + u * ADJ_TPB + threadIdx.x;
      pp[u] = -1;
      if (q < cnt) {
        rr[u] = rec[start + q];
        pp[u] = pidx[start + q];
      }
    }
#pragma unroll
    for (int u = 0; u < ADJ_U; ++u)
      if (pp[u] >= 0) {
#pragma unroll
        for (int a = 0; a < 3; ++a) vc[u][a] = vel_cot[3 * pp[u] + a];
      }
#pragma unroll
    for (int u = 0; u < ADJ_U; ++u) {
      if (pp[u] < 0) continue;
      const uint32_t q = q0 + u * ADJ_TPB + threadIdx.x;
      const float4 r = rr[u];
      const long long p = pp[u];
      int l[3];
      pmx_unpack_local(r.w, l[0], l[1], l[2]);
      if (l[0] >= s[0] || l[1] >= s[1] || l[2] >= s[2]) continue;
      const float d[3] = {r.x, r.y, r.z};
      const float v0 = vc[u][0], v1 = vc[u][1], v2 = vc[u][2];
      float o[3], ot[3];
      pmx_vjp_one(sh, ltot, LY, LZ, pmx_frac_of_d(d), l, vc[u], g.cell, g.inv_cell, o, ot);
      const float o0 = o[0], o1 = o[1], o2 = o[2];
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        acc[3 * p + a] = o[a];
        acc_cot[3 * p + a] = ot[a];
      }
      vrec[start + q] = v0;
      vrec[vstride + start + q] = v1;
      vrec[2 * vstride + start + q] = v2;
      vm0 = fmaxf(vm0, fabsf(v0));
      vm1 = fmaxf(vm1, fabsf(v1));
      vm2 = fmaxf(vm2, fabsf(v2));
      dotv += (double)v0 * o0 + (double)v1 * o1 + (double)v2 * o2;
    }
  }
  wave_max_store(vm0, vmax + 0);
  wave_max_store(vm1, vmax + 1);
  wave_max_store(vm2, vmax + 2);
  block_sum_store<ADJ_TPB>(dotv, part);
}

// ----------------------------------------- tile-staged gradient gather --
// Per record with the density cotangent R staged in LDS (scatter adjoint,
// scatter.py:86-121, scalar val):
//   acc_cot_p += sum_j (R[c_j] val) grad w_j / cell
__global__ void __launch_bounds__(ADJ_TPB)
    tile_grad_gather_k(const float4* __restrict__ rec, const uint32_t* __restrict__ pidx,
                       const uint32_t* __restrict__ tile_start,
                       const uint32_t* __restrict__ tile_count, MeshGeom g, TileGeom tg,
                       const float* __restrict__ R, float val, float* __restrict__ acc_cot) {
  extern __shared__ float sh[];
  int o[3], s[3], ti[3];
  const int t = (int)pmx_xcd_block(blockIdx.x, gridDim.x);
  tile_origin(t, tg, g, o, s, ti);
  const int LY = s[1] + 1, LZ = s[2] + 1;
  const int ltot = (s[0] + 1) * LY * LZ;
  const float* F[1] = {R};
  pmx_stage_tile<1, ADJ_TPB>(sh, o, LY, LZ, ltot, g, F);
  __syncthreads();
  const uint32_t start = tile_start[t];
  const uint32_t cnt = tile_count[t];
  for (uint32_t q0 = 0; q0 < cnt; q0 += ADJ_TPB * ADJ_U) {
    float4 rr[ADJ_U];
    long long pp[ADJ_U];
    float ac[ADJ_U][3];
#pragma unroll
    for (int u = 0; u < ADJ_U; ++u) {
      const uint32_t q = q0 + u * ADJ_TPB + threadIdx.x;
      pp[u] = -1;
      if (q < cnt) {
        rr[u] = rec[start + q];
        pp[u] = pidx[start + q];
      }
    }
#pragma unroll
    for (int u = 0; u < ADJ_U; ++u)
      if (pp[u] >= 0) {
#pragma unroll
        for (int a = 0; a < 3; ++a) ac[u][a] = acc_cot[3 * pp[u] + a];
      }
#pragma unroll
    for (int u = 0; u < ADJ_U; ++u) {
      if (pp[u] < 0) continue;
      const float4 r = rr[u];
      const long long p = pp[u];
      int l[3];
      pmx_unpack_local(r.w, l[0], l[1], l[2]);
      if (l[0] >= s[0] || l[1] >= s[1] || l[2] >= s[2]) continue;
      const float d[3] = {r.x, r.y, r.z};
      float gr[3];
      pmx_grad_one(sh, LY, LZ, pmx_frac_of_d(d), l, val, g.cell, g.inv_cell, gr);
#pragma unroll
      for (int a = 0; a < 3; ++a) acc_cot[3 * p + a] = ac[u][a] + gr[a];
    }
  }
}

// With the particles binned for the current disp and the force meshes in
// c->cbuf: acc and acc_cot (force VJP of vel_cot) and per-tile partials of
// vel_cot.acc at c->red + red_off (ntiles of them).
static int force_vjp_binned(pmx_ctx* c, long long n, float coef, float val, const float* vel_cot,
                            float* acc, float* acc_cot, long long red_off) {
  int rc;
  if ((rc = pmx_ensure_vrec(c))) return rc;
  if ((rc = pmx_fill_u32(c, reinterpret_cast<uint32_t*>(c->vmax), 3, 0u))) return rc;
  const MeshGeom g = pmx_geom_padded(c);
  const TileGeom& tg = c->tg;
  const unsigned nt = (unsigned)c->ntiles;
  const size_t cells = (size_t)(tg.tx + 1) * (tg.ty + 1) * (tg.tz + 1);
  pmx_tic(c, PMX_T_VJPG);
  tile_vjp_gather_k<<<nt, ADJ_TPB, 3 * cells * sizeof(float), c->stream>>>(
      c->rec, c->pidx, c->tile_start, c->tile_count, g, tg, (const float*)c->cbuf[0],
      (const float*)c->cbuf[1], (const float*)c->cbuf[2], vel_cot, acc, acc_cot, c->vrec,
      c->rec_cap, c->vmax, c->red + red_off);
  pmx_toc(c, PMX_T_VJPG);
  PMX_LAUNCH_CHECK();
  // the cotangent of each force mesh: scatter(vel_cot_d * w)
  float* meshes[3] = {(float*)c->cbuf[0], (float*)c->cbuf[1], (float*)c->cbuf[2]};
  if ((rc = pmx_deposit_binned3(c, n, meshes, c->rs, 1))) return rc;
  // density cotangent -coef * sum_d L_d(mesh_cot_d), L_d = irfftn (-ik_d) laplace rfftn
  if ((rc = pmx_poisson_adj_ctx(c, (float)(-(double)coef / (double)c->M), 1))) return rc;
  pmx_tic(c, PMX_T_GRADG);
  tile_grad_gather_k<<<nt, ADJ_TPB, cells * sizeof(float), c->stream>>>(
      c->rec, c->pidx, c->tile_start, c->tile_count, g, tg, (const float*)c->cbuf[0], val,
      acc_cot);
  pmx_toc(c, PMX_T_GRADG);
  PMX_LAUNCH_CHECK();
  return PMX_OK;
}

extern "C" int pmx_gravity_vjp(pmx_ctx* c, const int16_t* pmid, const float* disp, long long n,
                               float val, float coef, const float* vel_cot, float* acc,
                               float* acc_cot, double* dot) {
  PMX_ENTRY(c);
  PMX_REQUIRE(c && n >= 0, "pmx_gravity_vjp: bad argument");
  PMX_REQUIRE(n == 0 || (pmid && disp && acc && acc_cot && vel_cot), "pmx_gravity_vjp: NULL pointer");
  PMX_REQUIRE(n < (1LL << 32) - 2048, "pmx_gravity_vjp: too many particles");
  int rc;
  if (n == 0) {
    if (dot) *dot = 0;
    return PMX_OK;
  }
  if ((rc = pmx_ensure_red(c, c->ntiles))) return rc;
  if ((rc = pmx_gravity_fwd_meshes(c, pmid, disp, n, val, coef))) return rc;
  if ((rc = force_vjp_binned(c, n, coef, val, vel_cot, acc, acc_cot, 0))) return rc;
  if (!dot) return PMX_OK;
  if ((rc = pmx_red_reduce(c, 0, c->ntiles, 0))) return rc;
  return pmx_red_fetch(c, 1, dot);
}

// ------------------------------------------------ reverse kick / drift --
// kick_adj (nbody.py:80-99): vel += acc*k; disp_cot -= acc_cot*k;
// partials of vel_cot.acc
__global__ void __launch_bounds__(PMX_TPB)
    kick_adj_k(long long n3, float* __restrict__ vel, const float* __restrict__ acc,
               float* __restrict__ disp_cot, const float* __restrict__ acc_cot,
               const float* __restrict__ vel_cot, float k, double* __restrict__ part) {
  double s = 0;
  for (long long i = (long long)blockIdx.x * PMX_TPB + threadIdx.x; i < n3;
       i += (long long)gridDim.x * PMX_TPB) {
    const float a = acc[i];
    vel[i] = vel[i] + a * k;
    disp_cot[i] = disp_cot[i] - acc_cot[i] * k;
    s += (double)vel_cot[i] * a;
  }
  block_sum_store<PMX_TPB>(s, part);
}

// drift_adj (nbody.py:49-67): disp += vel*d; vel_cot -= disp_cot*d;
// partials of disp_cot.vel
__global__ void __launch_bounds__(PMX_TPB)
    drift_adj_k(long long n3, float* __restrict__ disp, const float* __restrict__ vel,
                const float* __restrict__ disp_cot, float* __restrict__ vel_cot, float d,
                double* __restrict__ part) {
  double s = 0;
  for (long long i = (long long)blockIdx.x * PMX_TPB + threadIdx.x; i < n3;
       i += (long long)gridDim.x * PMX_TPB) {
    const float v = vel[i];
    const float dc = disp_cot[i];
    disp[i] = disp[i] + v * d;
    vel_cot[i] = vel_cot[i] - dc * d;
    s += (double)dc * v;
  }
  block_sum_store<PMX_TPB>(s, part);
}

#define ADJ_GRID_CAP 65536

extern "C" int pmx_kick_adj(pmx_ctx* c, long long n, float* vel, const float* acc, float* disp_cot,
                            const float* acc_cot, const float* vel_cot, float k, double* dot) {
  PMX_ENTRY(c);
  PMX_REQUIRE(c && n >= 0, "pmx_kick_adj: bad argument");
  PMX_REQUIRE(n == 0 || (vel && acc && disp_cot && acc_cot && vel_cot), "pmx_kick_adj: NULL pointer");
  if (n == 0) {
    if (dot) *dot = 0;
    return PMX_OK;
  }
  const unsigned nblk = pmx_blocks_capped(3 * n, ADJ_GRID_CAP);
  int rc;
  if ((rc = pmx_ensure_red(c, nblk))) return rc;
  pmx_tic(c, PMX_T_KDADJ);
  kick_adj_k<<<nblk, PMX_TPB, 0, c->stream>>>(3 * n, vel, acc, disp_cot, acc_cot, vel_cot, k, c->red);
  pmx_toc(c, PMX_T_KDADJ);
  PMX_LAUNCH_CHECK();
  if (dot) return pmx_red_finish(c, nblk, dot);
  return PMX_OK;
}

extern "C" int pmx_drift_adj(pmx_ctx* c, long long n, float* disp, const float* vel,
                             const float* disp_cot, float* vel_cot, float d, double* dot) {
  PMX_ENTRY(c);
  PMX_REQUIRE(c && n >= 0, "pmx_drift_adj: bad argument");
  PMX_REQUIRE(n == 0 || (disp && vel && disp_cot && vel_cot), "pmx_drift_adj: NULL pointer");
  if (n == 0) {
    if (dot) *dot = 0;
    return PMX_OK;
  }
  const unsigned nblk = pmx_blocks_capped(3 * n, ADJ_GRID_CAP);
  int rc;
  if ((rc = pmx_ensure_red(c, nblk))) return rc;
  pmx_tic(c, PMX_T_KDADJ);
  drift_adj_k<<<nblk, PMX_TPB, 0, c->stream>>>(3 * n, disp, vel, disp_cot, vel_cot, d, c->red);
  pmx_toc(c, PMX_T_KDADJ);
  PMX_LAUNCH_CHECK();
  if (dot) return pmx_red_finish(c, nblk, dot);
  return PMX_OK;
}

// ------------------------------------------------- fused reverse step --
// Reverse kicks and drift fused with the tile count of the new positions:
//   [vel += acc*kp; disp_cot -= acc_cot*kp]   (HAS_KP: the previous step's
//                                             closing kick_adj, still pending)
//   vel += acc*ka; disp_cot -= acc_cot*ka; disp += vel*d; vel_cot -= disp_cot*d
// (disp replaced by snap when given, a stored checkpoint), with per-block
// partials of disp_cot.vel (the drift's cosmology term).  A kick's own
// vel_cot.acc equals the dot of the force VJP that produced acc.
template <int HAS_KP, bool CAP>
__global__ void __launch_bounds__(PMX_TPB)
    kd_adj_bin_k(const int16_t* __restrict__ pmid, float* __restrict__ disp, float* __restrict__ vel,
                 const float* __restrict__ acc, float* __restrict__ disp_cot,
                 float* __restrict__ vel_cot, const float* __restrict__ acc_cot,
                 const float* __restrict__ snap, long long n, float kp, float ka, float d,
                 MeshGeom g, TileGeom tg, uint32_t* __restrict__ tile_count, BinOut o,
                 double* __restrict__ part, int keep_x) {
  constexpr int T3 = 3 * PMX_TPB;
  // vel | acc | disp | disp_cot | vel_cot | acc_cot, staged through LDS
  __shared__ __align__(16) float sh[6 * T3];
  __shared__ __align__(16) int16_t sp[T3];
  const long long b0 = (long long)blockIdx.x * PMX_TPB;
  const int cnt = (int)(n - b0 < PMX_TPB ? n - b0 : PMX_TPB);
  pmx_stage_in3<6, PMX_TPB>(sh, {{vel, acc, disp, disp_cot, vel_cot, acc_cot}}, b0, cnt);
  pmx_stage_pmid<PMX_TPB>(sp, pmid, b0, cnt);
  __syncthreads();
  const int q = threadIdx.x;
  const bool active = q < cnt;
  double s = 0;
  uint32_t t = 0;
  CicAxis ax[3];
  float dp[3];
  if (active) {
    int pm[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const int e = 3 * q + a;
      pm[a] = sp[e];
      const float ac = sh[T3 + e], acot = sh[5 * T3 + e];
      float v = sh[e];
      float dc = sh[3 * T3 + e];
      if (HAS_KP) {
        v = v + ac * kp;
        dc = dc - acot * kp;
      }
      v = v + ac * ka;
      dc = dc - acot * ka;
      float x = sh[2 * T3 + e];
      // keep_x: disp already holds the stored state the drift would be replaced by
      if (!keep_x) x = x + v * d;
      float vc = sh[4 * T3 + e];
      vc = vc - dc * d;
      s += (double)dc * v;
      if (snap) x = snap[3 * (b0 + q) + a];
      sh[e] = v;
      sh[2 * T3 + e] = x;
      sh[3 * T3 + e] = dc;
      sh[4 * T3 + e] = vc;
      dp[a] = x;
    }
    cic_setup(pm, dp, g, ax);
    t = tile_of(ax, tg);
  }
  const uint32_t r = agg_slot(tile_count, t, active);
  if (active) pmx_bin_emit<CAP>(o, b0 + q, t, r, ax, tg);
  __syncthreads();
  pmx_stage_out3<1, PMX_TPB>({{vel}}, sh, b0, cnt);
  pmx_stage_out3<3, PMX_TPB>({{disp, disp_cot, vel_cot}}, sh + 2 * T3, b0, cnt);
  block_sum_store<PMX_TPB>(s, part);
}

int pmx_kd_adj_bin(pmx_ctx* c, const int16_t* pmid, float* disp, float* vel, const float* acc,
                   float* disp_cot, float* vel_cot, const float* acc_cot, const float* disp_snap,
                   long long n, float kp, float ka, float d, uint32_t* count, const BinOut& o,
                   int cap, double* part) {
  const MeshGeom g = pmx_geom_padded(c);
  const unsigned nb1 = pmx_blocks(n);
  // a snapshot already copied into disp: the drift's disp update is skipped
  const int keep_x = disp_snap != nullptr && disp_snap == disp;
  const float* snap = keep_x ? nullptr : disp_snap;
#define KD_ADJ_LAUNCH(KP, CAPV, KPV)                                                          \
  kd_adj_bin_k<KP, CAPV><<<nb1, PMX_TPB, 0, c->stream>>>(pmid, disp, vel, acc, disp_cot,       \
                                                         vel_cot, acc_cot, snap, n, KPV, ka, \
                                                         d, g, c->tg, count, o, part, keep_x)
  pmx_tic(c, PMX_T_KDADJ);
  if (kp != 0.f) {
    if (cap) KD_ADJ_LAUNCH(1, true, kp);
    else KD_ADJ_LAUNCH(1, false, kp);
  } else {
    if (cap) KD_ADJ_LAUNCH(0, true, 0.f);
    else KD_ADJ_LAUNCH(0, false, 0.f);
  }
#undef KD_ADJ_LAUNCH
  pmx_toc(c, PMX_T_KDADJ);
  PMX_LAUNCH_CHECK();
  return PMX_OK;
}

extern "C" int pmx_nbody_step_adj(pmx_ctx* c, const int16_t* pmid, float* disp, float* vel,
                                  float* acc, float* disp_cot, float* vel_cot, float* acc_cot,
                                  long long n, float val, float coef, float kp, float ka, float d,
                                  const float* disp_snap, double* dots) {
  PMX_ENTRY(c);
  PMX_REQUIRE(c && n >= 0, "pmx_nbody_step_adj: bad argument");
  PMX_REQUIRE(n == 0 || (pmid && disp && vel && acc && disp_cot && vel_cot && acc_cot),
              "pmx_nbody_step_adj: NULL pointer");
  PMX_REQUIRE(n < (1LL << 32) - 2048, "pmx_nbody_step_adj: too many particles");
  if (n == 0) {
    if (dots) dots[0] = dots[1] = 0;
    return PMX_OK;
  }
  int rc, cap;
  const unsigned nb1 = pmx_blocks(n);
  // every allocation (the force VJP's value channels and three-channel halo
  // shells included) before the first kernel writes the caller's arrays, so
  // that PMX_ENOMEM leaves them as they were
  if ((rc = pmx_ensure_red(c, (long long)nb1 + c->ntiles))) return rc;
  if ((rc = pmx_ensure_ptcl(c, n))) return rc;
  if ((rc = pmx_ensure_vrec(c))) return rc;
  if ((rc = pmx_ensure_halo(c, 3))) return rc;
  if ((rc = pmx_ensure_scan(c, c->ntiles + 1))) return rc;
  if ((rc = pmx_bin_begin(c, n, &cap))) return rc;
  const BinOut o = pmx_bin_out(c);
  if ((rc = pmx_kd_adj_bin(c, pmid, disp, vel, acc, disp_cot, vel_cot, acc_cot, disp_snap, n, kp,
                            ka, d, c->tile_count, o, cap, c->red)))
    return rc;
  // disp now holds the binned positions (disp_snap when given)
  if ((rc = pmx_bin_end(c, pmid, disp, n, cap))) return rc;
  if ((rc = pmx_deposit_binned(c, val, n, (float*)c->cbuf[0], c->rs, 0, 1))) return rc;
  if ((rc = pmx_poisson_ctx(c, coef, 1))) return rc;
  if ((rc = force_vjp_binned(c, n, coef, val, vel_cot, acc, acc_cot, nb1))) return rc;
  if (!dots) return PMX_OK;
  if ((rc = pmx_red_reduce(c, 0, nb1, 0))) return rc;
  if ((rc = pmx_red_reduce(c, nb1, c->ntiles, 1))) return rc;
  return pmx_red_fetch(c, 2, dots);
}
