// Generic CIC scatter/gather and adjoints for non-default arguments:
// offsets, a different output cell size (enmesh's a2 branch, computed in
// float64 as the reference does), target shapes with dropped indices,
// channels and 1-3 spatial dimensions.  Not the hot path: global float
// atomics are used for the scatter-type outputs.
#include "pmx_internal.h"

#include <math.h>

struct GenGeom {
  int dim;
  int s1[3];      // periodic shape
  float a1f;      // float32 cell (a2=None branch)
  double a1;      // float64 cell (a2 branch)
  int use_a2;
  double a2;
  float a2f;
  double b12[3];
  int ioff[3];    // int16(i12) of divmod(b12, a1) (a2=None branch)
  float doff[3];  // float32(d12)
  int s2[3];      // target shape (indices outside are dropped)
  float gcell;    // gradient divisor (a2 if given, else a1), float32
  int wrap;       // 1: wrap around the periodic s1 (enmesh's s1 given)
};

// JAX float divmod (lax.rem + sign fix, round of the quotient)
__device__ __forceinline__ void fdivmod(double x, double y, double* q, double* m) {
  double mod = fmod(x, y);
  double div = (x - mod) / y;
  bool ind = (mod != 0.0) && ((y < 0.0) != (mod < 0.0));
  if (ind) {
    mod += y;
    div -= 1.0;
  }
  *q = round(div);
  *m = mod;
}

struct GenAxis {
  int idx[2];
  float f[2];
  float s[2];
};

__device__ __forceinline__ GenAxis gen_axis(int pm, float dp, const GenGeom& G, int a) {
  GenAxis r;
  if (a >= G.dim) {  // padding axis of a lower-dimensional problem
    r.idx[0] = r.idx[1] = 0;
    r.f[0] = 1.f; r.f[1] = 0.f;
    r.s[0] = r.s[1] = 0.f;
    return r;
  }
  if (!G.use_a2) {
    int i1 = pm - G.ioff[a];
    float d1 = dp - G.doff[a];
    float d = d1 / G.a1f;
    int i = (int)floorf(d);
    for (int nb = 0; nb < 2; ++nb) {
      float d2 = d - (float)(i + nb);
      r.f[nb] = 1.f - fabsf(d2);
      r.s[nb] = pmx_sign(-d2);
      // int16 sum, as enmesh's index dtype
      r.idx[nb] = G.wrap ? pmx_wrap(i + nb + i1, G.s1[a]) : (int)(short)(i + nb + i1);
    }
  } else {
    double P = (double)pm * G.a1 + (double)dp - G.b12[a];
    double L = (double)G.s1[a] * G.a1;
    for (int nb = 0; nb < 2; ++nb) {
      double x = P + (double)nb * G.a2;
      double q, m;
      if (G.wrap) {
        fdivmod(x, L, &q, &m);
        x = m;
      }
      fdivmod(x, G.a2, &q, &m);
      double d2 = P - q * G.a2;
      if (G.wrap) d2 -= rint(d2 / L) * L;
      int ii = (int)(short)(long long)q;
      float d2f = (float)d2;
      d2f = d2f / G.a2f;
      r.f[nb] = 1.f - fabsf(d2f);
      r.s[nb] = pmx_sign(-d2f);
      r.idx[nb] = ii;
    }
  }
  return r;
}

// flat cell index of neighbour j or -1 when dropped
__device__ __forceinline__ long long gen_cell(const GenAxis* ax, int j, const GenGeom& G) {
  long long c = 0;
  for (int a = 0; a < 3; ++a) {
    int i = ax[a].idx[(j >> a) & 1];
    if (i < 0 || i >= G.s2[a]) return -1;
    c = c * G.s2[a] + i;
  }
  return c;
}

__device__ __forceinline__ float gen_w(const GenAxis* ax, int j, int dim) {
  float w = ax[0].f[j & 1];
  if (dim > 1) w = w * ax[1].f[(j >> 1) & 1];
  if (dim > 2) w = w * ax[2].f[(j >> 2) & 1];
  return w;
}

// grad of w_j, product order of pm_util.py:147-150
__device__ __forceinline__ void gen_wgrad(const GenAxis* ax, int j, int dim, float g[3]) {
  float f[3], s[3];
  for (int a = 0; a < 3; ++a) {
    f[a] = ax[a].f[(j >> a) & 1];
    s[a] = ax[a].s[(j >> a) & 1];
  }
  if (dim == 1) {
    g[0] = s[0];
    g[1] = g[2] = 0.f;
  } else if (dim == 2) {
    g[0] = s[0] * f[1];
    g[1] = s[1] * f[0];
    g[2] = 0.f;
  } else {
    g[0] = s[0] * (f[1] * f[2]);
    g[1] = s[1] * (f[2] * f[0]);
    g[2] = s[2] * (f[0] * f[1]);
  }
}

// pairwise tree over 2^dim terms
__device__ __forceinline__ float gen_tree(const float* v, int nn) {
  if (nn == 2) return v[0] + v[1];
  if (nn == 4) return (v[0] + v[1]) + (v[2] + v[3]);
  return ((v[0] + v[1]) + (v[2] + v[3])) + ((v[4] + v[5]) + (v[6] + v[7]));
}

__device__ __forceinline__ void gen_load(const int16_t* pmid, const float* disp, long long i, int dim,
                                         int pm[3], float dp[3]) {
  for (int a = 0; a < 3; ++a) {
    pm[a] = a < dim ? pmid[dim * i + a] : 0;
    dp[a] = a < dim ? disp[dim * i + a] : 0.f;
  }
}

__global__ void gen_scatter_k(GenGeom G, const int16_t* __restrict__ pmid,
                              const float* __restrict__ disp, long long n,
                              const float* __restrict__ val, float vs, int nchan,
                              float* __restrict__ mesh) {
  long long i = (long long)blockIdx.x * PMX_TPB + threadIdx.x;
  if (i >= n) return;
  int pm[3];
  float dp[3];
  gen_load(pmid, disp, i, G.dim, pm, dp);
  GenAxis ax[3];
  for (int a = 0; a < 3; ++a) ax[a] = gen_axis(pm[a], dp[a], G, a);
  int nn = 1 << G.dim;
  for (int j = 0; j < nn; ++j) {
    long long c = gen_cell(ax, j, G);
    if (c < 0) continue;
    float w = gen_w(ax, j, G.dim);
    for (int ch = 0; ch < nchan; ++ch) {
      float v = val ? val[i * nchan + ch] : vs;
      atomicAdd(&mesh[c * nchan + ch], v * w);
    }
  }
}

__global__ void gen_gather_k(GenGeom G, const int16_t* __restrict__ pmid,
                             const float* __restrict__ disp, long long n,
                             const float* __restrict__ mesh, int nchan,
                             const float* __restrict__ val, float vs, float* __restrict__ out) {
  long long i = (long long)blockIdx.x * PMX_TPB + threadIdx.x;
  if (i >= n) return;
  int pm[3];
  float dp[3];
  gen_load(pmid, disp, i, G.dim, pm, dp);
  GenAxis ax[3];
  for (int a = 0; a < 3; ++a) ax[a] = gen_axis(pm[a], dp[a], G, a);
  int nn = 1 << G.dim;
  long long cs[8];
  float ws[8];
  for (int j = 0; j < nn; ++j) {
    cs[j] = gen_cell(ax, j, G);
    ws[j] = gen_w(ax, j, G.dim);
  }
  for (int ch = 0; ch < nchan; ++ch) {
    float t[8];
    for (int j = 0; j < nn; ++j) t[j] = (cs[j] >= 0 ? mesh[cs[j] * nchan + ch] : 0.f) * ws[j];
    float v = val ? val[i * nchan + ch] : vs;
    out[i * nchan + ch] = v + gen_tree(t, nn);
  }
}

// scatter adjoint (scatter.py:86-121)
__global__ void gen_scatter_adj_k(GenGeom G, const int16_t* __restrict__ pmid,
                                  const float* __restrict__ disp, long long n,
                                  const float* __restrict__ mesh_cot, const float* __restrict__ val,
                                  float vs, int nchan, float* __restrict__ disp_cot,
                                  float* __restrict__ val_cot) {
  long long i = (long long)blockIdx.x * PMX_TPB + threadIdx.x;
  if (i >= n) return;
  int pm[3];
  float dp[3];
  gen_load(pmid, disp, i, G.dim, pm, dp);
  GenAxis ax[3];
  for (int a = 0; a < 3; ++a) ax[a] = gen_axis(pm[a], dp[a], G, a);
  int nn = 1 << G.dim;
  long long cs[8];
  for (int j = 0; j < nn; ++j) cs[j] = gen_cell(ax, j, G);
  if (disp_cot) {
    float gd[3][8];
    for (int j = 0; j < nn; ++j) {
      float dc = 0.f;
      for (int ch = 0; ch < nchan; ++ch) {
        float mc = cs[j] >= 0 ? mesh_cot[cs[j] * nchan + ch] : 0.f;
        float v = val ? val[i * nchan + ch] : vs;
        dc = dc + mc * v;
      }
      float g[3];
      gen_wgrad(ax, j, G.dim, g);
      for (int a = 0; a < 3; ++a) gd[a][j] = dc * g[a];
    }
    for (int a = 0; a < G.dim; ++a) disp_cot[i * G.dim + a] = gen_tree(gd[a], nn) / G.gcell;
  }
  if (val_cot) {
    for (int ch = 0; ch < nchan; ++ch) {
      float t[8];
      for (int j = 0; j < nn; ++j)
        t[j] = (cs[j] >= 0 ? mesh_cot[cs[j] * nchan + ch] : 0.f) * gen_w(ax, j, G.dim);
      val_cot[i * nchan + ch] = gen_tree(t, nn);
    }
  }
}

// gather adjoint (gather.py:80-116)
__global__ void gen_gather_adj_k(GenGeom G, const int16_t* __restrict__ pmid,
                                 const float* __restrict__ disp, long long n,
                                 const float* __restrict__ mesh, const float* __restrict__ val_cot,
                                 int nchan, float* __restrict__ disp_cot,
                                 float* __restrict__ mesh_cot) {
  long long i = (long long)blockIdx.x * PMX_TPB + threadIdx.x;
  if (i >= n) return;
  int pm[3];
  float dp[3];
  gen_load(pmid, disp, i, G.dim, pm, dp);
  GenAxis ax[3];
  for (int a = 0; a < 3; ++a) ax[a] = gen_axis(pm[a], dp[a], G, a);
  int nn = 1 << G.dim;
  long long cs[8];
  for (int j = 0; j < nn; ++j) cs[j] = gen_cell(ax, j, G);
  if (disp_cot) {
    float gd[3][8];
    for (int j = 0; j < nn; ++j) {
      float dc = 0.f;
      for (int ch = 0; ch < nchan; ++ch) {
        float m = cs[j] >= 0 ? mesh[cs[j] * nchan + ch] : 0.f;
        dc = dc + val_cot[i * nchan + ch] * m;
      }
      float g[3];
      gen_wgrad(ax, j, G.dim, g);
      for (int a = 0; a < 3; ++a) gd[a][j] = dc * g[a];
    }
    for (int a = 0; a < G.dim; ++a) disp_cot[i * G.dim + a] = gen_tree(gd[a], nn) / G.gcell;
  }
  if (mesh_cot) {
    for (int j = 0; j < nn; ++j) {
      if (cs[j] < 0) continue;
      float w = gen_w(ax, j, G.dim);
      for (int ch = 0; ch < nchan; ++ch)
        atomicAdd(&mesh_cot[cs[j] * nchan + ch], val_cot[i * nchan + ch] * w);
    }
  }
}

// host-side float64 divmod matching fdivmod
static void h_fdivmod(double x, double y, double* q, double* m) {
  double mod = fmod(x, y);
  double div = (x - mod) / y;
  bool ind = (mod != 0.0) && ((y < 0.0) != (mod < 0.0));
  if (ind) {
    mod += y;
    div -= 1.0;
  }
  *q = round(div);
  *m = mod;
}

static int make_gen(const pmx_cic_geom* g, GenGeom* G) {
  PMX_REQUIRE(g != nullptr, "pmx_cic_geom is NULL");
  int dim = 3;
  while (dim > 1 && g->mesh_shape[dim - 1] == 0) --dim;
  G->dim = dim;
  for (int a = 0; a < 3; ++a) {
    bool pad = a >= dim;
    G->s1[a] = pad ? 1 : g->mesh_shape[a];
    G->s2[a] = pad ? 1 : g->out_shape[a];
    PMX_REQUIRE(G->s1[a] > 0 && G->s2[a] > 0, "pmx_cic_geom: shapes must be positive");
    G->b12[a] = pad ? 0.0 : g->offset[a];
  }
  PMX_REQUIRE(g->cell_size > 0, "pmx_cic_geom: cell_size must be positive");
  G->a1 = g->cell_size;
  G->a1f = (float)g->cell_size;
  G->use_a2 = g->out_cell_size > 0;
  G->a2 = G->use_a2 ? g->out_cell_size : 0.0;
  G->a2f = (float)G->a2;
  G->gcell = G->use_a2 ? G->a2f : G->a1f;
  G->wrap = 1;
  for (int a = 0; a < 3; ++a) {
    double q, m;
    h_fdivmod(G->b12[a], (double)G->a1f, &q, &m);
    G->ioff[a] = (int)(short)(long long)q;
    G->doff[a] = (float)m;
  }
  return PMX_OK;
}

extern "C" int pmx_scatter_cic(pmx_ctx* c, const pmx_cic_geom* g, const int16_t* pmid,
                               const float* disp, long long n, const float* val, float val_scalar,
                               int nchan, float* mesh) {
  PMX_ENTRY(c);
  PMX_REQUIRE(c && n >= 0 && nchan >= 1 && mesh, "pmx_scatter_cic: bad argument");
  GenGeom G;
  int rc = make_gen(g, &G);
  if (rc) return rc;
  if (n == 0) return PMX_OK;
  PMX_REQUIRE(pmid && disp, "pmx_scatter_cic: NULL particle arrays");
  gen_scatter_k<<<pmx_blocks(n), PMX_TPB, 0, c->stream>>>(G, pmid, disp, n, val, val_scalar, nchan,
                                                          mesh);
  PMX_LAUNCH_CHECK();
  return PMX_OK;
}

extern "C" int pmx_gather_cic(pmx_ctx* c, const pmx_cic_geom* g, const int16_t* pmid,
                              const float* disp, long long n, const float* mesh, int nchan,
                              const float* val, float val_scalar, float* out) {
  PMX_ENTRY(c);
  PMX_REQUIRE(c && n >= 0 && nchan >= 1, "pmx_gather_cic: bad argument");
  GenGeom G;
  int rc = make_gen(g, &G);
  if (rc) return rc;
  if (n == 0) return PMX_OK;
  PMX_REQUIRE(pmid && disp && mesh && out, "pmx_gather_cic: NULL pointer");
  gen_gather_k<<<pmx_blocks(n), PMX_TPB, 0, c->stream>>>(G, pmid, disp, n, mesh, nchan, val,
                                                         val_scalar, out);
  PMX_LAUNCH_CHECK();
  return PMX_OK;
}

extern "C" int pmx_scatter_cic_adj(pmx_ctx* c, const pmx_cic_geom* g, const int16_t* pmid,
                                   const float* disp, long long n, const float* mesh_cot,
                                   const float* val, float val_scalar, int nchan, float* disp_cot,
                                   float* val_cot) {
  PMX_ENTRY(c);
  PMX_REQUIRE(c && n >= 0 && nchan >= 1, "pmx_scatter_cic_adj: bad argument");
  GenGeom G;
  int rc = make_gen(g, &G);
  if (rc) return rc;
  if (n == 0) return PMX_OK;
  PMX_REQUIRE(pmid && disp && mesh_cot, "pmx_scatter_cic_adj: NULL pointer");
  gen_scatter_adj_k<<<pmx_blocks(n), PMX_TPB, 0, c->stream>>>(G, pmid, disp, n, mesh_cot, val,
                                                              val_scalar, nchan, disp_cot, val_cot);
  PMX_LAUNCH_CHECK();
  return PMX_OK;
}

extern "C" int pmx_gather_cic_adj(pmx_ctx* c, const pmx_cic_geom* g, const int16_t* pmid,
                                  const float* disp, long long n, const float* mesh,
                                  const float* val_cot, int nchan, float* disp_cot,
                                  float* mesh_cot) {
  PMX_ENTRY(c);
  PMX_REQUIRE(c && n >= 0 && nchan >= 1, "pmx_gather_cic_adj: bad argument");
  GenGeom G;
  int rc = make_gen(g, &G);
  if (rc) return rc;
  if (n == 0) return PMX_OK;
  PMX_REQUIRE(pmid && disp && mesh && val_cot, "pmx_gather_cic_adj: NULL pointer");
  gen_gather_adj_k<<<pmx_blocks(n), PMX_TPB, 0, c->stream>>>(G, pmid, disp, n, mesh, val_cot, nchan,
                                                             disp_cot, mesh_cot);
  PMX_LAUNCH_CHECK();
  return PMX_OK;
}

// enmesh itself (pm_util.py:33-156): indices (n, 2^dim, dim) of the index
// dtype int16, fractions (n, 2^dim) and their gradients (n, 2^dim, dim).
// Without wrap, negative indices are replaced by the target extent when drop
// is set (pm_util.py:138-141), so that a later .at[] drops them.
__global__ void gen_enmesh_k(GenGeom G, int drop, const int16_t* __restrict__ pmid,
                             const float* __restrict__ disp, long long n, int16_t* __restrict__ ind,
                             float* __restrict__ frac, float* __restrict__ frac_grad) {
  long long i = (long long)blockIdx.x * PMX_TPB + threadIdx.x;
  if (i >= n) return;
  int pm[3];
  float dp[3];
  gen_load(pmid, disp, i, G.dim, pm, dp);
  GenAxis ax[3];
  for (int a = 0; a < 3; ++a) ax[a] = gen_axis(pm[a], dp[a], G, a);
  const int nn = 1 << G.dim;
  for (int j = 0; j < nn; ++j) {
    for (int a = 0; a < G.dim; ++a) {
      int v = ax[a].idx[(j >> a) & 1];
      if (!G.wrap && drop && v < 0) v = G.s2[a];
      ind[(i * nn + j) * G.dim + a] = (int16_t)v;
    }
    frac[i * nn + j] = gen_w(ax, j, G.dim);
    if (frac_grad) {
      float g[3];
      gen_wgrad(ax, j, G.dim, g);
      for (int a = 0; a < G.dim; ++a) frac_grad[(i * nn + j) * G.dim + a] = g[a];
    }
  }
}

extern "C" int pmx_enmesh(pmx_ctx* c, const pmx_cic_geom* g, int wrap, int drop,
                          const int16_t* pmid, const float* disp, long long n, int16_t* ind,
                          float* frac, float* frac_grad) {
  PMX_ENTRY(c);
  PMX_REQUIRE(c && n >= 0, "pmx_enmesh: bad argument");
  GenGeom G;
  int rc = make_gen(g, &G);
  if (rc) return rc;
  G.wrap = wrap != 0;
  if (n == 0) return PMX_OK;
  PMX_REQUIRE(pmid && disp && ind && frac, "pmx_enmesh: NULL pointer");
  gen_enmesh_k<<<pmx_blocks(n), PMX_TPB, 0, c->stream>>>(G, drop != 0, pmid, disp, n, ind, frac,
                                                         frac_grad);
  PMX_LAUNCH_CHECK();
  return PMX_OK;
}

// Three-component readout of the default-argument CIC (gather.py:33-77 with
// val=0, offset=0, cell_size=None) from component-major meshes
// (3, n0, n1, n2): acc[i, d] = sum_j mesh_d[c_j] w_j, the same sum order as
// three pmx_gather_cic calls.
__global__ void gen_gather3_k(GenGeom G, const int16_t* __restrict__ pmid,
                              const float* __restrict__ disp, long long n,
                              const float* __restrict__ mesh3, long long M,
                              float* __restrict__ acc) {
  long long i = (long long)blockIdx.x * PMX_TPB + threadIdx.x;
  if (i >= n) return;
  int pm[3];
  float dp[3];
  gen_load(pmid, disp, i, 3, pm, dp);
  GenAxis ax[3];
  for (int a = 0; a < 3; ++a) ax[a] = gen_axis(pm[a], dp[a], G, a);
  long long cs[8];
  float ws[8];
  for (int j = 0; j < 8; ++j) {
    cs[j] = gen_cell(ax, j, G);
    ws[j] = gen_w(ax, j, 3);
  }
  for (int d = 0; d < 3; ++d) {
    float t[8];
    for (int j = 0; j < 8; ++j) t[j] = (cs[j] >= 0 ? mesh3[d * M + cs[j]] : 0.f) * ws[j];
    acc[3 * i + d] = 0.f + gen_tree(t, 8);
  }
}

extern "C" int pmx_gather_cic3(pmx_ctx* c, const int16_t* pmid, const float* disp, long long n,
                               const float* mesh3, float* acc) {
  PMX_ENTRY(c);
  PMX_REQUIRE(c && n >= 0, "pmx_gather_cic3: bad argument");
  if (n == 0) return PMX_OK;
  PMX_REQUIRE(pmid && disp && mesh3 && acc, "pmx_gather_cic3: NULL pointer");
  pmx_cic_geom g = {};
  for (int a = 0; a < 3; ++a) g.mesh_shape[a] = g.out_shape[a] = c->n[a];
  g.cell_size = c->cell;
  GenGeom G;
  int rc = make_gen(&g, &G);
  if (rc) return rc;
  gen_gather3_k<<<pmx_blocks(n), PMX_TPB, 0, c->stream>>>(G, pmid, disp, n, mesh3, c->M, acc);
  PMX_LAUNCH_CHECK();
  return PMX_OK;
}
