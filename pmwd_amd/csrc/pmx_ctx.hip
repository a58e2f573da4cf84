// Context, errors, scans and reductions of the pmx library.
#include "pmx_internal.h"

#include <float.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <vector>

static thread_local char g_err[1024] = "";

void pmx_set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

extern "C" const char* pmx_last_error(void) { return g_err; }
extern "C" int pmx_version(void) { return 100; }

// Device allocator for contexts created from now on (both NULL: hipMalloc).
static pmx_alloc_fn g_dalloc = nullptr;
static pmx_free_fn g_dfree = nullptr;
static void* g_dalloc_user = nullptr;

extern "C" int pmx_set_allocator(pmx_alloc_fn alloc, pmx_free_fn release, void* user) {
  PMX_REQUIRE((alloc == nullptr) == (release == nullptr),
              "pmx_set_allocator: give both functions or neither");
  g_dalloc = alloc;
  g_dfree = release;
  g_dalloc_user = user;
  return PMX_OK;
}

// With a caller's allocator the context's stream is drained before a block
// is taken or handed back, so a block changes owner only when no queued work
// can still touch it: pmx_ctx_set_stream orders each new stream after the
// previous one, so the current stream's queue ends every stream's work the
// context queued.
int pmx_capturing(pmx_ctx* c) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(c->stream, &cs) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return cs != hipStreamCaptureStatusNone;
}

int pmx_no_capture(pmx_ctx* c, const char* what) {
  if (!pmx_capturing(c)) return PMX_OK;
  pmx_set_error("%s: the context's stream is under graph capture; run the call once before "
                "capturing so that the context holds its scratch, and read host values outside "
                "the capture", what);
  return PMX_EINVAL;
}

void pmx_mark_queued(pmx_ctx* c) {
  if (!c || !c->moved) return;
  if (pmx_capturing(c)) {
    // captured work runs when the graph is launched, which the capturing
    // caller orders
    c->moved_valid = 0;
    return;
  }
  c->moved_valid = hipEventRecord(c->moved, c->stream) == hipSuccess;
  if (!c->moved_valid) (void)hipGetLastError();
}

static int dev_alloc(pmx_ctx* c, void** p, size_t bytes) {
  if (bytes == 0) bytes = 256;
  if (c->no_alloc) {
    *p = nullptr;
    pmx_set_error("device allocation of %zu bytes in an entry point that allocates nothing",
                  bytes);
    return PMX_ENOMEM;
  }
  if (int rc = pmx_no_capture(c, "device allocation")) {
    *p = nullptr;
    return rc;
  }
  if (c->dalloc) {
    *p = c->dalloc(bytes, c->device, (void*)c->stream, c->dalloc_user);
    if (!*p) {
      pmx_set_error("device allocation of %zu bytes failed (caller's allocator)", bytes);
      return PMX_ENOMEM;
    }
    if (hipStreamSynchronize(c->stream) != hipSuccess) {
      (void)hipGetLastError();
      c->dfree(*p, bytes, c->device, (void*)c->stream, c->dalloc_user);
      *p = nullptr;
      pmx_set_error("device allocation: stream synchronisation failed");
      return PMX_EHIP;
    }
    c->bytes += bytes;
    return PMX_OK;
  }
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    pmx_set_error("hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    *p = nullptr;
    // the failure is reported here; clear it so that the next launch check
    // after a caller's fallback does not see it
    (void)hipGetLastError();
    return PMX_ENOMEM;
  }
  c->bytes += bytes;
  return PMX_OK;
}

static void dev_free(pmx_ctx* c, void** p, size_t bytes) {
  if (*p) {
    // a captured graph addresses the context's scratch: it goes with it
    if (c->gexec) {
      (void)hipStreamSynchronize(c->stream);
      (void)hipGraphExecDestroy(c->gexec);
      c->gexec = nullptr;
    }
    if (c->graph) {
      (void)hipGraphDestroy(c->graph);
      c->graph = nullptr;
    }
    if (c->dfree) {
      (void)hipStreamSynchronize(c->stream);
      c->dfree(*p, bytes, c->device, (void*)c->stream, c->dalloc_user);
    } else {
      hipFree(*p);
    }
    c->bytes -= bytes;
  }
  *p = nullptr;
}

int pmx_dev_alloc(pmx_ctx* c, void** p, size_t bytes) { return dev_alloc(c, p, bytes); }
void pmx_dev_free(pmx_ctx* c, void** p, size_t bytes) { dev_free(c, p, bytes); }

static void free_adj_state(pmx_ctx* c);

// Scatter tile extent: 16^3 cells, clipped to the mesh (the 17^3 LDS block
// of 64-bit accumulators is 39 KB, four workgroups per CU).
static TileGeom make_tiles(const int n[3]);

void pmx_ctx_geometry(pmx_ctx* c, int n0, int n1, int n2, double cell_size, hipStream_t stream) {
  c->n[0] = n0; c->n[1] = n1; c->n[2] = n2;
  c->n2h = n2 / 2 + 1;
  // 16 complex = one 128-byte cache line: column blocks of the strided passes
  // and rows of the row passes then start on line boundaries
  c->n2p = c->n2h <= 16 ? c->n2h : (c->n2h + 15) / 16 * 16;
  c->rs = 2 * c->n2p;
  c->M = (long long)n0 * n1 * n2;
  c->Mc = (long long)n0 * n1 * c->n2p;
  c->cell = cell_size;
  c->cellf = (float)cell_size;
  c->stream = stream;
  hipGetDevice(&c->device);
  c->tg = make_tiles(c->n);
  c->ntiles = (long long)c->tg.ntx * c->tg.nty * c->tg.ntz;
}

static TileGeom make_tiles(const int n[3]) {
  TileGeom t;
  t.tx = n[0] < 16 ? n[0] : 16;
  t.ty = n[1] < 16 ? n[1] : 16;
  t.tz = n[2] < 16 ? n[2] : 16;
  t.ntx = (n[0] + t.tx - 1) / t.tx;
  t.nty = (n[1] + t.ty - 1) / t.ty;
  t.ntz = (n[2] + t.tz - 1) / t.tz;
  t.shell = (t.tx + 1) * (t.ty + 1) * (t.tz + 1) - t.tx * t.ty * t.tz;
  // exact for extents <= 16 and mesh indices < 4096 (pmx_tdiv)
  t.mx = ((1u << 20) + t.tx - 1) / t.tx;
  t.my = ((1u << 20) + t.ty - 1) / t.ty;
  t.mz = ((1u << 20) + t.tz - 1) / t.tz;
  return t;
}

// angular wavenumbers as jnp.fft.fftfreq(n) * (2 pi / cell), float64, then
// float32 (pm_util.py:185-197); rfft axis uses rfftfreq.
static void kvec_host(int n, double cell, bool rfft, std::vector<float>& k,
                      std::vector<float>& km) {
  int len = rfft ? n / 2 + 1 : n;
  k.resize(len);
  km.resize(len);
  double period = 2 * M_PI / cell;
  double nyq = M_PI / cell;
  float nyqf = (float)nyq;
  float eps = (float)(nyq * (double)FLT_EPSILON);
  for (int i = 0; i < len; ++i) {
    long long j = i;
    if (!rfft && i >= (n + 1) / 2) j = i - n;  // fftfreq ordering
    double f = (double)j / (double)n;
    float kf = (float)(f * period);
    k[i] = kf;
    bool is_nyq = fabsf(fabsf(kf) - nyqf) <= eps;
    km[i] = is_nyq ? 0.f : kf;
  }
}

extern "C" int pmx_ctx_create(int n0, int n1, int n2, double cell_size, void* stream,
                              pmx_ctx** out) {
  PMX_REQUIRE(out != nullptr, "pmx_ctx_create: out is NULL");
  *out = nullptr;
  PMX_REQUIRE(n0 > 0 && n1 > 0 && n2 > 0, "pmx_ctx_create: mesh shape must be positive");
  PMX_REQUIRE(n0 <= 4096 && n1 <= 4096 && n2 <= 4096,
              "pmx_ctx_create: mesh extent > 4096 is not supported");
  PMX_REQUIRE((long long)n0 * n1 * n2 < (1LL << 31), "pmx_ctx_create: mesh must have < 2^31 cells");
  PMX_REQUIRE(cell_size > 0, "pmx_ctx_create: cell_size must be positive");
  pmx_ctx* c = (pmx_ctx*)calloc(1, sizeof(pmx_ctx));
  if (!c) {
    pmx_set_error("pmx_ctx_create: out of host memory");
    return PMX_ENOMEM;
  }
  pmx_ctx_geometry(c, n0, n1, n2, cell_size, (hipStream_t)stream);
  c->dalloc = g_dalloc;
  c->dfree = g_dfree;
  c->dalloc_user = g_dalloc_user;

  int rc = PMX_OK;
  if (hipEventCreateWithFlags(&c->moved, hipEventDisableTiming) != hipSuccess) {
    c->moved = nullptr;
    pmx_set_error("pmx_ctx_create: hipEventCreate failed");
    rc = PMX_EHIP;
  }
  for (int a = 0; a < 3 && rc == PMX_OK; ++a) {
    std::vector<float> k, km;
    kvec_host(c->n[a], cell_size, a == 2, k, km);
    rc |= dev_alloc(c, (void**)&c->k[a], k.size() * 4);
    rc |= dev_alloc(c, (void**)&c->km[a], km.size() * 4);
    if (rc) break;
    hipMemcpy(c->k[a], k.data(), k.size() * 4, hipMemcpyHostToDevice);
    hipMemcpy(c->km[a], km.data(), km.size() * 4, hipMemcpyHostToDevice);
    int n = c->n[a];
    std::vector<float2> tw(n);
    for (int j = 0; j < n; ++j) {
      double ang = -2.0 * M_PI * (double)j / (double)n;
      tw[j] = make_float2((float)cos(ang), (float)sin(ang));
    }
    rc |= dev_alloc(c, (void**)&c->tw[a], n * sizeof(float2));
    if (rc) break;
    hipMemcpy(c->tw[a], tw.data(), n * sizeof(float2), hipMemcpyHostToDevice);
  }
  for (int b = 0; b < 3 && rc == PMX_OK; ++b)
    rc |= dev_alloc(c, (void**)&c->cbuf[b], (size_t)c->Mc * sizeof(float2));

  c->tg = make_tiles(c->n);
  const char* fol = getenv("PMX_FOLD_ON_LOAD");
  c->fold_on_load = !(fol && strcmp(fol, "0") == 0);
  const char* dsp = getenv("PMX_DEPOSIT_SPLIT");
  c->deposit_split = !(dsp && strcmp(dsp, "0") == 0);
  c->ntiles = (long long)c->tg.ntx * c->tg.nty * c->tg.ntz;
  if (rc == PMX_OK) rc |= dev_alloc(c, (void**)&c->tile_count, (size_t)(c->ntiles + 1) * 4);
  if (rc == PMX_OK) rc |= dev_alloc(c, (void**)&c->tile_start, (size_t)(c->ntiles + 1) * 4);
  if (rc == PMX_OK) rc |= pmx_ensure_halo(c, 1);
  if (rc == PMX_OK) rc |= dev_alloc(c, (void**)&c->tile_cap, (size_t)(c->ntiles + 1) * 4);
  if (rc == PMX_OK) rc |= dev_alloc(c, (void**)&c->flags, 4 * sizeof(uint32_t));
  if (rc == PMX_OK) rc |= dev_alloc(c, (void**)&c->vmax, 3 * sizeof(float));
  if (rc == PMX_OK) rc |= pmx_ensure_red(c, 4096);
  if (rc == PMX_OK) rc |= dev_alloc(c, (void**)&c->red_out, PMX_RED_SLOTS * sizeof(double));
  if (rc == PMX_OK && hipHostMalloc((void**)&c->red_host, PMX_RED_SLOTS * sizeof(double),
                                    hipHostMallocDefault) != hipSuccess) {
    pmx_set_error("pmx_ctx_create: hipHostMalloc failed");
    rc = PMX_ENOMEM;
  }
  if (rc != PMX_OK) {
    pmx_ctx_destroy(c);
    return rc;
  }
  // the table uploads are synchronous copies; surface any error of the set
  // up without waiting for other streams' work
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) {
    pmx_set_error("pmx_ctx_create: %s", hipGetErrorString(e));
    pmx_ctx_destroy(c);
    return PMX_EHIP;
  }
  *out = c;
  return PMX_OK;
}

// A stream of the caller's own, created here at the default priority and
// non-blocking (no implicit waits on the null stream), so that a capture on
// it cannot meet another thread's work: a host's stream pool may hand the
// same stream to several users.
extern "C" int pmx_stream_create(int device, void** out) {
  PMX_REQUIRE(out != nullptr, "pmx_stream_create: out is NULL");
  *out = nullptr;
  int cur = 0;
  PMX_CHECK_HIP(hipGetDevice(&cur));
  PMX_CHECK_HIP(hipSetDevice(device));
  hipStream_t st = nullptr;
  const hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
  (void)hipSetDevice(cur);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    pmx_set_error("pmx_stream_create: %s", hipGetErrorString(e));
    return PMX_EHIP;
  }
  *out = (void*)st;
  return PMX_OK;
}

extern "C" int pmx_stream_destroy(void* stream) {
  PMX_REQUIRE(stream != nullptr, "pmx_stream_destroy: stream is NULL");
  PMX_CHECK_HIP(hipStreamSynchronize((hipStream_t)stream));
  PMX_CHECK_HIP(hipStreamDestroy((hipStream_t)stream));
  return PMX_OK;
}

void pmx_timer_destroy(pmx_ctx* c);

extern "C" int pmx_ctx_destroy(pmx_ctx* c) {
  if (!c) return PMX_OK;
  hipDeviceSynchronize();
  pmx_timer_destroy(c);
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  // sizes are not needed here: the byte count ends with the context
  auto rel = [c](void* p) {
    if (!p) return;
    if (c->dfree)
      c->dfree(p, 0, c->device, (void*)c->stream, c->dalloc_user);
    else
      hipFree(p);
  };
  for (int a = 0; a < 3; ++a) {
    rel(c->k[a]);
    rel(c->km[a]);
    rel(c->tw[a]);
    rel(c->cbuf[a]);
  }
  void* ps[] = {c->tile_count, c->tile_start, c->tile_cap, c->flags, c->halo, c->rec, c->pkey,
                c->prank, c->pidx, c->vrec, c->vmax, c->red,
                c->red_out, c->sort_k[0], c->sort_k[1], c->sort_v[0], c->sort_v[1], c->sort_hist,
                c->scan_tmp};
  for (void* p : ps) rel(p);
  // pl[1], al[1], pl_tile and pl_rank alias the scratch freed above
  void* qs[] = {c->pl[0].s0, c->pl[0].s1, c->pl_count[0], c->pl_start[0], c->pl_count[1],
                c->pl_start[1], c->al[0].p.s0, c->al[0].p.s1, c->al[0].c0, c->al[0].c1,
                c->abuf[0], c->abuf[1], c->abuf[2], c->al_dots, c->lbuf, c->tab};
  for (void* p : qs) rel(p);
  if (c->red_host) hipHostFree(c->red_host);
  if (c->moved) hipEventDestroy(c->moved);
  free(c);
  return PMX_OK;
}

extern "C" int pmx_ctx_set_stream(pmx_ctx* c, void* stream) {
  PMX_REQUIRE(c != nullptr, "pmx_ctx_set_stream: ctx is NULL");
  hipStream_t s = (hipStream_t)stream;
  if (s == c->stream) return PMX_OK;
  // the context's scratch may still be in use by work queued on the old
  // stream: the new one waits for the event recorded after the last call
  // (the old stream's handle is not used again: its owner may have
  // destroyed it).  A stream under capture cannot wait on work outside its
  // capture; a capturing caller orders the streams itself.
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &cs) != hipSuccess) {
    (void)hipGetLastError();
    pmx_set_error("pmx_ctx_set_stream: stream capture query failed");
    return PMX_EHIP;
  }
  if (cs == hipStreamCaptureStatusNone && c->moved_valid &&
      hipStreamWaitEvent(s, c->moved, 0) != hipSuccess) {
    (void)hipGetLastError();
    pmx_set_error("pmx_ctx_set_stream: stream ordering failed");
    return PMX_EHIP;
  }
  c->stream = s;
  return PMX_OK;
}

extern "C" long long pmx_ctx_bytes(const pmx_ctx* c) { return c ? (long long)c->bytes : 0; }

extern "C" int pmx_ctx_trim(pmx_ctx* c, int what) {
  PMX_REQUIRE(c != nullptr, "pmx_ctx_trim: NULL context");
  PMX_REQUIRE((what & ~(PMX_TRIM_PIPELINE | PMX_TRIM_ALL)) == 0, "pmx_ctx_trim: bad flags");
  if (int rc = pmx_no_capture(c, "pmx_ctx_trim")) return rc;
  // work still queued on the context's stream may use the scratch
  hipError_t e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) {
    pmx_set_error("pmx_ctx_trim: %s", hipGetErrorString(e));
    return PMX_EHIP;
  }
  if (!(what & PMX_TRIM_ALL)) {
    if (what & PMX_TRIM_PIPELINE) {
      dev_free(c, (void**)&c->pl[0].s0, c->pl_cap * 16);
      dev_free(c, (void**)&c->pl[0].s1, c->pl_cap * 16);
      c->pl_cap = 0;
      free_adj_state(c);
    }
    return PMX_OK;
  }
  free_adj_state(c);
  dev_free(c, (void**)&c->lbuf, (size_t)c->Mc * sizeof(float2));
  dev_free(c, (void**)&c->rec, c->rec_cap * 16);
  dev_free(c, (void**)&c->pidx, c->rec_cap * 4);
  dev_free(c, (void**)&c->pkey, c->rec_cap * 4);
  dev_free(c, (void**)&c->prank, c->rec_cap * 4);
  dev_free(c, (void**)&c->vrec, c->vrec_cap * 4);
  dev_free(c, (void**)&c->pl[0].s0, c->pl_cap * 16);
  dev_free(c, (void**)&c->pl[0].s1, c->pl_cap * 16);
  c->pl[1] = PState{};
  c->pl_tile = c->pl_rank = nullptr;
  c->ptcl_cap = c->rec_cap = c->vrec_cap = c->pl_cap = 0;
  c->caps_valid = 0;
  if (c->halo_nch > 1) {
    dev_free(c, (void**)&c->halo, (size_t)c->halo_nch * c->ntiles * c->tg.shell * 4);
    c->halo_nch = 0;
    return pmx_ensure_halo(c, 1);
  }
  return PMX_OK;
}

extern "C" int pmx_tile_shape(const pmx_ctx* c, int t[3]) {
  PMX_REQUIRE(c != nullptr && t != nullptr, "pmx_tile_shape: NULL argument");
  t[0] = c->tg.tx; t[1] = c->tg.ty; t[2] = c->tg.tz;
  return PMX_OK;
}

int pmx_ensure_ptcl(pmx_ctx* c, long long n) {
  if (n <= c->ptcl_cap) return PMX_OK;
  if (int rc = pmx_no_capture(c, "particle scratch")) return rc;
  const long long cap = n + n / 16 + 1024;
  // records also hold the per-tile slack of the capacity binning
  const long long rcap = pmx_record_cap(c, cap);
  dev_free(c, (void**)&c->rec, c->rec_cap * 16);
  dev_free(c, (void**)&c->pidx, c->rec_cap * 4);
  dev_free(c, (void**)&c->pkey, c->rec_cap * 4);
  dev_free(c, (void**)&c->prank, c->rec_cap * 4);
  dev_free(c, (void**)&c->vrec, c->vrec_cap * 4);
  dev_free(c, (void**)&c->pl[0].s0, c->pl_cap * 16);
  dev_free(c, (void**)&c->pl[0].s1, c->pl_cap * 16);
  free_adj_state(c);
  c->ptcl_cap = c->rec_cap = c->vrec_cap = c->pl_cap = 0;
  c->caps_valid = 0;
  // pkey/prank hold a tile and rank per particle, or (multi-step re-bin) per record slot
  int rc = dev_alloc(c, (void**)&c->rec, rcap * 16);
  rc |= dev_alloc(c, (void**)&c->pidx, rcap * 4);
  rc |= dev_alloc(c, (void**)&c->pkey, rcap * 4);
  rc |= dev_alloc(c, (void**)&c->prank, rcap * 4);
  if (rc) {
    // give back the part that was allocated (the failed pointers are NULL)
    dev_free(c, (void**)&c->rec, rcap * 16);
    dev_free(c, (void**)&c->pidx, rcap * 4);
    dev_free(c, (void**)&c->pkey, rcap * 4);
    dev_free(c, (void**)&c->prank, rcap * 4);
    return PMX_ENOMEM;
  }
  c->ptcl_cap = cap;
  c->rec_cap = rcap;
  return PMX_OK;
}

int pmx_ensure_halo(pmx_ctx* c, int nch) {
  if (nch <= c->halo_nch) return PMX_OK;
  if (int rc = pmx_no_capture(c, "halo shells")) return rc;
  dev_free(c, (void**)&c->halo, (size_t)c->halo_nch * c->ntiles * c->tg.shell * 4);
  c->halo_nch = 0;
  if (dev_alloc(c, (void**)&c->halo, (size_t)nch * c->ntiles * c->tg.shell * 4)) return PMX_ENOMEM;
  c->halo_nch = nch;
  return PMX_OK;
}

// The multi-step pipeline's two tile-ordered particle copies of rec_cap
// slots: the first is its own memory, the second lives in the per-step
// path's scratch (s0 in rec, s1 in the four vrec channels), and
// the exact re-bin's per-slot tile and rank in pkey/prank.  The per-step
// path rewrites all of these before it reads them.
int pmx_ensure_state(pmx_ctx* c) {
  int rc;
  if ((rc = pmx_ensure_vrec(c))) return rc;
  const long long cap = c->rec_cap;
  if (cap > c->pl_cap) {
    if ((rc = pmx_no_capture(c, "pipeline particle copy"))) return rc;
    dev_free(c, (void**)&c->pl[0].s0, c->pl_cap * 16);
    dev_free(c, (void**)&c->pl[0].s1, c->pl_cap * 16);
    c->pl_cap = 0;
    rc = dev_alloc(c, (void**)&c->pl[0].s0, cap * 16);
    rc |= dev_alloc(c, (void**)&c->pl[0].s1, cap * 16);
    for (int b = 0; b < 2; ++b) {
      if (!c->pl_count[b]) rc |= dev_alloc(c, (void**)&c->pl_count[b], (size_t)(c->ntiles + 1) * 4);
      if (!c->pl_start[b]) rc |= dev_alloc(c, (void**)&c->pl_start[b], (size_t)(c->ntiles + 1) * 4);
    }
    if (rc) {
      dev_free(c, (void**)&c->pl[0].s0, cap * 16);
      dev_free(c, (void**)&c->pl[0].s1, cap * 16);
      return PMX_ENOMEM;
    }
    c->pl_cap = cap;
  }
  c->pl[1].s0 = c->rec;
  c->pl[1].s1 = reinterpret_cast<float4*>(c->vrec);
  c->pl_tile = c->pkey;
  c->pl_rank = c->prank;
  return PMX_OK;
}

// The reverse-time pipeline's own particle copy and per-slot gather
// results (rec_cap slots each), its per-step dot products, and its second
// copy in the per-step scratch and the forward pipeline's copy: p.s0 in rec,
// p.s1 in the four vrec channels, c0/c1 in the forward copy's s0/s1.
int pmx_ensure_adj_state(pmx_ctx* c, int nsteps) {
  int rc;
  if ((rc = pmx_ensure_state(c))) return rc;
  const long long cap = c->rec_cap;
  if (cap > c->al_cap) {
    if ((rc = pmx_no_capture(c, "reverse pipeline particle copy"))) return rc;
    AState& a = c->al[0];
    const long long old = c->al_cap;
    dev_free(c, (void**)&a.p.s0, old * 16);
    dev_free(c, (void**)&a.p.s1, old * 16);
    dev_free(c, (void**)&a.c0, old * 16);
    dev_free(c, (void**)&a.c1, old * 8);
    c->al_cap = 0;
    rc = dev_alloc(c, (void**)&a.p.s0, cap * 16);
    rc |= dev_alloc(c, (void**)&a.p.s1, cap * 16);
    rc |= dev_alloc(c, (void**)&a.c0, cap * 16);
    rc |= dev_alloc(c, (void**)&a.c1, cap * 8);
    if (rc) {
      dev_free(c, (void**)&a.p.s0, cap * 16);
      dev_free(c, (void**)&a.p.s1, cap * 16);
      dev_free(c, (void**)&a.c0, cap * 16);
      dev_free(c, (void**)&a.c1, cap * 8);
      return PMX_ENOMEM;
    }
    c->al_cap = cap;
  }
  for (int b = 0; b < 3; ++b)
    if (!c->abuf[b] && dev_alloc(c, (void**)&c->abuf[b], (size_t)c->Mc * sizeof(float2)))
      return PMX_ENOMEM;
  if (2LL * nsteps > c->al_dots_cap) {
    if ((rc = pmx_no_capture(c, "reverse pipeline dot products"))) return rc;
    dev_free(c, (void**)&c->al_dots, c->al_dots_cap * 8);
    c->al_dots_cap = 0;
    const long long k = 2LL * nsteps + 64;
    if (dev_alloc(c, (void**)&c->al_dots, k * 8)) return PMX_ENOMEM;
    c->al_dots_cap = k;
  }
  c->al[1].p = c->pl[1];
  c->al[1].c0 = c->pl[0].s0;
  c->al[1].c1 = reinterpret_cast<float2*>(c->pl[0].s1);
  return PMX_OK;
}

static void free_adj_state(pmx_ctx* c) {
  AState& a = c->al[0];
  const long long cap = c->al_cap;
  dev_free(c, (void**)&a.p.s0, cap * 16);
  dev_free(c, (void**)&a.p.s1, cap * 16);
  dev_free(c, (void**)&a.c0, cap * 16);
  dev_free(c, (void**)&a.c1, cap * 8);
  for (int b = 0; b < 3; ++b) dev_free(c, (void**)&c->abuf[b], (size_t)c->Mc * sizeof(float2));
  c->al_cap = 0;
  c->al[1] = AState{};
}

// per-record value channels, rec_cap floats apart: three for the adjoint's
// deposits, a fourth so the four hold a float4 per slot for the pipeline
int pmx_ensure_vrec(pmx_ctx* c) {
  const long long cap = 4 * c->rec_cap;
  if (cap <= c->vrec_cap) return PMX_OK;
  if (int rc = pmx_no_capture(c, "value channels")) return rc;
  dev_free(c, (void**)&c->vrec, c->vrec_cap * 4);
  c->vrec_cap = 0;
  if (dev_alloc(c, (void**)&c->vrec, cap * 4)) return PMX_ENOMEM;
  c->vrec_cap = cap;
  return PMX_OK;
}

int pmx_ensure_red(pmx_ctx* c, long long nblk) {
  if (nblk <= c->red_cap) return PMX_OK;
  if (int rc = pmx_no_capture(c, "reduction partials")) return rc;
  dev_free(c, (void**)&c->red, (size_t)c->red_cap * 8);
  c->red_cap = 0;
  if (dev_alloc(c, (void**)&c->red, (size_t)nblk * 8)) return PMX_ENOMEM;
  c->red_cap = nblk;
  return PMX_OK;
}

MeshGeom pmx_geom_padded(const pmx_ctx* c) {
  MeshGeom g;
  g.n0 = c->n[0]; g.n1 = c->n[1]; g.n2 = c->n[2];
  g.rs = c->rs;
  g.cell = c->cellf;
  int e = 0;
  const float m = frexpf(c->cellf, &e);  // cell = m 2^e, m in [0.5, 1)
  const float inv = ldexpf(1.f, 1 - e);
  g.inv_cell = (m == 0.5f && isfinite(inv) && inv >= FLT_MIN) ? inv : 0.f;
  return g;
}

// ------------------------------------------------------------------ scan --
// Exclusive prefix sum of uint32, three phases: per-block sums, a
// single-block scan of the block sums, then per-block scan plus offset.
#define SCAN_ITEMS 8
#define SCAN_TILE (PMX_TPB * SCAN_ITEMS)

__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* sh, uint32_t* total) {
  int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int off = 1; off < PMX_TPB; off <<= 1) {
    uint32_t x = t >= off ? sh[t - off] : 0;
    __syncthreads();
    sh[t] += x;
    __syncthreads();
  }
  uint32_t incl = sh[t];
  *total = sh[PMX_TPB - 1];
  __syncthreads();
  return incl - v;
}

// The scan kernels return at once when cond is given and *cond == 0 (a
// device-side decision, no host round trip).
__global__ void scan_reduce_k(const uint32_t* __restrict__ in, long long n, uint32_t* __restrict__ bsum,
                              const uint32_t* __restrict__ cond) {
  if (cond && *cond == 0) return;
  __shared__ uint32_t sh[PMX_TPB];
  long long base = (long long)blockIdx.x * SCAN_TILE;
  uint32_t s = 0;
  for (int i = 0; i < SCAN_ITEMS; ++i) {
    long long j = base + (long long)i * PMX_TPB + threadIdx.x;
    if (j < n) s += in[j];
  }
  uint32_t tot;
  block_exclusive_scan(s, sh, &tot);
  if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

__global__ void scan_single_k(uint32_t* __restrict__ data, long long n,
                              const uint32_t* __restrict__ cond) {
  if (cond && *cond == 0) return;
  __shared__ uint32_t sh[PMX_TPB];
  uint32_t carry = 0;
  for (long long base = 0; base < n; base += PMX_TPB) {
    long long j = base + threadIdx.x;
    uint32_t v = j < n ? data[j] : 0;
    uint32_t tot;
    uint32_t ex = block_exclusive_scan(v, sh, &tot);
    if (j < n) data[j] = ex + carry;
    carry += tot;
  }
}

__global__ void scan_apply_k(const uint32_t* __restrict__ in, long long n,
                             const uint32_t* __restrict__ boff, uint32_t* __restrict__ out,
                             const uint32_t* __restrict__ cond) {
  if (cond && *cond == 0) return;
  __shared__ uint32_t sh[PMX_TPB];
  long long base = (long long)blockIdx.x * SCAN_TILE + (long long)threadIdx.x * SCAN_ITEMS;
  uint32_t v[SCAN_ITEMS];
  uint32_t s = 0;
  for (int i = 0; i < SCAN_ITEMS; ++i) {
    long long j = base + i;
    v[i] = j < n ? in[j] : 0;
    s += v[i];
  }
  uint32_t tot;
  uint32_t ex = block_exclusive_scan(s, sh, &tot) + boff[blockIdx.x];
  for (int i = 0; i < SCAN_ITEMS; ++i) {
    long long j = base + i;
    if (j < n) out[j] = ex;
    ex += v[i];
  }
}

long long pmx_scan_blocks(long long n) {
  long long nb = (n + SCAN_TILE - 1) / SCAN_TILE;
  return nb < 1 ? 1 : nb;
}

// Sizes the scan scratch for scans of up to n entries, so that entry points
// can take it before their first kernel writes the caller's arrays.
int pmx_ensure_scan(pmx_ctx* c, long long n) {
  long long nb = (n + SCAN_TILE - 1) / SCAN_TILE;
  if (nb < 1) nb = 1;
  if (nb > c->scan_tmp_cap) {
    if (int rc = pmx_no_capture(c, "scan scratch")) return rc;
    dev_free(c, (void**)&c->scan_tmp, c->scan_tmp_cap * 4);
    c->scan_tmp_cap = 0;
    if (dev_alloc(c, (void**)&c->scan_tmp, nb * 4)) return PMX_ENOMEM;
    c->scan_tmp_cap = nb;
  }
  return PMX_OK;
}

int pmx_scan_exclusive(pmx_ctx* c, const uint32_t* in, uint32_t* out, long long n,
                       const uint32_t* cond) {
  long long nb = (n + SCAN_TILE - 1) / SCAN_TILE;
  if (nb < 1) nb = 1;
  if (int rc = pmx_ensure_scan(c, n)) return rc;
  scan_reduce_k<<<(unsigned)nb, PMX_TPB, 0, c->stream>>>(in, n, c->scan_tmp, cond);
  scan_single_k<<<1, PMX_TPB, 0, c->stream>>>(c->scan_tmp, nb, cond);
  scan_apply_k<<<(unsigned)nb, PMX_TPB, 0, c->stream>>>(in, n, c->scan_tmp, out, cond);
  PMX_LAUNCH_CHECK();
  return PMX_OK;
}

// -------------------------------------------------------------- dot ---
__global__ void dot_k(const float* __restrict__ a, const float* __restrict__ b, long long n,
                      double* __restrict__ part) {
  __shared__ double sh[PMX_TPB];
  double s = 0;
  for (long long i = (long long)blockIdx.x * PMX_TPB + threadIdx.x; i < n;
       i += (long long)gridDim.x * PMX_TPB)
    s += (double)a[i] * (double)b[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int o = PMX_TPB / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// One workgroup sums part[0, nblk) in a fixed order (strided per thread,
// then a tree), so the result does not depend on timing.
#define RED_TPB 1024
__global__ void __launch_bounds__(RED_TPB)
    red_sum_k(const double* __restrict__ part, long long nblk, double* __restrict__ out) {
  __shared__ double sh[RED_TPB];
  double s = 0;
  for (long long i = threadIdx.x; i < nblk; i += RED_TPB) s += part[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int o = RED_TPB / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sh[0];
}

int pmx_red_reduce(pmx_ctx* c, long long off, long long nblk, int slot) {
  PMX_REQUIRE(slot >= 0 && slot < PMX_RED_SLOTS && off >= 0 && off + nblk <= c->red_cap,
              "pmx_red_reduce: bad range");
  red_sum_k<<<1, RED_TPB, 0, c->stream>>>(c->red + off, nblk, c->red_out + slot);
  PMX_LAUNCH_CHECK();
  return PMX_OK;
}

int pmx_red_reduce_to(pmx_ctx* c, long long off, long long nblk, double* dst) {
  PMX_REQUIRE(dst && off >= 0 && off + nblk <= c->red_cap, "pmx_red_reduce_to: bad range");
  red_sum_k<<<1, RED_TPB, 0, c->stream>>>(c->red + off, nblk, dst);
  PMX_LAUNCH_CHECK();
  return PMX_OK;
}

int pmx_red_fetch(pmx_ctx* c, int k, double* out) {
  PMX_REQUIRE(k > 0 && k <= PMX_RED_SLOTS, "pmx_red_fetch: bad count");
  if (int rc = pmx_no_capture(c, "host value of a dot product")) return rc;
  PMX_CHECK_HIP(hipMemcpyAsync(c->red_host, c->red_out, (size_t)k * 8, hipMemcpyDeviceToHost,
                               c->stream));
  PMX_CHECK_HIP(hipStreamSynchronize(c->stream));
  for (int i = 0; i < k; ++i) out[i] = c->red_host[i];
  return PMX_OK;
}

int pmx_red_finish(pmx_ctx* c, long long nblk, double* out) {
  int rc = pmx_red_reduce(c, 0, nblk, 0);
  if (rc) return rc;
  return pmx_red_fetch(c, 1, out);
}

extern "C" int pmx_dot(pmx_ctx* c, const float* a, const float* b, long long n, double* out) {
  PMX_ENTRY(c);
  PMX_REQUIRE(c && a && b && out && n >= 0, "pmx_dot: bad argument");
  int nblk = 512;
  dot_k<<<nblk, PMX_TPB, 0, c->stream>>>(a, b, n, c->red);
  PMX_LAUNCH_CHECK();
  return pmx_red_finish(c, nblk, out);
}

// ------------------------------------------------------------- timing ---
#include <vector>

struct PmxTimer {
  struct Rec {
    int cls;
    hipEvent_t a, b;
  };
  std::vector<Rec> recs;
  std::vector<hipEvent_t> pool;
  hipEvent_t open[PMX_NTIMER];
  double ms[PMX_NTIMER];
  long long cnt[PMX_NTIMER];
};

static hipEvent_t timer_event(PmxTimer* t) {
  hipEvent_t e = nullptr;
  if (!t->pool.empty()) {
    e = t->pool.back();
    t->pool.pop_back();
  } else if (hipEventCreate(&e) != hipSuccess) {
    e = nullptr;
  }
  return e;
}

void pmx_tic(pmx_ctx* c, int cls) {
  if (!c->timer) return;
  hipEvent_t e = timer_event(c->timer);
  if (e && hipEventRecord(e, c->stream) == hipSuccess) c->timer->open[cls] = e;
  else c->timer->open[cls] = nullptr;
}

void pmx_toc(pmx_ctx* c, int cls) {
  if (!c->timer || !c->timer->open[cls]) return;
  hipEvent_t e = timer_event(c->timer);
  if (e && hipEventRecord(e, c->stream) == hipSuccess)
    c->timer->recs.push_back({cls, c->timer->open[cls], e});
  c->timer->open[cls] = nullptr;
}

static void timer_free(PmxTimer* t) {
  for (auto& r : t->recs) {
    hipEventDestroy(r.a);
    hipEventDestroy(r.b);
  }
  for (auto e : t->pool) hipEventDestroy(e);
  delete t;
}

extern "C" int pmx_timing_enable(pmx_ctx* c, int on) {
  PMX_REQUIRE(c != nullptr, "pmx_timing_enable: ctx is NULL");
  if (on && !c->timer) {
    c->timer = new PmxTimer();
    for (int i = 0; i < PMX_NTIMER; ++i) {
      c->timer->open[i] = nullptr;
      c->timer->ms[i] = 0;
      c->timer->cnt[i] = 0;
    }
  } else if (!on && c->timer) {
    hipStreamSynchronize(c->stream);
    timer_free(c->timer);
    c->timer = nullptr;
  }
  return PMX_OK;
}

extern "C" int pmx_timing_read(pmx_ctx* c, double* ms, long long* count) {
  PMX_REQUIRE(c && ms && count, "pmx_timing_read: NULL argument");
  for (int i = 0; i < PMX_NTIMER; ++i) {
    ms[i] = 0;
    count[i] = 0;
  }
  if (!c->timer) return PMX_OK;
  if (int rc = pmx_no_capture(c, "pmx_timing_read")) return rc;
  PMX_CHECK_HIP(hipStreamSynchronize(c->stream));
  PmxTimer* t = c->timer;
  for (auto& r : t->recs) {
    float e = 0.f;
    if (hipEventElapsedTime(&e, r.a, r.b) == hipSuccess) {
      ms[r.cls] += e;
      count[r.cls] += 1;
    }
    t->pool.push_back(r.a);
    t->pool.push_back(r.b);
  }
  t->recs.clear();
  return PMX_OK;
}

void pmx_timer_destroy(pmx_ctx* c) {
  if (c->timer) {
    timer_free(c->timer);
    c->timer = nullptr;
  }
}
