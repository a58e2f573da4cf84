// Stable LSD radix sort of uint32 keys producing the argsort permutation
// (8-bit digits; histogram -> exclusive scan -> stable scatter per pass).
// Stability inside a block comes from ranking equal digits with wave
// ballots in index order, then ordering waves and rounds through LDS.
#include "pmx_internal.h"

#define RS_ROUNDS 16
#define RS_TILE (PMX_TPB * RS_ROUNDS)
#define RS_WAVES (PMX_TPB / 64)

__global__ void rs_hist_k(const uint32_t* __restrict__ kin, long long n, int shift, int nblocks,
                          uint32_t* __restrict__ hist) {
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  long long base = (long long)blockIdx.x * RS_TILE;
  for (int r = 0; r < RS_ROUNDS; ++r) {
    long long i = base + r * PMX_TPB + threadIdx.x;
    if (i < n) atomicAdd(&h[(kin[i] >> shift) & 255u], 1u);
  }
  __syncthreads();
  hist[(long long)threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

__global__ void rs_scatter_k(const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin,
                             long long n, int shift, const uint32_t* __restrict__ hscan, int nblocks,
                             uint32_t* __restrict__ kout, uint32_t* __restrict__ vout) {
  __shared__ uint32_t base[256];
  __shared__ uint32_t wcnt[RS_WAVES][256];
  const int tid = threadIdx.x;
  const int lane = __lane_id();
  const int w = tid >> 6;
  base[tid] = hscan[(long long)tid * nblocks + blockIdx.x];
  const unsigned long long lt = (1ull << lane) - 1ull;
  for (int r = 0; r < RS_ROUNDS; ++r) {
    for (int q = 0; q < RS_WAVES; ++q) wcnt[q][tid] = 0;
    __syncthreads();
    long long i = (long long)blockIdx.x * RS_TILE + (long long)r * PMX_TPB + tid;
    bool act = i < n;
    uint32_t key = act ? kin[i] : 0u;
    uint32_t val = act ? (vin ? vin[i] : (uint32_t)i) : 0u;
    uint32_t dg = (key >> shift) & 255u;
    unsigned long long m = __ballot(act);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      bool bit = (dg >> b) & 1u;
      unsigned long long bb = __ballot(bit);
      m &= bit ? bb : ~bb;
    }
    uint32_t rk = (uint32_t)__popcll(m & lt);
    if (act && rk == 0) wcnt[w][dg] = (uint32_t)__popcll(m);
    __syncthreads();
    if (act) {
      uint32_t pos = base[dg] + rk;
      for (int q = 0; q < w; ++q) pos += wcnt[q][dg];
      kout[pos] = key;
      vout[pos] = val;
    }
    __syncthreads();
    uint32_t add = 0;
    for (int q = 0; q < RS_WAVES; ++q) add += wcnt[q][tid];
    base[tid] += add;
    __syncthreads();
  }
}

__global__ void max_u32_k(const uint32_t* __restrict__ k, long long n, uint32_t* __restrict__ out) {
  __shared__ uint32_t sh[PMX_TPB];
  uint32_t m = 0;
  for (long long i = (long long)blockIdx.x * PMX_TPB + threadIdx.x; i < n;
       i += (long long)gridDim.x * PMX_TPB)
    m = k[i] > m ? k[i] : m;
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int o = PMX_TPB / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) sh[threadIdx.x] = sh[threadIdx.x] > sh[threadIdx.x + o] ? sh[threadIdx.x] : sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicMax(out, sh[0]);
}

static int ensure_sort(pmx_ctx* c, long long n, long long nhist) {
  if (n > c->sort_cap) {
    if (int rc = pmx_no_capture(c, "sort scratch")) return rc;
    for (int b = 0; b < 2; ++b) {
      pmx_dev_free(c, (void**)&c->sort_k[b], (size_t)c->sort_cap * 4);
      pmx_dev_free(c, (void**)&c->sort_v[b], (size_t)c->sort_cap * 4);
    }
    c->sort_cap = 0;
    int rc = PMX_OK;
    for (int b = 0; b < 2; ++b) {
      rc |= pmx_dev_alloc(c, (void**)&c->sort_k[b], (size_t)n * 4);
      rc |= pmx_dev_alloc(c, (void**)&c->sort_v[b], (size_t)n * 4);
    }
    if (rc) {
      for (int b = 0; b < 2; ++b) {
        pmx_dev_free(c, (void**)&c->sort_k[b], (size_t)n * 4);
        pmx_dev_free(c, (void**)&c->sort_v[b], (size_t)n * 4);
      }
      return PMX_ENOMEM;
    }
    c->sort_cap = n;
  }
  if (nhist > c->sort_hist_cap) {
    if (int rc = pmx_no_capture(c, "sort histograms")) return rc;
    pmx_dev_free(c, (void**)&c->sort_hist, (size_t)c->sort_hist_cap * 8);
    c->sort_hist_cap = 0;
    if (pmx_dev_alloc(c, (void**)&c->sort_hist, (size_t)nhist * 8)) return PMX_ENOMEM;
    c->sort_hist_cap = nhist;
  }
  return PMX_OK;
}

extern "C" int pmx_sort_perm(pmx_ctx* c, const uint32_t* key, long long n, uint32_t* perm) {
  PMX_ENTRY(c);
  PMX_REQUIRE(c && n >= 0, "pmx_sort_perm: bad argument");
  if (n == 0) return PMX_OK;
  PMX_REQUIRE(key && perm, "pmx_sort_perm: NULL pointer");
  PMX_REQUIRE(n < (1LL << 32), "pmx_sort_perm: n must be < 2^32");
  long long nblocks = (n + RS_TILE - 1) / RS_TILE;
  long long nhist = 256 * nblocks;
  int rc;
  // the digit count is read back on the host
  if ((rc = pmx_no_capture(c, "pmx_sort_perm"))) return rc;
  if ((rc = ensure_sort(c, n, nhist))) return rc;
  uint32_t* hist = c->sort_hist;
  uint32_t* hscan = c->sort_hist + nhist;
  // number of digit passes from the largest key
  uint32_t* dmax = hscan;  // scratch word, overwritten by the first scan
  if ((rc = pmx_fill_u32(c, dmax, 1, 0u))) return rc;
  max_u32_k<<<256, PMX_TPB, 0, c->stream>>>(key, n, dmax);
  PMX_LAUNCH_CHECK();
  uint32_t kmax = 0;
  PMX_CHECK_HIP(hipMemcpyAsync(&kmax, dmax, 4, hipMemcpyDeviceToHost, c->stream));
  PMX_CHECK_HIP(hipStreamSynchronize(c->stream));
  int passes = 0;
  while (passes < 4 && (kmax >> (8 * passes)) != 0) ++passes;
  if (passes == 0) passes = 1;
  const uint32_t* kin = key;
  const uint32_t* vin = nullptr;
  int cur = 0;
  for (int p = 0; p < passes; ++p) {
    int shift = 8 * p;
    rs_hist_k<<<(unsigned)nblocks, PMX_TPB, 0, c->stream>>>(kin, n, shift, (int)nblocks, hist);
    PMX_LAUNCH_CHECK();
    if ((rc = pmx_scan_exclusive(c, hist, hscan, nhist))) return rc;
    uint32_t* vout = (p == passes - 1) ? perm : c->sort_v[cur];
    rs_scatter_k<<<(unsigned)nblocks, PMX_TPB, 0, c->stream>>>(kin, vin, n, shift, hscan,
                                                               (int)nblocks, c->sort_k[cur], vout);
    PMX_LAUNCH_CHECK();
    kin = c->sort_k[cur];
    vin = vout;
    cur ^= 1;
  }
  return PMX_OK;
}
