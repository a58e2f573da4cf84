// A default-argument CIC mass assignment that allocates nothing: the same
// binning and LDS-tile deposit kernels as pmx_scatter_mass, with their
// scratch (tile counts, starts and capacities, the overflow flag, the scan
// partials, the +1 halo shells, and per record slot the record, index, tile
// and rank words) carved from one file-scope device array instead of a
// context's allocations.  For callers that cannot allocate device memory
// (code under graph capture, tight serving loops, hosts that forbid it).
//
// Each device holds its own copy of the array until the process ends; its
// address is taken once per device with hipGetSymbolAddress and kept here.
// Every call on a device uses that device's array, so calls that overlap on
// one device (two host threads, or two streams) race and give wrong results:
// the caller serialises them.  The kernels clear every counter they read
// before use, so a call does not depend on what the last one left.
#include "pmx_internal.h"

#include <mutex>

// 64 MiB: about 1.9 million particles on a 128^3 mesh, 1.5 million on 256^3
#define PMX_NOALLOC_BYTES (64ull << 20)
#define PMX_NOALLOC_MAX_DEVICES 64

static __device__ __align__(256) unsigned char scratch[PMX_NOALLOC_BYTES];

static std::mutex g_scratch_lock;
static void* g_scratch_ptr[PMX_NOALLOC_MAX_DEVICES];

// The device address of this device's copy of the array (the array's name
// is only a host-side stand-in).
static int scratch_ptr(int device, unsigned char** out) {
  PMX_REQUIRE(device >= 0 && device < PMX_NOALLOC_MAX_DEVICES,
              "pmx_scatter_mass_noalloc: device %d out of range", device);
  std::lock_guard<std::mutex> lock(g_scratch_lock);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      pmx_set_error("pmx_scatter_mass_noalloc: hipGetSymbolAddress failed: %s",
                    hipGetErrorString(e));
      return PMX_EHIP;
    }
    g_scratch_ptr[device] = p;
  }
  *out = (unsigned char*)g_scratch_ptr[device];
  return PMX_OK;
}

// Bump allocation of 256-byte aligned pieces of the array.
struct Carve {
  unsigned char* base;
  size_t used;
  template <class T>
  T* take(size_t count) {
    T* p = reinterpret_cast<T*>(base + used);
    used += (count * sizeof(T) + 255) / 256 * 256;
    return p;
  }
};

extern "C" int pmx_scatter_mass_noalloc(const int16_t* pmid, const float* disp, long long n,
                                        int n0, int n1, int n2, double cell_size, float val,
                                        float* mesh, int accumulate, void* stream) {
  PMX_REQUIRE(mesh && n >= 0, "pmx_scatter_mass_noalloc: bad argument");
  PMX_REQUIRE(n == 0 || (pmid && disp), "pmx_scatter_mass_noalloc: NULL particle arrays");
  PMX_REQUIRE(n0 > 0 && n1 > 0 && n2 > 0 && n0 <= 4096 && n1 <= 4096 && n2 <= 4096,
              "pmx_scatter_mass_noalloc: mesh extents must be in 1..4096");
  PMX_REQUIRE(cell_size > 0, "pmx_scatter_mass_noalloc: cell_size must be positive");
  PMX_REQUIRE(n < (1LL << 31), "pmx_scatter_mass_noalloc: too many particles");
  // a context on the host stack: geometry only, every scratch pointer into
  // the static array, and no_alloc so that nothing can grow it
  pmx_ctx c = {};
  pmx_ctx_geometry(&c, n0, n1, n2, cell_size, (hipStream_t)stream);
  c.no_alloc = 1;
  const long long nt = c.ntiles;
  const long long cap = n + n / 16 + 1024;
  const long long rcap = pmx_record_cap(&c, cap);
  const long long nscan = pmx_scan_blocks(nt + 1);
  Carve sz = {nullptr, 0};
  sz.take<uint32_t>(nt + 1);                       // tile_count
  sz.take<uint32_t>(nt + 1);                       // tile_start
  sz.take<uint32_t>(nt + 1);                       // tile_cap
  sz.take<uint32_t>(4);                            // flags
  sz.take<uint32_t>(nscan);                        // scan partials
  sz.take<float>((size_t)nt * c.tg.shell);         // one channel of halo shells
  sz.take<float4>(rcap);                           // records
  sz.take<uint32_t>(rcap);                         // pidx
  sz.take<uint32_t>(rcap);                         // pkey
  sz.take<uint32_t>(rcap);                         // prank
  PMX_REQUIRE(sz.used <= PMX_NOALLOC_BYTES,
              "pmx_scatter_mass_noalloc: %lld particles on a %dx%dx%d mesh need %zu bytes of "
              "scratch, more than the %llu it holds (use pmx_scatter_mass)", n, n0, n1, n2,
              sz.used, (unsigned long long)PMX_NOALLOC_BYTES);
  unsigned char* base = nullptr;
  if (int rc = scratch_ptr(c.device, &base)) return rc;
  Carve cv = {base, 0};
  c.tile_count = cv.take<uint32_t>(nt + 1);
  c.tile_start = cv.take<uint32_t>(nt + 1);
  c.tile_cap = cv.take<uint32_t>(nt + 1);
  c.flags = cv.take<uint32_t>(4);
  c.scan_tmp = cv.take<uint32_t>(nscan);
  c.scan_tmp_cap = nscan;
  c.halo = cv.take<float>((size_t)nt * c.tg.shell);
  c.halo_nch = 1;
  c.rec = cv.take<float4>(rcap);
  c.pidx = cv.take<uint32_t>(rcap);
  c.pkey = cv.take<uint32_t>(rcap);
  c.prank = cv.take<uint32_t>(rcap);
  c.ptcl_cap = cap;
  c.rec_cap = rcap;
  c.caps_valid = 0;   // exact count, scan and fill: nothing read from the last call
  int rc;
  if ((rc = pmx_bin_particles(&c, pmid, disp, n))) return rc;
  return pmx_deposit_binned(&c, val, n, mesh, n2, accumulate);
}
