// White noise on the device: standard normals from a counter-based
// generator, so that any element can be drawn on its own and the whole
// field is one embarrassingly parallel pass (written once, HBM-bound).
//
// Generator: Philox4x32-10 (Salmon, Moraes, Dror and Shaw, "Parallel random
// numbers: as easy as 1, 2, 3", SC'11): key = the 64-bit seed as two 32-bit
// words (low first), counter = (block index low, high, 0, 0); each block
// gives four 32-bit words.  Normals: the Box-Muller transform of word pairs
// (a, b) in float64, u = (w + 1/2) 2^-32, z0 = sqrt(-2 ln u_a) cos(2 pi u_b),
// z1 = sqrt(-2 ln u_a) sin(2 pi u_b), rounded to float32.  Element i is
// normal i % 4 of block i / 4.
//
// pmwd draws its modes with jax.random.normal (threefry, modes.py:33-36);
// that bit stream depends on the jax version and is not reproduced here.
#include "pmx_internal.h"

#include <math.h>

__device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint32_t hi0 = __umulhi(0xD2511F53u, c[0]), lo0 = 0xD2511F53u * c[0];
    const uint32_t hi1 = __umulhi(0xCD9E8D57u, c[2]), lo1 = 0xCD9E8D57u * c[2];
    const uint32_t n0 = hi1 ^ c[1] ^ k0;
    const uint32_t n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0;
    c[1] = lo1;
    c[2] = n2;
    c[3] = lo0;
  }
}

__device__ __forceinline__ void box_muller(uint32_t a, uint32_t b, float& z0, float& z1) {
  const double u1 = ((double)a + 0.5) * 0x1p-32;
  const double u2 = ((double)b + 0.5) * 0x1p-32;
  const double r = sqrt(-2.0 * log(u1));
  const double t = 6.283185307179586 * u2;
  z0 = (float)(r * cos(t));
  z1 = (float)(r * sin(t));
}

__global__ void __launch_bounds__(PMX_TPB) white_noise_k(uint32_t k0, uint32_t k1, long long n,
                                                         float* __restrict__ out) {
  const long long nb = (n + 3) / 4;
  for (long long b = (long long)blockIdx.x * PMX_TPB + threadIdx.x; b < nb;
       b += (long long)gridDim.x * PMX_TPB) {
    uint32_t c[4] = {(uint32_t)b, (uint32_t)((unsigned long long)b >> 32), 0u, 0u};
    philox4x32_10(c, k0, k1);
    float4 z;
    box_muller(c[0], c[1], z.x, z.y);
    box_muller(c[2], c[3], z.z, z.w);
    const long long i = 4 * b;
    if (i + 3 < n) {
      *reinterpret_cast<float4*>(out + i) = z;
    } else {
      const float zs[4] = {z.x, z.y, z.z, z.w};
      for (int j = 0; i + j < n; ++j) out[i + j] = zs[j];
    }
  }
}

__global__ void __launch_bounds__(PMX_TPB) random_bits_k(uint32_t k0, uint32_t k1, long long nb,
                                                         uint4* __restrict__ out) {
  for (long long b = (long long)blockIdx.x * PMX_TPB + threadIdx.x; b < nb;
       b += (long long)gridDim.x * PMX_TPB) {
    uint32_t c[4] = {(uint32_t)b, (uint32_t)((unsigned long long)b >> 32), 0u, 0u};
    philox4x32_10(c, k0, k1);
    out[b] = make_uint4(c[0], c[1], c[2], c[3]);
  }
}

extern "C" int pmx_white_noise(pmx_ctx* c, unsigned long long seed, long long n, float* out) {
  PMX_ENTRY(c);
  PMX_REQUIRE(c && n >= 0, "pmx_white_noise: bad argument");
  if (n == 0) return PMX_OK;
  PMX_REQUIRE(out != nullptr, "pmx_white_noise: out is NULL");
  PMX_REQUIRE(((uintptr_t)out & 15) == 0, "pmx_white_noise: out must be 16-byte aligned");
  const long long nb = (n + 3) / 4;
  white_noise_k<<<pmx_blocks_capped(nb, 1 << 16), PMX_TPB, 0, c->stream>>>(
      (uint32_t)seed, (uint32_t)(seed >> 32), n, out);
  PMX_LAUNCH_CHECK();
  return PMX_OK;
}

extern "C" int pmx_random_bits(pmx_ctx* c, unsigned long long seed, long long nblocks,
                               uint32_t* out) {
  PMX_ENTRY(c);
  PMX_REQUIRE(c && nblocks >= 0, "pmx_random_bits: bad argument");
  if (nblocks == 0) return PMX_OK;
  PMX_REQUIRE(out != nullptr && ((uintptr_t)out & 15) == 0,
              "pmx_random_bits: out must be a 16-byte aligned device pointer");
  random_bits_k<<<pmx_blocks_capped(nblocks, 1 << 16), PMX_TPB, 0, c->stream>>>(
      (uint32_t)seed, (uint32_t)(seed >> 32), nblocks, reinterpret_cast<uint4*>(out));
  PMX_LAUNCH_CHECK();
  return PMX_OK;
}
