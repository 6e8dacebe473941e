// avbd_scan.hpp — exclusive prefix sum of uint32 arrays on the device.
//
// out[0..n] receives the exclusive scan of in[0..n) with out[n] = total.
// Blocks of 256 threads scan 1024 elements each in LDS; block totals are
// scanned recursively and added back.  Integer addition is associative, so
// the result does not depend on the blocking.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#define SCAN_BLOCK 256
#define SCAN_PER_THREAD 4
#define SCAN_TILE (SCAN_BLOCK * SCAN_PER_THREAD)

__global__ void __launch_bounds__(SCAN_BLOCK) k_scan_tile(const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
                                                          uint32_t* __restrict__ tile_sum, uint32_t n) {
  __shared__ uint32_t s[SCAN_BLOCK];
  const uint32_t tid = threadIdx.x;
  const uint32_t base = blockIdx.x * SCAN_TILE + tid * SCAN_PER_THREAD;
  uint32_t v[SCAN_PER_THREAD];
  uint32_t t = 0;
#pragma unroll
  for (int k = 0; k < SCAN_PER_THREAD; ++k) {
    v[k] = (base + k < n) ? in[base + k] : 0u;
    t += v[k];
  }
  s[tid] = t;
  __syncthreads();
  for (uint32_t off = 1; off < SCAN_BLOCK; off <<= 1) {
    uint32_t add = tid >= off ? s[tid - off] : 0u;
    __syncthreads();
    s[tid] += add;
    __syncthreads();
  }
  uint32_t run = s[tid] - t;
#pragma unroll
  for (int k = 0; k < SCAN_PER_THREAD; ++k) {
    if (base + k < n) out[base + k] = run;
    run += v[k];
  }
  if (tid == SCAN_BLOCK - 1) tile_sum[blockIdx.x] = s[tid];
}

__global__ void __launch_bounds__(SCAN_BLOCK) k_scan_add(uint32_t* __restrict__ out, const uint32_t* __restrict__ tile_off,
                                                         uint32_t n) {
  const uint32_t base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_PER_THREAD;
  const uint32_t add = tile_off[blockIdx.x];
#pragma unroll
  for (int k = 0; k < SCAN_PER_THREAD; ++k)
    if (base + k < n) out[base + k] += add;
}

__global__ void k_scan_total(uint32_t* __restrict__ out, uint32_t n, const uint32_t* __restrict__ src) { out[n] = src[0]; }
__global__ void k_scan_zero(uint32_t* __restrict__ out) { out[0] = 0u; }

// Scratch for the recursive levels: sized once for the largest n.
struct ScanScratch {
  uint32_t* buf = nullptr;
  size_t words = 0;
  static size_t need(size_t n) {
    size_t w = 0;
    while (n > 1) {
      size_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
      w += 2 * nb + 2;
      n = nb;
    }
    return w + 4;
  }
};

// out must hold n + 1 entries.
static inline void scan_excl(const uint32_t* in, uint32_t* out, uint32_t n, hipStream_t st, uint32_t* scratch) {
  if (n == 0) {
    k_scan_zero<<<1, 1, 0, st>>>(out);
    return;
  }
  uint32_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
  uint32_t* tsum = scratch;
  uint32_t* toff = scratch + nb;
  k_scan_tile<<<nb, SCAN_BLOCK, 0, st>>>(in, out, tsum, n);
  if (nb == 1) {
    k_scan_total<<<1, 1, 0, st>>>(out, n, tsum);
    return;
  }
  scan_excl(tsum, toff, nb, st, scratch + 2 * nb + 2);
  k_scan_add<<<nb, SCAN_BLOCK, 0, st>>>(out, toff, n);
  k_scan_total<<<1, 1, 0, st>>>(out, n, toff + nb);
}
