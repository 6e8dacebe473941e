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

// Single-launch scan (decoupled look-back).  Workgroups take tiles in the
// order they start (an atomic ticket), so every tile a workgroup waits on
// belongs to a workgroup that is already running.  Each tile publishes its
// total, then its inclusive prefix, in one 64-bit word: epoch (30 bits) |
// status (2 bits: 1 total, 2 prefix) | value (32 bits); words of an older
// epoch read as not yet published, so the state array is never cleared.
// A look-back that spins past a fixed bound sets *error and gives up, so
// the kernel always ends; the caller then reports the failure.
#define SCAN_SPIN_LIMIT (1u << 22)

__device__ __forceinline__ void scan_publish(unsigned long long* w, uint32_t epoch, uint32_t status, uint32_t value) {
  unsigned long long word = ((unsigned long long)(epoch & 0x3fffffffu) << 34) | ((unsigned long long)status << 32) | value;
  __hip_atomic_store(w, word, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ void __launch_bounds__(SCAN_BLOCK) k_scan_lookback(const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
                                                              uint32_t n, uint32_t ntiles, unsigned long long* state,
                                                              uint32_t* ticket, uint32_t epoch, uint32_t* error) {
  __shared__ uint32_t s[SCAN_BLOCK];
  __shared__ uint32_t s_tile, s_excl;
  const uint32_t tid = threadIdx.x;
  if (tid == 0) {
    uint32_t t = atomicAdd(ticket, 1u);
    if (t == ntiles - 1) atomicExch(ticket, 0u);  // last ticket of this launch: reset for the next scan
    s_tile = t;
  }
  __syncthreads();
  const uint32_t tile = s_tile;
  const uint32_t base = tile * SCAN_TILE + tid * SCAN_PER_THREAD;
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
  const uint32_t total = s[SCAN_BLOCK - 1];
  // Look-back by the first wavefront: lane i reads tile top - i, so 64
  // predecessors are examined per step; a window is usable once every tile
  // up to the nearest published prefix has published at least its total.
  if (tid < 64) {
    uint32_t excl = 0;
    if (tile == 0) {
      if (tid == 0) scan_publish(&state[0], epoch, 2u, total);
    } else {
      if (tid == 0) scan_publish(&state[tile], epoch, 1u, total);
      const uint32_t ep = epoch & 0x3fffffffu;
      int64_t top = (int64_t)tile - 1;
      uint32_t spins = 0;
      for (;;) {
        const int64_t j = top - (int64_t)tid;
        uint32_t st = 2u, val = 0u;  // before tile 0: an empty prefix
        if (j >= 0) {
          unsigned long long w = __hip_atomic_load(&state[j], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
          st = ((uint32_t)(w >> 34) == ep) ? ((uint32_t)(w >> 32) & 3u) : 0u;
          val = (uint32_t)w;
        }
        const unsigned long long pmask = __ballot(st == 2u);
        const unsigned long long vmask = __ballot(st != 0u);
        const int first = pmask ? __ffsll(pmask) - 1 : 64;  // nearest prefix in the window
        const unsigned long long need = first >= 63 ? ~0ull : ((2ull << first) - 1ull);
        if ((vmask & need) != need) {
          if (++spins > SCAN_SPIN_LIMIT) {
            if (tid == 0) atomicOr(error, 1u);
            break;
          }
          continue;
        }
        uint32_t part = (int)tid <= first ? val : 0u;
        for (int o = 32; o > 0; o >>= 1) part += __shfl_xor(part, o);
        excl += part;
        if (first < 64) break;
        top -= 64;
      }
      if (tid == 0) scan_publish(&state[tile], epoch, 2u, excl + total);
    }
    if (tid == 0) s_excl = excl;
  }
  __syncthreads();
  uint32_t run = s_excl + s[tid] - t;
#pragma unroll
  for (int k = 0; k < SCAN_PER_THREAD; ++k) {
    if (base + k < n) out[base + k] = run;
    run += v[k];
  }
  if (tile == ntiles - 1 && tid == 0) out[n] = s_excl + total;
}

// Device state of the single-launch scan (one per stream).
struct LookbackScan {
  unsigned long long* state = nullptr;  // one word per tile, zero-initialised
  uint32_t* ticket = nullptr;           // zero-initialised
  uint32_t max_tiles = 0;
  uint32_t epoch = 0;
};

// out must hold n + 1 entries; n <= max_tiles * SCAN_TILE.  *error becomes
// non-zero if a look-back gave up.
static inline void scan_lookback(LookbackScan& ls, const uint32_t* in, uint32_t* out, uint32_t n, hipStream_t st,
                                 uint32_t* error) {
  if (n == 0) {
    k_scan_zero<<<1, 1, 0, st>>>(out);
    return;
  }
  uint32_t nt = (n + SCAN_TILE - 1) / SCAN_TILE;
  ls.epoch = (ls.epoch + 1u) & 0x3fffffffu;
  if (ls.epoch == 0u) {  // epochs wrapped: clear words of the previous cycle
    (void)hipMemsetAsync(ls.state, 0, (size_t)ls.max_tiles * sizeof(unsigned long long), st);
    ls.epoch = 1u;
  }
  k_scan_lookback<<<nt, SCAN_BLOCK, 0, st>>>(in, out, n, nt, ls.state, ls.ticket, ls.epoch, error);
}

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
