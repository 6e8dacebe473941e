// avbd_scan.hpp — exclusive prefix sum of uint32 arrays on the device.
//
// out[0..n] receives the exclusive scan of in[0..n) with out[n] = total.
// Blocks of SCAN_BLOCK threads take tiles of SCAN_TILE elements, each thread
// a run of SCAN_PER_THREAD.  Integer addition is associative, so the result
// does not depend on the blocking.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define SCAN_BLOCK 256
#define SCAN_PER_THREAD 16
#define SCAN_TILE (SCAN_BLOCK * SCAN_PER_THREAD)

__global__ void k_scan_zero(uint32_t* __restrict__ out) { out[0] = 0u; }

// Two-launch scan (reduce, then scan).  Launch 1: every tile writes its
// sum; the last tile to finish (an atomic ticket after a release fence)
// turns the tile sums into exclusive tile offsets.  Launch 2: every tile
// scans itself in LDS and adds its offset.  No workgroup waits on another.
#define SCAN_MAX_TILES (SCAN_BLOCK * 64)  // tile offsets scanned by one workgroup: n <= 2^24

// Exclusive scan over the workgroup: wavefront scans by shuffles, then the
// wavefront totals through LDS (s holds SCAN_BLOCK / 64 words).  Integer
// sums, so the grouping does not matter.
__device__ __forceinline__ uint32_t block_scan_excl(uint32_t t, uint32_t* s, uint32_t* total) {
  const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
  uint32_t x = t;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o, 64);
    if (lane >= (uint32_t)o) x += y;
  }
  if (lane == 63u) s[w] = x;
  __syncthreads();
  uint32_t pre = 0, tot = 0;
#pragma unroll
  for (uint32_t k = 0; k < SCAN_BLOCK / 64; ++k) {
    const uint32_t v = s[k];
    if (k < w) pre += v;
    tot += v;
  }
  *total = tot;
  __syncthreads();  // s may be reused at once
  return pre + x - t;
}

// The SCAN_PER_THREAD consecutive entries of a thread (zero past n); whole
// 16-byte vectors when the run is in range and aligned.
__device__ __forceinline__ void load_run(const uint32_t* __restrict__ in, uint32_t base, uint32_t n,
                                         uint32_t (&v)[SCAN_PER_THREAD]) {
  if (base + SCAN_PER_THREAD <= n && ((reinterpret_cast<uintptr_t>(in + base) & 15u) == 0u)) {
    const uint4* p = reinterpret_cast<const uint4*>(in + base);
#pragma unroll
    for (int k = 0; k < SCAN_PER_THREAD / 4; ++k) {
      const uint4 q = p[k];
      v[4 * k] = q.x;
      v[4 * k + 1] = q.y;
      v[4 * k + 2] = q.z;
      v[4 * k + 3] = q.w;
    }
  } else {
#pragma unroll
    for (int k = 0; k < SCAN_PER_THREAD; ++k) v[k] = (base + k < n) ? in[base + k] : 0u;
  }
}

// Up to SCAN_JOBS arrays of the same length scanned by one launch pair
// (blockIdx.y selects the array); each has its own tile offsets and ticket.
#define SCAN_JOBS 3
struct ScanJobs {
  const uint32_t* in[SCAN_JOBS];
  uint32_t* out[SCAN_JOBS];
  uint32_t* pick = nullptr;  // optional, job 0 of 0/1 flags: pick[out[i]] = i for every set flag
  // optional: job j's total is also written to post[post_word + j] (pinned
  // host memory); the last job to get there sets post[post_flag] = post_seq,
  // for a host that polls (post_ticket: a zeroed device word)
  uint32_t* post = nullptr;
  uint32_t* post_ticket = nullptr;
  uint32_t post_word = 0, post_flag = 0, post_seq = 0;
  // optional: the length is n_dev[0] when that is at most the launch's n
  // (the launch is sized for n), else zero -- a count only the device holds
  const uint32_t* n_dev = nullptr;
  // optional: the scan runs only when 0 < gate[0] <= gate_max
  const uint32_t* gate = nullptr;
  uint32_t gate_max = 0;
};

// The scan's length on the device (see ScanJobs::n_dev, ::gate); false
// when the scan is skipped.
__device__ __forceinline__ bool scan_length(const ScanJobs& jobs, uint32_t* n, uint32_t* ntiles) {
  if (jobs.gate) {
    const uint32_t g = *jobs.gate;
    if (g == 0u || g > jobs.gate_max) return false;
  }
  if (jobs.n_dev) {
    const uint32_t m = *jobs.n_dev;
    *n = m <= *n ? m : 0u;
    *ntiles = (*n + SCAN_TILE - 1) / SCAN_TILE;
  }
  return true;
}

__global__ void __launch_bounds__(SCAN_BLOCK) k_scan_reduce(ScanJobs jobs, uint32_t n, uint32_t ntiles,
                                                            uint32_t* tiles, uint32_t* dones) {
  __shared__ uint32_t s[SCAN_BLOCK / 64];
  if (!scan_length(jobs, &n, &ntiles) || blockIdx.x >= ntiles) return;  // whole workgroups
  const uint32_t* __restrict__ in = jobs.in[blockIdx.y];
  uint32_t* tile = tiles + (size_t)blockIdx.y * (SCAN_MAX_TILES + 1);
  uint32_t* done = dones + blockIdx.y;
  __shared__ bool last;
  const uint32_t tid = threadIdx.x;
  const uint32_t base = blockIdx.x * SCAN_TILE + tid * SCAN_PER_THREAD;
  uint32_t v[SCAN_PER_THREAD];
  load_run(in, base, n, v);
  uint32_t t = 0;
#pragma unroll
  for (int k = 0; k < SCAN_PER_THREAD; ++k) t += v[k];
  uint32_t total;
  (void)block_scan_excl(t, s, &total);
  if (tid == 0) {
    __hip_atomic_store(&tile[blockIdx.x], total, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __threadfence();
    last = atomicAdd(done, 1u) == ntiles - 1;
  }
  __syncthreads();
  if (!last) return;
  // the last tile: exclusive offsets of all tiles (SCAN_MAX_TILES at most),
  // SCAN_BLOCK at a time
  __threadfence();
  uint32_t carry = 0;
  for (uint32_t c0 = 0; c0 < ntiles; c0 += SCAN_BLOCK) {
    const uint32_t q = c0 + tid;
    const uint32_t x = q < ntiles ? __hip_atomic_load(&tile[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0u;
    uint32_t sum;
    const uint32_t ex = block_scan_excl(x, s, &sum);
    if (q < ntiles) tile[q] = carry + ex;
    carry += sum;
  }
  if (tid == 0) {
    tile[ntiles] = carry;
    *done = 0u;  // ready for the next scan on this stream
  }
}

__global__ void __launch_bounds__(SCAN_BLOCK) k_scan_apply(ScanJobs jobs, uint32_t n, uint32_t ntiles,
                                                           const uint32_t* tiles) {
  __shared__ uint32_t s[SCAN_BLOCK / 64];
  const uint32_t* __restrict__ in = jobs.in[blockIdx.y];
  uint32_t* __restrict__ out = jobs.out[blockIdx.y];
  if (!scan_length(jobs, &n, &ntiles)) return;
  if (blockIdx.x >= ntiles) {  // whole workgroups; an empty scan still writes its total
    if (n == 0u && blockIdx.x == 0 && threadIdx.x == 0) out[0] = 0u;
    return;
  }
  const uint32_t* tile = tiles + (size_t)blockIdx.y * (SCAN_MAX_TILES + 1);
  const uint32_t tid = threadIdx.x;
  const uint32_t base = blockIdx.x * SCAN_TILE + tid * SCAN_PER_THREAD;
  uint32_t v[SCAN_PER_THREAD];
  load_run(in, base, n, v);
  uint32_t t = 0;
#pragma unroll
  for (int k = 0; k < SCAN_PER_THREAD; ++k) t += v[k];
  uint32_t total;
  uint32_t run = tile[blockIdx.x] + block_scan_excl(t, s, &total);
  if (jobs.pick && blockIdx.y == 0) {
    uint32_t r = run;
#pragma unroll
    for (int k = 0; k < SCAN_PER_THREAD; ++k) {
      if (v[k]) jobs.pick[r] = base + k;
      r += v[k];
    }
  }
  if (base + SCAN_PER_THREAD <= n && ((reinterpret_cast<uintptr_t>(out + base) & 15u) == 0u)) {
    uint4* p = reinterpret_cast<uint4*>(out + base);
#pragma unroll
    for (int k = 0; k < SCAN_PER_THREAD / 4; ++k) {
      uint4 q;
      q.x = run;
      run += v[4 * k];
      q.y = run;
      run += v[4 * k + 1];
      q.z = run;
      run += v[4 * k + 2];
      q.w = run;
      run += v[4 * k + 3];
      p[k] = q;
    }
  } else {
#pragma unroll
    for (int k = 0; k < SCAN_PER_THREAD; ++k) {
      if (base + k < n) out[base + k] = run;
      run += v[k];
    }
  }
  if (blockIdx.x == ntiles - 1 && tid == 0) {
    out[n] = tile[ntiles];
    if (jobs.post) {
      jobs.post[jobs.post_word + blockIdx.y] = tile[ntiles];
      __threadfence_system();
      // The ticket is taken with acquire-release at system scope: the job
      // that takes the last ticket sees the other jobs' post words (their
      // release) before its own release store raises the flag.
      if (__hip_atomic_fetch_add(jobs.post_ticket, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_SYSTEM) ==
          gridDim.y - 1u) {
        *jobs.post_ticket = 0u;  // ready for the next post
        __hip_atomic_store(jobs.post + jobs.post_flag, jobs.post_seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
      }
    }
  }
}

// Device state of the two-launch scan (one per stream).
struct TileScan {
  uint32_t* tile = nullptr;  // SCAN_JOBS x (SCAN_MAX_TILES + 1) tile sums / offsets
  uint32_t* done = nullptr;  // SCAN_JOBS zero-initialised tickets
};

// Scans the first `jobs` arrays of js, n entries each (outputs hold n + 1);
// n <= SCAN_MAX_TILES * SCAN_TILE.
static inline void scan_tiles(TileScan& ts, const ScanJobs& js, int jobs, uint32_t n, hipStream_t st) {
  if (n == 0) {
    for (int j = 0; j < jobs; ++j) k_scan_zero<<<1, 1, 0, st>>>(js.out[j]);
    return;
  }
  uint32_t nt = (n + SCAN_TILE - 1) / SCAN_TILE;
  k_scan_reduce<<<dim3(nt, jobs), SCAN_BLOCK, 0, st>>>(js, n, nt, ts.tile, ts.done);
  k_scan_apply<<<dim3(nt, jobs), SCAN_BLOCK, 0, st>>>(js, n, nt, ts.tile);
}
