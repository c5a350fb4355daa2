// Teams: the threads that step one world together.  A block team is one workgroup (a batch launch steps one world
// per workgroup); a grid team is every workgroup of a cooperative launch stepping one large world.  The stage bodies
// (broadphase, contact sort, prepare, solve) are written once against a team's rank, size, barrier, prefix scan
// and radix sort, so both forms compute the same results: every cross-thread result of a body is an exact integer
// scan, a stable sort or an order-free atomic, whatever the number of threads.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#ifndef DEV
#define DEV __device__ __forceinline__
#endif

template <int BS>
DEV uint32_t block_excl_scan(uint32_t v, uint32_t* s_w, uint32_t& total);
template <int BS>
DEV uint32_t block_scan_array(uint32_t* a, uint32_t n, uint32_t* s_w);
template <int BS>
DEV void block_radix_sort(uint64_t* k0, uint32_t* v0, uint64_t* k1, uint32_t* v1, uint32_t n, uint32_t* s_hist, uint32_t* s_cnt,
                          uint64_t* s_red);

// LDS a team's scans use (per-wave partials), and the radix sort's digit counts, declared once per kernel; kernels
// whose bodies do not sort pass no sort LDS.
template <int BS>
struct TeamLds {
  uint32_t w[BS / 64];
};
template <int BS>
struct SortLds {
  uint32_t hist[256];
  uint32_t cnt[(BS / 64) * 256];
  uint64_t red[2 * (BS / 64)];
};

template <int BS>
struct BlockTeam {
  static constexpr bool GRID = false;
  static constexpr uint32_t NT = BS;
  uint32_t tid;
  TeamLds<BS>* L;
  SortLds<BS>* SL;
  DEV void sync() const { __syncthreads(); }
  DEV uint32_t scan(uint32_t* a, uint32_t n) const { return block_scan_array<BS>(a, n, L->w); }
  DEV void sort(uint64_t* k0, uint32_t* v0, uint64_t* k1, uint32_t* v1, uint32_t n) const {
    block_radix_sort<BS>(k0, v0, k1, v1, n, SL->hist, SL->cnt, SL->red);
  }
};

// Grid team scratch in global memory (words): [0] arrivals, [1] barrier generation, [2] set when a barrier gave up
// waiting, [3..10] words of the bodies (counters, flags, reductions), [11..15] reserved, then GRID_PART words of
// per-workgroup partials and 256 * gridDim.x radix counts.
static constexpr uint32_t GRID_HDR = 16, GRID_PART = 2048, GRID_MAX_GROUPS = 512;
__host__ __device__ inline size_t grid_scratch_words(uint32_t groups) { return GRID_HDR + GRID_PART + 256 * (size_t)groups; }

// Barrier over the workgroups of a cooperative launch.  Writes before it are visible to every workgroup after it
// (device-scope release and acquire fences around the arrival).  The wait is bounded: a barrier that waits past
// about a second records g[2] and returns, and every later barrier of the launch returns at once, so a launch
// whose workgroups disagree about the barriers ends instead of hanging (the host reports g[2] as an error).
DEV void grid_barrier(uint32_t* g, uint32_t groups) {
  __syncthreads();
  if (threadIdx.x == 0) {
    if (__hip_atomic_load(&g[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) {
      const uint32_t gen = __hip_atomic_load(&g[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __threadfence();
      if (atomicAdd(&g[0], 1u) == groups - 1) {
        atomicExch(&g[0], 0u);
        __threadfence();
        atomicAdd(&g[1], 1u);
      } else {
        for (uint32_t spin = 0; __hip_atomic_load(&g[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gen; spin++) {
          if (spin > (1u << 24)) { atomicOr(&g[2], 1u); break; }
          __builtin_amdgcn_s_sleep(2);
        }
      }
    }
    __threadfence();
  }
  __syncthreads();
}

template <int BS>
struct GridTeam {
  static constexpr bool GRID = true;
  uint32_t NT;    // threads of the launch
  uint32_t tid;   // this thread's rank in the launch
  TeamLds<BS>* L;
  SortLds<BS>* SL;
  uint32_t* g;    // scratch (grid_scratch_words)
  DEV void sync() const { grid_barrier(g, gridDim.x); }
  // In-place exclusive scan of a[0..n) (each thread a contiguous chunk, as block_scan_array); returns the total.
  DEV uint32_t scan(uint32_t* a, uint32_t n) const {
    const uint32_t per = (n + NT - 1) / NT;
    const uint32_t lo_ = min(tid * per, n), hi = min(lo_ + per, n);
    uint32_t s = 0;
    for (uint32_t i = lo_; i < hi; i++) s += a[i];
    uint32_t btot;
    uint32_t off = block_excl_scan<BS>(s, L->w, btot);
    uint32_t* part = g + GRID_HDR;
    if (threadIdx.x == 0) part[blockIdx.x] = btot;
    sync();
    uint32_t pre = 0, tot = 0;
    for (uint32_t b = 0; b < gridDim.x; b++) { const uint32_t t = part[b]; if (b < blockIdx.x) pre += t; tot += t; }
    off += pre;
    for (uint32_t i = lo_; i < hi; i++) { const uint32_t v = a[i]; a[i] = off; off += v; }
    sync();
    return tot;
  }
  // Stable LSD radix sort of n (key, value) pairs, 8-bit digits, digits no key varies in skipped; the sorted pairs
  // land in (k0, v0).  Workgroup b ranks the contiguous chunk b of the input; within a chunk, tiles of BS keep their
  // order and lanes rank by ballots, as block_radix_sort does, so the order is the stable one.
  DEV void sort(uint64_t* k0, uint32_t* v0, uint64_t* k1, uint32_t* v1, uint32_t n) const {
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    constexpr uint32_t NW = BS / 64;
    const uint32_t G = gridDim.x, chunk = (n + G - 1) / G;
    const uint32_t c0 = min(blockIdx.x * chunk, n), c1 = min(c0 + chunk, n);
    uint64_t kor = 0, kand = ~0ull;
    for (uint32_t i = tid; i < n; i += NT) { kor |= k0[i]; kand &= k0[i]; }
    for (int off = 32; off >= 1; off >>= 1) {
      kor |= (uint64_t)__shfl_xor((unsigned long long)kor, off, 64);
      kand &= (uint64_t)__shfl_xor((unsigned long long)kand, off, 64);
    }
    if (lane == 0) { SL->red[wave] = kor; SL->red[NW + wave] = kand; }
    __syncthreads();
    uint64_t* part = (uint64_t*)(g + GRID_HDR);   // per-workgroup (or, and) of the keys
    if (threadIdx.x == 0) {
      uint64_t o = 0, a = ~0ull;
      for (uint32_t k = 0; k < NW; k++) { o |= SL->red[k]; a &= SL->red[NW + k]; }
      part[2 * blockIdx.x] = o; part[2 * blockIdx.x + 1] = a;
    }
    sync();
    kor = 0; kand = ~0ull;
    for (uint32_t b = 0; b < G; b++) { kor |= part[2 * b]; kand &= part[2 * b + 1]; }
    const uint64_t vary = kor ^ kand;
    sync();  // every workgroup has read the partials before the scans reuse their words
    uint32_t* cnt = g + GRID_HDR + GRID_PART;       // counts, digit-major: cnt[d * G + b]
    uint64_t* src = k0; uint32_t* srcv = v0; uint64_t* dst = k1; uint32_t* dstv = v1;
    for (uint32_t sh = 0; sh < 64; sh += 8) {
      if (((vary >> sh) & 0xffull) == 0) continue;
      for (uint32_t d = threadIdx.x; d < 256; d += BS) SL->hist[d] = 0;
      __syncthreads();
      for (uint32_t i = c0 + threadIdx.x; i < c1; i += BS) atomicAdd(&SL->hist[(uint32_t)(src[i] >> sh) & 0xffu], 1u);
      __syncthreads();
      for (uint32_t d = threadIdx.x; d < 256; d += BS) cnt[d * G + blockIdx.x] = SL->hist[d];
      sync();  // every workgroup's counts are in place before the scan reads them
      scan(cnt, 256 * G);
      for (uint32_t d = threadIdx.x; d < 256; d += BS) SL->hist[d] = cnt[d * G + blockIdx.x];
      for (uint32_t i = threadIdx.x; i < NW * 256; i += BS) SL->cnt[i] = 0;
      __syncthreads();
      for (uint32_t base = c0; base < c1; base += BS) {
        const uint32_t i = base + threadIdx.x;
        const bool valid = i < c1;
        const uint64_t key = valid ? src[i] : 0;
        const uint32_t val = valid ? srcv[i] : 0;
        const uint32_t dg = (uint32_t)(key >> sh) & 0xffu;
        unsigned long long peers = __ballot(valid);
        for (uint32_t b = 0; b < 8; b++) {
          unsigned long long m = __ballot((dg >> b) & 1u);
          peers &= ((dg >> b) & 1u) ? m : ~m;
        }
        const uint32_t rank = (uint32_t)__popcll(peers & ((1ull << lane) - 1ull));
        if (valid && rank == 0) SL->cnt[wave * 256 + dg] = (uint32_t)__popcll(peers);
        __syncthreads();
        if (valid) {
          uint32_t pos = SL->hist[dg] + rank;
          for (uint32_t w2 = 0; w2 < wave; w2++) pos += SL->cnt[w2 * 256 + dg];
          dst[pos] = key; dstv[pos] = val;
        }
        __syncthreads();
        for (uint32_t d = threadIdx.x; d < 256; d += BS) {
          uint32_t s = 0;
          for (uint32_t w2 = 0; w2 < NW; w2++) { s += SL->cnt[w2 * 256 + d]; SL->cnt[w2 * 256 + d] = 0; }
          SL->hist[d] += s;
        }
        __syncthreads();
      }
      sync();
      uint64_t* tk = src; src = dst; dst = tk;
      uint32_t* tv = srcv; srcv = dstv; dstv = tv;
    }
    if (src != k0) {
      for (uint32_t i = tid; i < n; i += NT) { k0[i] = src[i]; v0[i] = srcv[i]; }
      sync();
    }
  }
};
