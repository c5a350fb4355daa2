// Step kernels for gfx950.  Every kernel runs one world per workgroup over an array of worlds, so the
// plain per-world step is a one-block launch and a batch of thousands of uneven worlds is one launch of
// the same code.  Per substep: k_collide (shapes, broadphase, narrowphase, contact sort), k_prepare
// (articulation factor, grouping, reduction, rows, coloring, dense rows), k_solve (colored PGS with
// articulation propagation, integration, friction anchors), following the reference's kernels
// shapes/broad/narrow/csort (wgsl.js:117-1215, 3895-3980), prepare (1787-2494) and solve (3633-3891).
#include "steel_collide.cuh"
#include "steel_launch.h"
#include "steel_team.cuh"

#define AS_GLOBAL __attribute__((address_space(1)))
#define AS_LDS __attribute__((address_space(3)))

// ------------------------------------------------------------------------------------------------
// Block helpers
// ------------------------------------------------------------------------------------------------
// Exclusive scan of one value per thread; returns the exclusive prefix, writes the block total.
template <int BS>
DEV uint32_t block_excl_scan(uint32_t v, uint32_t* s_w, uint32_t& total) {
  uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t x = v;
  for (int off = 1; off < 64; off <<= 1) {
    uint32_t y = __shfl_up(x, off, 64);
    if (lane >= (uint32_t)off) x += y;
  }
  if (lane == 63) s_w[wave] = x;
  __syncthreads();
  uint32_t pre = 0, tot = 0;
  for (int k = 0; k < BS / 64; k++) { uint32_t t = s_w[k]; if (k < (int)wave) pre += t; tot += t; }
  __syncthreads();
  total = tot;
  return pre + x - v;
}
// In-place exclusive scan of a global array a[0..n); returns the total.
template <int BS>
DEV uint32_t block_scan_array(uint32_t* a, uint32_t n, uint32_t* s_w) {
  uint32_t per = (n + BS - 1) / BS;
  uint32_t lo_ = min(threadIdx.x * per, n), hi = min(lo_ + per, n);
  uint32_t s = 0;
  for (uint32_t i = lo_; i < hi; i++) s += a[i];
  uint32_t total;
  uint32_t off = block_excl_scan<BS>(s, s_w, total);
  for (uint32_t i = lo_; i < hi; i++) { uint32_t v = a[i]; a[i] = off; off += v; }
  __syncthreads();
  return total;
}

// Stable LSD radix sort of n (key, value) pairs, 8-bit digits, digits that no key varies in skipped.
// Ranks within a wave come from ballots (a peer mask per digit value), across waves from LDS counts.
// Sorted output lands in (k0, v0).  LDS: s_hist[256], s_cnt[BS/64][256], s_red[2*BS/64].
template <int BS>
DEV void block_radix_sort(uint64_t* k0, uint32_t* v0, uint64_t* k1, uint32_t* v1, uint32_t n, uint32_t* s_hist, uint32_t* s_cnt,
                          uint64_t* s_red) {
  const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  constexpr uint32_t NW = BS / 64;
  // bits any two keys differ in
  uint64_t kor = 0, kand = ~0ull;
  for (uint32_t i = tid; i < n; i += BS) { kor |= k0[i]; kand &= k0[i]; }
  for (int off = 32; off >= 1; off >>= 1) {
    kor |= (uint64_t)__shfl_xor((unsigned long long)kor, off, 64);
    kand &= (uint64_t)__shfl_xor((unsigned long long)kand, off, 64);
  }
  if (lane == 0) { s_red[wave] = kor; s_red[NW + wave] = kand; }
  __syncthreads();
  kor = 0; kand = ~0ull;
  for (uint32_t k = 0; k < NW; k++) { kor |= s_red[k]; kand &= s_red[NW + k]; }
  __syncthreads();
  uint64_t vary = kor ^ kand;
  uint64_t* src = k0; uint32_t* srcv = v0; uint64_t* dst = k1; uint32_t* dstv = v1;
  for (uint32_t sh = 0; sh < 64; sh += 8) {
    if (((vary >> sh) & 0xffull) == 0) continue;
    for (uint32_t d = tid; d < 256; d += BS) s_hist[d] = 0;
    __syncthreads();
    for (uint32_t i = tid; i < n; i += BS) atomicAdd(&s_hist[(uint32_t)(src[i] >> sh) & 0xffu], 1u);
    __syncthreads();
    if (tid == 0) { uint32_t acc = 0; for (uint32_t d = 0; d < 256; d++) { uint32_t c = s_hist[d]; s_hist[d] = acc; acc += c; } }
    for (uint32_t i = tid; i < NW * 256; i += BS) s_cnt[i] = 0;
    __syncthreads();
    for (uint32_t base = 0; base < n; base += BS) {
      uint32_t i = base + tid;
      bool valid = i < n;
      uint64_t key = valid ? src[i] : 0;
      uint32_t val = valid ? srcv[i] : 0;
      uint32_t dg = (uint32_t)(key >> sh) & 0xffu;
      unsigned long long peers = __ballot(valid);
      for (uint32_t b = 0; b < 8; b++) {
        unsigned long long m = __ballot((dg >> b) & 1u);
        peers &= ((dg >> b) & 1u) ? m : ~m;
      }
      uint32_t rank = (uint32_t)__popcll(peers & ((1ull << lane) - 1ull));
      if (valid && rank == 0) s_cnt[wave * 256 + dg] = (uint32_t)__popcll(peers);
      __syncthreads();
      if (valid) {
        uint32_t pos = s_hist[dg] + rank;
        for (uint32_t w2 = 0; w2 < wave; w2++) pos += s_cnt[w2 * 256 + dg];
        dst[pos] = key; dstv[pos] = val;
      }
      __syncthreads();
      for (uint32_t d = tid; d < 256; d += BS) {
        uint32_t s = 0;
        for (uint32_t w2 = 0; w2 < NW; w2++) { s += s_cnt[w2 * 256 + d]; s_cnt[w2 * 256 + d] = 0; }
        s_hist[d] += s;
      }
      __syncthreads();
    }
    uint64_t* tk = src; src = dst; dst = tk;
    uint32_t* tv = srcv; srcv = dstv; dstv = tv;
  }
  if (src != k0) {
    for (uint32_t i = tid; i < n; i += BS) { k0[i] = src[i]; v0[i] = srcv[i]; }
    __syncthreads();
  }
}

DEV void load_header(uint32_t* Hs, const WorldDev& wd) {
  for (uint32_t i = threadIdx.x; i < STEEL_WORLD_HEADER_WORDS; i += blockDim.x) Hs[i] = wd.hdr[i];
  __syncthreads();
}

// ------------------------------------------------------------------------------------------------
// k_collide
// ------------------------------------------------------------------------------------------------
DEV void shape_pose(const WC& w, uint32_t s) {  // wgsl.js:117-150
  uint32_t o = w.so(s);
  float4 r0 = w.sf(o);
  int body = ti(r0.x);
  uint32_t hull = tu(r0.y);
  f3 bp = mk3(0, 0, 0);
  q4 bq = mk4(0, 0, 0, 1);
  if (body >= 0) {
    f3 c = w.pose_c(body);
    bq = w.pose_q(body);
    bp = c - qrot(bq, w.body_com_local(body));
  }
  f3 sp = bp + qrot(bq, xyz(w.sf(o + 1)));
  q4 sq = nrm4(qmul(bq, Q(w.sf(o + 2))));
  uint32_t PS = w.h(W_P_SHAPE), PA = w.h(W_P_AABB);
  w.d.poses[PS + s * 2] = F4(sp, 0.0f);
  w.d.poses[PS + s * 2 + 1] = F4(sq);
  if (w.sf(o + 3).x == 4.0f) {
    w.d.poses[PA + s * 2] = make_float4(-1e6f, -1e6f, -1e6f, 0.0f);
    w.d.poses[PA + s * 2 + 1] = make_float4(1e6f, 1e6f, 1e6f, 0.0f);
    return;
  }
  f3 hc = xyz(w.sf(w.h(W_SF_HAABB) + hull * 2));
  f3 hh = xyz(w.sf(w.h(W_SF_HAABB) + hull * 2 + 1));
  m3 m = qmat(sq);
  f3 c = sp + mul(m, hc);
  f3 e = vabs(m.c[0]) * hh.x + vabs(m.c[1]) * hh.y + vabs(m.c[2]) * hh.z + mk3(r0.w, r0.w, r0.w);
  w.d.poses[PA + s * 2] = F4(c - e, 0.0f);
  w.d.poses[PA + s * 2 + 1] = F4(c + e, 0.0f);
}

// Shape tag for the pair predicate: body index + 1 in bits 0-30, bit 31 set when the shape cannot move
// (static shape, or static or kinematic body).
DEV uint32_t shape_tag(const WC& w, uint32_t s) {
  int b = ti(w.sf(w.so(s)).x);
  bool fixed = b < 0 || w.body_kind(b) == 3 || (w.body_flags(b) & 1) != 0;
  return (uint32_t)(b + 1) | (fixed ? 0x80000000u : 0u);
}

// The reference's pair predicate (wgsl.js:158-172) for shapes i < j with overlapping AABBs: different
// bodies, not both unmovable, not filtered (explicit list, joint parent/child, world and group rules, all in
// the filter bitset).
DEV bool pair_allowed(const WC& w, uint32_t i, uint32_t j, uint32_t tag_i, uint32_t tag_j) {
  if ((tag_i & 0x7fffffffu) == (tag_j & 0x7fffffffu)) return false;
  if ((tag_i & tag_j) >> 31) return false;
  const uint32_t F = w.h(W_SU_FILTER);
  if (w.h(W_FILT_LIST)) {  // shape i's sorted list of filtered partners j > i
    uint32_t lo_ = w.su(F + i), hi_ = w.su(F + i + 1);
    const uint32_t L = F + w.NS() + 1;
    while (lo_ < hi_) {
      const uint32_t m = (lo_ + hi_) >> 1, v = w.su(L + m);
      if (v == j) return false;
      if (v < j) lo_ = m + 1; else hi_ = m;
    }
    return true;
  }
  const uint64_t k = (uint64_t)i * w.NS() + j;  // past 65536 shapes the bit index needs 64 bits
  return (w.su(F + (uint32_t)(k >> 5)) & (1u << (uint32_t)(k & 31))) == 0;
}

// Waves per SIMD each kernel is compiled for (a VGPR budget: 512 / waves).  Every phase of a world is a
// latency-bound chain, so resident waves, not registers per lane, set the batch's throughput.
#ifndef STEEL_WPE_COLLIDE
#define STEEL_WPE_COLLIDE 4
#endif
#ifndef STEEL_WPE_PREPARE
#define STEEL_WPE_PREPARE 4
#endif
#ifndef STEEL_WPE_SOLVE
#define STEEL_WPE_SOLVE 4
#endif

// ---- cell broadphase (worlds with a W_CELL bound) ----
// A shape is binned when its AABB extent on every axis is at most the bound HB (a NaN extent is not).  With H the
// largest binned extent and cells of side Hc > H, two binned shapes whose AABBs overlap have min corners less than
// Hc apart on every axis, so their cells differ by at most one on each axis; cell coordinates are counted from the
// least binned cell and clamped to 21 bits, which keeps neighbours neighbours.  Cells are sorted by Morton code and
// found through an open-addressing table; each binned shape tests the shapes of its own cell that follow it and of
// the 13 cells after its own in (z, y, x) order, the shapes not binned test every shape.  Every candidate passes the
// reference's full test (wgsl.js:154-177), so the pair set is the reference's.
DEV bool cell_binned(const float4& lo, const float4& hi, float HB) {
  return hi.x - lo.x <= HB && hi.y - lo.y <= HB && hi.z - lo.z <= HB;
}
DEV int cell_floor(float x, double Hc) {  // floor(x / Hc) clamped to [-2^30, 2^30), in fp64 (error far below the margin)
  return (int)floor(fmin(fmax((double)x / Hc, -1073741824.0), 1073741823.0));
}
DEV uint64_t morton_spread(uint32_t v) {  // bit i of v (21 bits) to bit 3i
  uint64_t x = v & 0x1fffffu;
  x = (x | x << 32) & 0x1f00000000ffffull;
  x = (x | x << 16) & 0x1f0000ff0000ffull;
  x = (x | x << 8) & 0x100f00f00f00f00full;
  x = (x | x << 4) & 0x10c30c30c30c30c3ull;
  x = (x | x << 2) & 0x1249249249249249ull;
  return x;
}
DEV uint32_t morton_compact(uint64_t x) {  // inverse of morton_spread
  x &= 0x1249249249249249ull;
  x = (x ^ (x >> 2)) & 0x10c30c30c30c30c3ull;
  x = (x ^ (x >> 4)) & 0x100f00f00f00f00full;
  x = (x ^ (x >> 8)) & 0x1f0000ff0000ffull;
  x = (x ^ (x >> 16)) & 0x1f00000000ffffull;
  x = (x ^ (x >> 32)) & 0x1fffffull;
  return (uint32_t)x;
}
DEV uint64_t morton3(uint32_t x, uint32_t y, uint32_t z) { return morton_spread(x) | morton_spread(y) << 1 | morton_spread(z) << 2; }
static constexpr uint64_t CELL_NONE = 1ull << 63;  // key of a shape not binned: after every Morton code
// Table of 2^lg entries of two words: Morton code + 1 (0: empty), then first | end << 32 of the cell's sorted run.
// Reads go to the device-coherent level (the keys were placed by atomics).
DEV uint32_t cell_slot(uint64_t m, uint32_t lg) { return (uint32_t)(m ^ (m >> lg) ^ (m >> (2 * lg))) & ((1u << lg) - 1u); }
DEV void cell_insert(uint64_t* T, uint32_t lg, uint64_t m, uint32_t first, uint32_t end) {
  for (uint32_t e = cell_slot(m, lg);; e = (e + 1) & ((1u << lg) - 1u)) {
    if (atomicCAS((unsigned long long*)&T[2 * e], 0ull, (unsigned long long)(m + 1)) == 0ull) {
      atomicExch((unsigned long long*)&T[2 * e + 1], (unsigned long long)((uint64_t)first | (uint64_t)end << 32));
      return;
    }
  }
}
DEV bool cell_find(const uint64_t* T, uint32_t lg, uint64_t m, uint32_t& first, uint32_t& end) {
  for (uint32_t e = cell_slot(m, lg);; e = (e + 1) & ((1u << lg) - 1u)) {
    const uint64_t k = __hip_atomic_load(&T[2 * e], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (k == 0) return false;
    if (k == m + 1) {
      const uint64_t v = __hip_atomic_load(&T[2 * e + 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      first = (uint32_t)v; end = (uint32_t)(v >> 32);
      return true;
    }
  }
}

// The cell broadphase's phases, each a loop over its items by thread `tid` of `NT`; a team runs them with a barrier
// between phases, the device-wide form (a large world alone in its batch) as one launch each.  aux: the phases'
// words, set before the first: [0] pair count 0, [1] largest binned extent 0, [2..4] least cell per axis
// 0xffffffff, [5] binned count 0.
DEV double cell_side(const uint32_t* aux, float HB) { const float H = ubits(aux[1]); return (double)(H > 0.0f ? H : HB) * 1.0078125; }
DEV uint32_t cell_lg(uint32_t nsm) { uint32_t lg = 1; while ((1u << lg) < 2 * nsm) lg++; return lg; }
// Each thread reduces over its shapes, then each wave (all lanes reach the end of the loop) before one lane updates
// the shared words: a word in global memory is then updated once per wave, not once per shape.
DEV uint32_t wave_max_u(uint32_t x) { for (int o = 32; o >= 1; o >>= 1) x = max(x, (uint32_t)__shfl_xor((int)x, o, 64)); return x; }
DEV uint32_t wave_min_u(uint32_t x) { for (int o = 32; o >= 1; o >>= 1) x = min(x, (uint32_t)__shfl_xor((int)x, o, 64)); return x; }
DEV uint32_t wave_sum_u(uint32_t x) { for (int o = 32; o >= 1; o >>= 1) x += (uint32_t)__shfl_xor((int)x, o, 64); return x; }
DEV void cells_extent(uint32_t tid, uint32_t NT, const WC& w, const WorldDev& wd, uint32_t* aux, float HB) {
  const uint32_t NS = w.NS(), PA = w.h(W_P_AABB);
  uint32_t m = 0;  // extents are >= 0, so their bits order as the floats do
  for (uint32_t s = tid; s < NS; s += NT) {
    const float4 lo = wd.poses[PA + s * 2], hi = wd.poses[PA + s * 2 + 1];
    if (cell_binned(lo, hi, HB)) m = max(m, fbits(fmaxf(fmaxf(hi.x - lo.x, hi.y - lo.y), hi.z - lo.z)));
  }
  m = wave_max_u(m);
  if ((threadIdx.x & 63) == 0 && m) atomicMax(&aux[1], m);
}
DEV void cells_origin(uint32_t tid, uint32_t NT, const WC& w, const WorldDev& wd, uint32_t* aux, float HB) {
  const uint32_t NS = w.NS(), PA = w.h(W_P_AABB);
  const double Hc = cell_side(aux, HB);
  uint32_t mx_ = 0xffffffffu, my_ = 0xffffffffu, mz_ = 0xffffffffu, n = 0;
  for (uint32_t s = tid; s < NS; s += NT) {
    const float4 lo = wd.poses[PA + s * 2], hi = wd.poses[PA + s * 2 + 1];
    if (!cell_binned(lo, hi, HB)) continue;
    mx_ = min(mx_, (uint32_t)(cell_floor(lo.x, Hc) + (1 << 30)));
    my_ = min(my_, (uint32_t)(cell_floor(lo.y, Hc) + (1 << 30)));
    mz_ = min(mz_, (uint32_t)(cell_floor(lo.z, Hc) + (1 << 30)));
    n++;
  }
  mx_ = wave_min_u(mx_); my_ = wave_min_u(my_); mz_ = wave_min_u(mz_); n = wave_sum_u(n);
  if ((threadIdx.x & 63) == 0 && n) {
    atomicMin(&aux[2], mx_); atomicMin(&aux[3], my_); atomicMin(&aux[4], mz_);
    atomicAdd(&aux[5], n);
  }
}
// vary (device-wide form): or and and of the keys, for the sort's digit skipping.
DEV void cells_keys(uint32_t tid, uint32_t NT, const WC& w, const WorldDev& wd, const uint32_t* aux, float HB, uint64_t* vary) {
  const uint32_t NS = w.NS(), PA = w.h(W_P_AABB);
  const double Hc = cell_side(aux, HB);
  const uint32_t ox = aux[2], oy = aux[3], oz = aux[4];
  uint64_t kor = 0, kand = ~0ull;
  for (uint32_t s = tid; s < NS; s += NT) {
    const float4 lo = wd.poses[PA + s * 2], hi = wd.poses[PA + s * 2 + 1];
    uint64_t key = CELL_NONE;
    if (cell_binned(lo, hi, HB))
      key = morton3(min((uint32_t)(cell_floor(lo.x, Hc) + (1 << 30)) - ox, 0x1fffffu),
                    min((uint32_t)(cell_floor(lo.y, Hc) + (1 << 30)) - oy, 0x1fffffu),
                    min((uint32_t)(cell_floor(lo.z, Hc) + (1 << 30)) - oz, 0x1fffffu));
    wd.skey[s] = key; wd.sval[s] = s;
    kor |= key; kand &= key;
  }
  if (vary) {
    for (int o = 32; o >= 1; o >>= 1) {
      kor |= (uint64_t)__shfl_xor((unsigned long long)kor, o, 64);
      kand &= (uint64_t)__shfl_xor((unsigned long long)kand, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
      atomicOr((unsigned long long*)&vary[0], (unsigned long long)kor);
      atomicAnd((unsigned long long*)&vary[1], (unsigned long long)kand);
    }
  }
}
// Sorted bounds (lo.w: shape id, hi.w: pair tag) in the raw contact buffer, the cell table (at most 4 NS entries of
// 16 bytes) cleared in the contact buffer; both buffers hold 96 MAXC bytes and MAXC >= NS.
DEV void cells_gather(uint32_t tid, uint32_t NT, const WC& w, const WorldDev& wd, const uint32_t* aux) {
  const uint32_t NS = w.NS(), PA = w.h(W_P_AABB), lg = cell_lg(aux[5]);
  float4* s_lo = wd.craw;
  float4* s_hi = s_lo + NS;
  uint64_t* T = (uint64_t*)wd.contacts;
  for (uint32_t p = tid; p < NS; p += NT) {
    const uint32_t si = wd.sval[p];
    float4 lo = wd.poses[PA + si * 2], hi = wd.poses[PA + si * 2 + 1];
    lo.w = ubits(si);
    hi.w = ubits(shape_tag(w, si));
    s_lo[p] = lo; s_hi[p] = hi;
  }
  for (uint32_t e = tid; e < (1u << lg); e += NT) T[2 * e] = 0;
}
DEV void cells_insert(uint32_t tid, uint32_t NT, const WorldDev& wd, const uint32_t* aux) {
  const uint32_t nsm = aux[5], lg = cell_lg(nsm);
  uint64_t* T = (uint64_t*)wd.contacts;
  for (uint32_t p = tid; p < nsm; p += NT) {
    const uint64_t m = wd.skey[p];
    if (p > 0 && wd.skey[p - 1] == m) continue;
    uint32_t q = p + 1;
    while (q < nsm && wd.skey[q] == m) q++;
    cell_insert(T, lg, m, p, q);
  }
}
// The reference's test of one candidate; the pairs that pass are appended by their wave with one counter update
// (slots in lane order), so a counter in global memory is not hit once per pair.  Called by every active lane.
DEV void cell_test(const WC& w, const WorldDev& wd, uint32_t* np, uint32_t MAXP, const float4& lo_i, const float4& hi_i,
                   const float4& lo_j, const float4& hi_j) {
  const uint32_t si = fbits(lo_i.w), sj = fbits(lo_j.w), a = min(si, sj), b = max(si, sj);
  const bool hit = !(lo_i.x > hi_j.x || lo_i.y > hi_j.y || lo_i.z > hi_j.z || lo_j.x > hi_i.x || lo_j.y > hi_i.y || lo_j.z > hi_i.z) &&
                   pair_allowed(w, a, b, fbits(hi_i.w), fbits(hi_j.w));
  const unsigned long long m = __ballot(hit);
  if (m == 0) return;
  const uint32_t lane = threadIdx.x & 63, leader = (uint32_t)__ffsll((long long)m) - 1;
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(np, (uint32_t)__popcll(m));
  base = __shfl(base, (int)leader, 64);
  if (hit) {
    const uint32_t slot = base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
    if (slot < MAXP) { wd.pairs[slot * 2] = a; wd.pairs[slot * 2 + 1] = b; }
  }
}
DEV void cells_pairs(uint32_t tid, uint32_t NT, const WC& w, const WorldDev& wd, uint32_t* aux) {
  const uint32_t NS = w.NS(), MAXP = w.MAXP(), nsm = aux[5], lg = cell_lg(nsm);
  const float4* s_lo = wd.craw;
  const float4* s_hi = s_lo + NS;
  const uint64_t* T = (const uint64_t*)wd.contacts;
  for (uint32_t p = tid; p < nsm; p += NT) {
    const uint64_t m = wd.skey[p];
    const uint32_t cx = morton_compact(m), cy = morton_compact(m >> 1), cz = morton_compact(m >> 2);
    const float4 lo_p = s_lo[p], hi_p = s_hi[p];
    for (uint32_t t = 0; t < 14; t++) {  // own cell, then the 13 cells after it: dz = 1 (9), dz = 0 dy = 1 (3), dx = 1
      const int dz = t == 0 ? 0 : (t <= 9 ? 1 : 0);
      const int dy = t == 0 ? 0 : (t <= 9 ? (int)((t - 1) / 3) - 1 : (t <= 12 ? 1 : 0));
      const int dx = t == 0 ? 0 : (t <= 9 ? (int)((t - 1) % 3) - 1 : (t <= 12 ? (int)(t - 10) - 1 : 1));
      const uint32_t nx = cx + (uint32_t)dx, ny = cy + (uint32_t)dy, nz = cz + (uint32_t)dz;
      if ((nx | ny | nz) > 0x1fffffu) continue;
      uint32_t q0 = p + 1, q1 = 0;
      if (t == 0) {
        q1 = q0;
        while (q1 < nsm && wd.skey[q1] == m) q1++;
      } else if (!cell_find(T, lg, morton3(nx, ny, nz), q0, q1)) {
        continue;
      }
      for (uint32_t q = q0; q < q1; q++) cell_test(w, wd, &aux[0], MAXP, lo_p, hi_p, s_lo[q], s_hi[q]);
    }
  }
}
DEV void cells_big(uint32_t tid, uint32_t NT, const WC& w, const WorldDev& wd, uint32_t* aux) {
  const uint32_t NS = w.NS(), MAXP = w.MAXP(), nsm = aux[5];
  const float4* s_lo = wd.craw;
  const float4* s_hi = s_lo + NS;
  for (uint32_t b = nsm; b < NS; b++) {  // shapes not binned, against every shape (a pair of two of them once)
    const float4 lo_b = s_lo[b], hi_b = s_hi[b];
    for (uint32_t q = tid; q < NS; q += NT)
      if (q < nsm || q > b) cell_test(w, wd, &aux[0], MAXP, lo_b, hi_b, s_lo[q], s_hi[q]);
  }
}

template <int BS, class TM>
DEV void broad_cells(const TM& tm, const WC& w, const WorldDev& wd, uint32_t* aux, float HB) {
  const uint32_t tid = tm.tid, NT = tm.NT;
  cells_extent(tid, NT, w, wd, aux, HB);
  tm.sync();
  cells_origin(tid, NT, w, wd, aux, HB);
  tm.sync();
  cells_keys(tid, NT, w, wd, aux, HB, nullptr);
  tm.sync();
  tm.sort(wd.skey, wd.sval, wd.skey + w.MAXC(), wd.sval + w.MAXC(), w.NS());
  cells_gather(tid, NT, w, wd, aux);
  tm.sync();
  cells_insert(tid, NT, wd, aux);
  tm.sync();
  cells_pairs(tid, NT, w, wd, aux);
  cells_big(tid, NT, w, wd, aux);
  tm.sync();
  if (tid == 0) { wd.ctr[0] = aux[0]; wd.ctr[1] = 0; }
}

// Shapes and broadphase of one world by a team: sort shapes by the order-preserving integer of AABB min x, then
// sweep the sorted list; pair (i, j) survives iff it passes the reference's all-pairs predicate (wgsl.js:154-177),
// tested in full for every candidate the sweep reaches.  A block team keeps the sorted bounds and runs of worlds of
// up to `cap` shapes in LDS (s_box); larger worlds, and a grid team, keep them in the world's contact buffers (free
// until the narrowphase).  The flattened candidates are dealt out over the team.
// np: the team's pair counter and, for the cell broadphase, five more words (LDS for a block, global for a grid).
template <int BS, class TM>
DEV void broad_world(const TM& tm, const WorldDev wd, uint32_t* Hs, uint32_t cap, float4* s_box, uint32_t* np) {
  load_header(Hs, wd);
  WC w = make_world<false>(Hs, wd);
  const uint32_t tid = tm.tid, NT = tm.NT, NS = w.NS();
  const float HB = w.hf(W_CELL);
  if (tid == 0) {
    np[0] = 0;
    if (HB > 0.0f) { np[1] = 0; np[2] = np[3] = np[4] = 0xffffffffu; np[5] = 0; }
  }
  for (uint32_t s = tid; s < NS; s += NT) shape_pose(w, s);
  tm.sync();
  if (HB > 0.0f) {
    broad_cells<BS>(tm, w, wd, np, HB);
    return;
  }
  const uint32_t PA = w.h(W_P_AABB);
  for (uint32_t s = tid; s < NS; s += NT) { wd.skey[s] = fflip_u(fbits(wd.poses[PA + s * 2].x)); wd.sval[s] = s; }
  tm.sync();
  tm.sort(wd.skey, wd.sval, wd.skey + w.MAXC(), wd.sval + w.MAXC(), NS);
  const bool glob = TM::GRID || NS > cap;
  float4* s_lo = glob ? wd.craw : s_box;
  float4* s_hi = s_lo + (glob ? NS : cap);
  uint32_t* s_run = glob ? (uint32_t*)wd.contacts : (uint32_t*)(s_box + 2 * cap);
  const uint32_t MAXP = w.MAXP();
  // Candidate (i, j), j after i in x order with min x_j <= max x_i.  The x intervals overlap (min x_i <= min
  // x_j <= max x_j), so the AABB test is y and z, then the rest of the predicate.
  auto test = [&](const float4& lo_i, const float4& hi_i, uint32_t si, uint32_t tag_i, const float4& lo_j, const float4& hi_j,
                  uint32_t sj, uint32_t tag_j) {
    if (lo_i.y > hi_j.y || lo_i.z > hi_j.z || lo_j.y > hi_i.y || lo_j.z > hi_i.z) return;
    uint32_t a = min(si, sj), b = max(si, sj);
    if (!pair_allowed(w, a, b, tag_i, tag_j)) return;
    uint32_t slot = atomicAdd(np, 1u);
    if (slot < MAXP) { wd.pairs[slot * 2] = a; wd.pairs[slot * 2 + 1] = b; }
  };
  for (uint32_t i = tid; i < NS; i += NT) {
    uint32_t si = wd.sval[i];
    float4 lo = wd.poses[PA + si * 2], hi = wd.poses[PA + si * 2 + 1];
    lo.w = ubits(si);
    hi.w = ubits(shape_tag(w, si));
    s_lo[i] = lo; s_hi[i] = hi;
  }
  tm.sync();
  // Run length of each sorted shape (binary search for the first later min x above its max x), then the
  // candidates are flattened by an exclusive scan and dealt out to the lanes in chunks of 8, so one shape
  // with a long run (a ground plane spans every x) is shared by the whole team.
  for (uint32_t i = tid; i < NS; i += NT) {
    const float hx = s_hi[i].x;
    uint32_t lo_ = i + 1, hi_ = NS;
    while (lo_ < hi_) {
      uint32_t m = (lo_ + hi_) >> 1;
      if (s_lo[m].x > hx) hi_ = m; else lo_ = m + 1;
    }
    s_run[i] = lo_ - i - 1;
  }
  tm.sync();
  const uint32_t T = tm.scan(s_run, NS);
  if (tid == 0) s_run[NS] = T;
  tm.sync();
  constexpr uint32_t CH = 8;
  for (uint32_t c = tid; c < (T + CH - 1) / CH; c += NT) {
    const uint32_t k0 = c * CH, k1 = min(k0 + CH, T);
    uint32_t lo_ = 0, hi_ = NS - 1;  // largest i with s_run[i] <= k0 (s_run[0] = 0, s_run[NS] = T > k0)
    while (lo_ < hi_) {
      uint32_t m = (lo_ + hi_ + 1) >> 1;
      if (s_run[m] <= k0) lo_ = m; else hi_ = m - 1;
    }
    uint32_t i = lo_;
    float4 lo_i = s_lo[i], hi_i = s_hi[i];
    for (uint32_t k = k0; k < k1; k++) {
      if (s_run[i + 1] <= k) {
        do { i++; } while (s_run[i + 1] <= k);
        lo_i = s_lo[i]; hi_i = s_hi[i];
      }
      const uint32_t j = i + 1 + (k - s_run[i]);
      const float4 lo_j = s_lo[j], hi_j = s_hi[j];
      test(lo_i, hi_i, fbits(lo_i.w), fbits(hi_i.w), lo_j, hi_j, fbits(lo_j.w), fbits(hi_j.w));
    }
  }
  tm.sync();
  if (tid == 0) { wd.ctr[0] = *np; wd.ctr[1] = 0; }
}

// One world per workgroup.  Dynamic LDS for worlds of up to `cap` shapes: the sorted AABBs (lo.w holds the shape
// id) and each sorted shape's first candidate in the flattened candidate list (cap + 1 words).
template <int BS>
__global__ void __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(STEEL_WPE_COLLIDE))) k_broad(const WorldDev* __restrict__ worlds, uint32_t nworlds, uint32_t cap) {
  __shared__ uint32_t Hs[STEEL_WORLD_HEADER_WORDS];
  __shared__ TeamLds<BS> tl;
  __shared__ SortLds<BS> sl;
  __shared__ uint32_t s_np[6];
  extern __shared__ float4 s_box[];
  const uint32_t wid = blockIdx.x;
  if (wid >= nworlds) return;
  const BlockTeam<BS> tm{threadIdx.x, &tl, &sl};
  broad_world<BS>(tm, worlds[wid], Hs, cap, s_box, s_np);
}

// One large world over a cooperative launch (g: grid team scratch; g[3..8] the pair counter and cell words).
template <int BS>
__global__ void __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(STEEL_WPE_COLLIDE))) k_broad_grid(const WorldDev* __restrict__ world, uint32_t* g) {
  __shared__ uint32_t Hs[STEEL_WORLD_HEADER_WORDS];
  __shared__ TeamLds<BS> tl;
  __shared__ SortLds<BS> sl;
  const GridTeam<BS> tm{gridDim.x * BS, blockIdx.x * BS + threadIdx.x, &tl, &sl, g};
  broad_world<BS>(tm, world[0], Hs, 0, nullptr, g + 3);
}

// ---- Device-wide cell broadphase of a large world alone in its batch (world 0): the phases above, a launch each,
// and a stable LSD radix sort of the Morton keys by digit-count, scan and scatter launches (digits no key varies in
// skipped).  Scratch: the contact buffer past the cell table (which takes at most 64 NS of its 96 MAXC bytes):
// [0..15] the phases' words, [16..19] the keys' or and and, [64..] 256 digit counts per sort chunk.
static constexpr uint32_t CELLW_BS = 256, CELLW_MAXG = 128;
DEV uint32_t* cellw_scratch(const WorldDev& wd, uint32_t MAXC) { return (uint32_t*)(wd.contacts + 4 * (size_t)MAXC); }
DEV uint32_t cellw_chunk(uint32_t NS) {  // keys per sort chunk: a multiple of the block, at most CELLW_MAXG chunks
  uint32_t ch = 4 * CELLW_BS;
  while ((NS + ch - 1) / ch > CELLW_MAXG) ch += CELLW_BS;
  return ch;
}
#define CELLW_VIEW                                          \
  __shared__ uint32_t Hs[STEEL_WORLD_HEADER_WORDS];         \
  const WorldDev wd = world[0];                             \
  load_header(Hs, wd);                                      \
  const WC w = make_world<false>(Hs, wd);                   \
  const float HB = w.hf(W_CELL);                            \
  uint32_t* aux = cellw_scratch(wd, w.MAXC());              \
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x, NT = gridDim.x * blockDim.x;
__global__ void __launch_bounds__(64) k_cellw_init(const WorldDev* __restrict__ world) {
  CELLW_VIEW
  (void)HB; (void)NT;
  if (tid == 0) {
    aux[0] = 0; aux[1] = 0; aux[2] = aux[3] = aux[4] = 0xffffffffu; aux[5] = 0;
    uint64_t* vary = (uint64_t*)(aux + 16);
    vary[0] = 0; vary[1] = ~0ull;
  }
}
__global__ void __launch_bounds__(CELLW_BS) k_cellw_pose(const WorldDev* __restrict__ world) {
  CELLW_VIEW
  for (uint32_t s = tid; s < w.NS(); s += NT) shape_pose(w, s);
  cells_extent(tid, NT, w, wd, aux, HB);  // the same thread reads back the boxes it wrote
}
__global__ void __launch_bounds__(CELLW_BS) k_cellw_origin(const WorldDev* __restrict__ world) { CELLW_VIEW cells_origin(tid, NT, w, wd, aux, HB); }
__global__ void __launch_bounds__(CELLW_BS) k_cellw_keys(const WorldDev* __restrict__ world) {
  CELLW_VIEW
  cells_keys(tid, NT, w, wd, aux, HB, (uint64_t*)(aux + 16));
}
// The pass on bits [sh, sh + 8): whether any key varies there, and whether the keys are in (skey, sval) before it
// (an even number of earlier passes ran) or in the second halves.
DEV bool cellw_pass(const uint32_t* aux, uint32_t sh, bool& in_first) {
  const uint64_t* vary = (const uint64_t*)(aux + 16);
  const uint64_t v = vary[0] ^ vary[1];
  uint32_t before = 0;
  for (uint32_t d = 0; d < sh; d += 8) before += ((v >> d) & 0xffull) != 0;
  in_first = (before & 1u) == 0;
  return ((v >> sh) & 0xffull) != 0;
}
__global__ void __launch_bounds__(CELLW_BS) k_cellw_count(const WorldDev* __restrict__ world, uint32_t sh) {
  CELLW_VIEW
  (void)HB; (void)tid; (void)NT;
  __shared__ uint32_t hist[256];
  bool in_first;
  if (!cellw_pass(aux, sh, in_first)) return;
  const uint32_t NS = w.NS(), MAXC = w.MAXC(), ch = cellw_chunk(NS), G = (NS + ch - 1) / ch, b = blockIdx.x;
  if (b >= G) return;
  const uint64_t* src = in_first ? wd.skey : wd.skey + MAXC;
  for (uint32_t d = threadIdx.x; d < 256; d += CELLW_BS) hist[d] = 0;
  __syncthreads();
  const uint32_t c0 = b * ch, c1 = min(c0 + ch, NS);
  for (uint32_t i = c0 + threadIdx.x; i < c1; i += CELLW_BS) atomicAdd(&hist[(uint32_t)(src[i] >> sh) & 0xffu], 1u);
  __syncthreads();
  for (uint32_t d = threadIdx.x; d < 256; d += CELLW_BS) aux[64 + d * G + b] = hist[d];
}
__global__ void __launch_bounds__(1024) k_cellw_scan(const WorldDev* __restrict__ world, uint32_t sh) {
  CELLW_VIEW
  (void)HB; (void)tid; (void)NT;
  __shared__ uint32_t s_w[1024 / 64];
  bool in_first;
  if (!cellw_pass(aux, sh, in_first)) return;
  const uint32_t NS = w.NS(), ch = cellw_chunk(NS), G = (NS + ch - 1) / ch;
  block_scan_array<1024>(aux + 64, 256 * G, s_w);
}
__global__ void __launch_bounds__(CELLW_BS) k_cellw_scatter(const WorldDev* __restrict__ world, uint32_t sh) {
  CELLW_VIEW
  (void)HB; (void)tid; (void)NT;
  constexpr uint32_t NW = CELLW_BS / 64;
  __shared__ uint32_t hist[256];
  __shared__ uint32_t wcnt[NW * 256];
  bool in_first;
  if (!cellw_pass(aux, sh, in_first)) return;
  const uint32_t NS = w.NS(), MAXC = w.MAXC(), ch = cellw_chunk(NS), G = (NS + ch - 1) / ch, b = blockIdx.x;
  if (b >= G) return;
  const uint64_t* src = in_first ? wd.skey : wd.skey + MAXC;
  const uint32_t* srcv = in_first ? wd.sval : wd.sval + MAXC;
  uint64_t* dst = in_first ? wd.skey + MAXC : wd.skey;
  uint32_t* dstv = in_first ? wd.sval + MAXC : wd.sval;
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (uint32_t d = threadIdx.x; d < 256; d += CELLW_BS) hist[d] = aux[64 + d * G + b];
  for (uint32_t i = threadIdx.x; i < NW * 256; i += CELLW_BS) wcnt[i] = 0;
  __syncthreads();
  const uint32_t c0 = b * ch, c1 = min(c0 + ch, NS);
  for (uint32_t base = c0; base < c1; base += CELLW_BS) {  // tiles keep their order; lanes rank by ballots
    const uint32_t i = base + threadIdx.x;
    const bool valid = i < c1;
    const uint64_t key = valid ? src[i] : 0;
    const uint32_t val = valid ? srcv[i] : 0;
    const uint32_t dg = (uint32_t)(key >> sh) & 0xffu;
    unsigned long long peers = __ballot(valid);
    for (uint32_t k = 0; k < 8; k++) {
      unsigned long long m = __ballot((dg >> k) & 1u);
      peers &= ((dg >> k) & 1u) ? m : ~m;
    }
    const uint32_t rank = (uint32_t)__popcll(peers & ((1ull << lane) - 1ull));
    if (valid && rank == 0) wcnt[wave * 256 + dg] = (uint32_t)__popcll(peers);
    __syncthreads();
    if (valid) {
      uint32_t pos = hist[dg] + rank;
      for (uint32_t w2 = 0; w2 < wave; w2++) pos += wcnt[w2 * 256 + dg];
      dst[pos] = key; dstv[pos] = val;
    }
    __syncthreads();
    for (uint32_t d = threadIdx.x; d < 256; d += CELLW_BS) {
      uint32_t sum = 0;
      for (uint32_t w2 = 0; w2 < NW; w2++) { sum += wcnt[w2 * 256 + d]; wcnt[w2 * 256 + d] = 0; }
      hist[d] += sum;
    }
    __syncthreads();
  }
}
__global__ void __launch_bounds__(CELLW_BS) k_cellw_sortfix(const WorldDev* __restrict__ world) {
  CELLW_VIEW
  (void)HB;
  bool last_first;  // where the keys were before the last digit's pass, and whether that pass ran
  const bool last = cellw_pass(aux, 56, last_first);
  const bool sorted_in_first = last ? !last_first : last_first;
  if (sorted_in_first) return;
  const uint32_t MAXC = w.MAXC();
  for (uint32_t i = tid; i < w.NS(); i += NT) { wd.skey[i] = wd.skey[MAXC + i]; wd.sval[i] = wd.sval[MAXC + i]; }
}
__global__ void __launch_bounds__(CELLW_BS) k_cellw_gather(const WorldDev* __restrict__ world) { CELLW_VIEW (void)HB; cells_gather(tid, NT, w, wd, aux); }
__global__ void __launch_bounds__(CELLW_BS) k_cellw_insert(const WorldDev* __restrict__ world) { CELLW_VIEW (void)HB; cells_insert(tid, NT, wd, aux); }
__global__ void __launch_bounds__(CELLW_BS) k_cellw_pairs(const WorldDev* __restrict__ world) { CELLW_VIEW (void)HB; cells_pairs(tid, NT, w, wd, aux); }
__global__ void __launch_bounds__(CELLW_BS) k_cellw_big(const WorldDev* __restrict__ world) { CELLW_VIEW (void)HB; cells_big(tid, NT, w, wd, aux); }
__global__ void __launch_bounds__(64) k_cellw_finish(const WorldDev* __restrict__ world) {
  CELLW_VIEW
  (void)HB; (void)NT;
  if (tid == 0) { wd.ctr[0] = aux[0]; wd.ctr[1] = 0; }
}

// Exclusive prefix of the worlds' pair counts (min(ctr[0], MAXP)) into offs[0..n], one block.
template <int BS>
__global__ void __launch_bounds__(BS) k_pair_offsets(const WorldDev* __restrict__ worlds, uint32_t nworlds, uint32_t* __restrict__ offs) {
  __shared__ uint32_t s_w[BS / 64];
  const uint32_t per = (nworlds + BS - 1) / BS;
  const uint32_t lo_ = min(threadIdx.x * per, nworlds), hi = min(lo_ + per, nworlds);
  uint32_t sum = 0;
  for (uint32_t i = lo_; i < hi; i++) sum += min(worlds[i].ctr[0], worlds[i].hdr[W_MAXP]);
  uint32_t total;
  uint32_t off = block_excl_scan<BS>(sum, s_w, total);
  for (uint32_t i = lo_; i < hi; i++) { offs[i] = off; off += min(worlds[i].ctr[0], worlds[i].hdr[W_MAXP]); }
  if (threadIdx.x == 0) offs[nworlds] = total;
}

// Narrowphase over every candidate pair of every world: one wave64 per pair (wgsl.js:1063-1195).  The
// pairs are laid end to end in world order (offs) and cut into chunks of CH = clamp(total / waves, 1, 8)
// consecutive pairs; wave g takes chunks g, g + G, g + 2G, ... so every wave draws a similar mix of cheap and
// expensive pairs from all worlds, and a world with many pairs is spread over many CUs.  Contacts take slots
// from the world's counter ctr[1]; their order is made canonical by k_csort.
static constexpr uint32_t NARROW_LDS_OFFS = 4097;
// A wave's view of a world whose header it copies into its own LDS slot hs (faster than reading the header from
// global memory through the pair loop).  Called by a whole wave; the slot's earlier reads are complete before the
// copy overwrites it (a wave's LDS operations execute in order).
DEV WC wave_world(uint32_t* hs, const WorldDev& wd, uint32_t lane) {
  __builtin_amdgcn_wave_barrier();
  for (uint32_t k = lane; k < STEEL_WORLD_HEADER_WORDS; k += 64) hs[k] = wd.hdr[k];
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  return make_world<false>(hs, wd);
}
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(STEEL_WPE_COLLIDE))) k_narrow(const WorldDev* __restrict__ worlds, uint32_t nworlds, const uint32_t* __restrict__ offs) {
  __shared__ uint32_t s_offs[NARROW_LDS_OFFS];
  __shared__ ManifoldLds s_man[4];
  __shared__ uint32_t s_hdr[4][STEEL_WORLD_HEADER_WORDS];  // each wave's copy of its current world's header
  const bool lds_offs = nworlds + 1 <= NARROW_LDS_OFFS;
  if (lds_offs)
    for (uint32_t i = threadIdx.x; i <= nworlds; i += 256) s_offs[i] = offs[i];
  __syncthreads();
  const uint32_t* O = lds_offs ? s_offs : offs;
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t G = gridDim.x * 4, g = blockIdx.x * 4 + (threadIdx.x >> 6);
  const uint32_t total = O[nworlds];
  const uint32_t CH = max(1u, min(8u, total / G));
  const uint32_t nch = (total + CH - 1) / CH;
  for (uint32_t c = g; c < nch; c += G) {
    const uint32_t i0 = c * CH, i1 = min(i0 + CH, total);
    uint32_t lo = 0, hi = nworlds;  // largest w with O[w] <= i0 (then O[w + 1] > i0, so w < nworlds)
    while (lo < hi) {
      uint32_t mid = (lo + hi + 1) >> 1;
      if (O[mid] <= i0) lo = mid; else hi = mid - 1;
    }
    uint32_t w = lo, wend = O[w + 1];
    WorldDev wd = worlds[w];
    Collider col;
    col.w = wave_world(s_hdr[threadIdx.x >> 6], wd, lane);
    col.s_nc = wd.ctr + 1;
    col.mlds = s_man + (threadIdx.x >> 6);
    for (uint32_t i = i0; i < i1; i++) {
      if (i >= wend) {
        do { w++; wend = O[w + 1]; } while (i >= wend);
        wd = worlds[w];
        col.w = wave_world(s_hdr[threadIdx.x >> 6], wd, lane);
        col.s_nc = wd.ctr + 1;
      }
      col.narrow_pair(i - O[w], lane);
    }
  }
}

// Canonical contact order of one world by a team: (body pair, feature key) by the radix sort; equal pairs of
// those (a feature-hash collision) are ordered by (point bits, emission index) afterwards, as the reference's key
// does (wgsl.js:3906-3980).  tie: the team's tie flag (LDS for a block, global for a grid).
template <int BS, class TM>
DEV void csort_world(const TM& tm, const WorldDev wd, uint32_t* Hs, uint32_t* tie) {
  load_header(Hs, wd);
  WC w = make_world<false>(Hs, wd);
  const uint32_t tid = tm.tid, NT = tm.NT;
  if (tid == 0) *tie = 0;
  const uint32_t nc = min(wd.ctr[1], w.MAXC());
  // Key (body pair, feature): one 64-bit radix key while body ids + 1 fit 16 bits (the reference's packing);
  // past that, the feature pass then a stable pass over the 64-bit pair key.
  const bool wide = w.NB() > 0xffffu;
  auto feat_of = [&](uint32_t c) { const float4 m = wd.craw[c * 6 + 2]; return tu(m.z) | (tu(m.w) << 16); };
  auto pair_of = [&](uint32_t c) { const float4 m = wd.craw[c * 6 + 2]; return pair_key(ti(m.x), ti(m.y)); };
  for (uint32_t c = tid; c < nc; c += NT) {
    float4 m = wd.craw[c * 6 + 2];
    const uint32_t feat = tu(m.z) | (tu(m.w) << 16);
    wd.skey[c] = wide ? (uint64_t)feat : ((uint64_t)((uint32_t)(ti(m.x) + 1) | ((uint32_t)(ti(m.y) + 1) << 16)) << 32) | feat;
    wd.sval[c] = c;
  }
  tm.sync();
  tm.sort(wd.skey, wd.sval, wd.skey + w.MAXC(), wd.sval + w.MAXC(), nc);
  if (wide) {
    for (uint32_t i = tid; i < nc; i += NT) wd.skey[i] = pair_of(wd.sval[i]);
    tm.sync();
    tm.sort(wd.skey, wd.sval, wd.skey + w.MAXC(), wd.sval + w.MAXC(), nc);
  }
  // equal (pair, feature) of sorted positions i and j
  auto same = [&](uint32_t i, uint32_t j) {
    return wd.skey[i] == wd.skey[j] && (!wide || feat_of(wd.sval[i]) == feat_of(wd.sval[j]));
  };
  for (uint32_t i = tid; i + 1 < nc; i += NT) if (same(i, i + 1)) *tie = 1;
  tm.sync();
  if (*tie && tid == 0) {
    auto bits_of = [&](uint32_t c) {
      float4 p = wd.craw[c * 6];
      return hash_u32(hash_u32(hash_u32(fbits(p.x)) ^ fbits(p.y)) ^ fbits(p.z)) ^ fbits(p.w);
    };
    for (uint32_t i = 1; i < nc; i++) {
      uint64_t k = wd.skey[i];
      uint32_t v = wd.sval[i];
      uint32_t bv = bits_of(v), fv = wide ? feat_of(v) : 0u;
      int j = (int)i - 1;
      while (j >= 0 && wd.skey[j] == k && (!wide || feat_of(wd.sval[j]) == fv) &&
             (bits_of(wd.sval[j]) > bv || (bits_of(wd.sval[j]) == bv && wd.sval[j] > v))) {
        wd.skey[j + 1] = wd.skey[j]; wd.sval[j + 1] = wd.sval[j]; j--;
      }
      wd.skey[j + 1] = k; wd.sval[j + 1] = v;
    }
  }
  tm.sync();
  for (uint32_t i = tid; i < nc * 6; i += NT) {
    uint32_t r = i / 6, q = i % 6;
    wd.contacts[r * 6 + q] = wd.craw[wd.sval[r] * 6 + q];
  }
}

template <int BS>
__global__ void __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(STEEL_WPE_COLLIDE))) k_csort(const WorldDev* __restrict__ worlds, uint32_t nworlds) {
  __shared__ uint32_t Hs[STEEL_WORLD_HEADER_WORDS];
  __shared__ TeamLds<BS> tl;
  __shared__ SortLds<BS> sl;
  __shared__ uint32_t s_tie;
  const uint32_t wid = blockIdx.x;
  if (wid >= nworlds) return;
  const BlockTeam<BS> tm{threadIdx.x, &tl, &sl};
  csort_world<BS>(tm, worlds[wid], Hs, &s_tie);
}

// One large world over a cooperative launch (g: grid team scratch; g[3] is the tie flag).
template <int BS>
__global__ void __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(STEEL_WPE_COLLIDE))) k_csort_grid(const WorldDev* __restrict__ world, uint32_t* g) {
  __shared__ uint32_t Hs[STEEL_WORLD_HEADER_WORDS];
  __shared__ TeamLds<BS> tl;
  __shared__ SortLds<BS> sl;
  const GridTeam<BS> tm{gridDim.x * BS, blockIdx.x * BS + threadIdx.x, &tl, &sl, g};
  csort_world<BS>(tm, world[0], Hs, g + 3);
}

// ------------------------------------------------------------------------------------------------
// k_prepare (wgsl.js:1787-2494)
// ------------------------------------------------------------------------------------------------
struct Prep {
  W w;
  uint32_t parity;
  DEV float4 con(uint32_t c, uint32_t q) const { const gv4f v = ((const AS_GLOBAL gv4f*)w.d.contacts)[c * 6 + q]; return make_float4(v.x, v.y, v.z, v.w); }
  DEV uint64_t contact_pair(uint32_t c) const { float4 m = con(c, 2); return pair_key(ti(m.x), ti(m.y)); }
  DEV uint32_t unit_of(uint32_t p, uint32_t nu) const {
    const uint32_t* US = w.sch(w.S.USTART);
    uint32_t lo_ = 0, hi = nu;
    for (uint32_t it = 0; it < 32 && hi - lo_ > 1; it++) { uint32_t mid = (lo_ + hi) / 2; if (US[mid] <= p) lo_ = mid; else hi = mid; }
    return lo_;
  }
  struct BPC { uint32_t k0, k1, k2; f2 pos; float gap; uint32_t pk; };
  DEV BPC bp_contact(uint32_t c) const {  // wgsl.js:1475-1495
    const float BINS[20][3] = {
        {-0.57095543f, 0.75026579f, 0.33333337f}, {-0.64404845f, 0.17218052f, 0.74535593f}, {-0.86918069f, -0.36525859f, 0.33333330f},
        {-0.93522697f, -0.11932883f, -0.33333339f}, {-0.75091368f, 0.57010308f, -0.33333336f}, {0.36427171f, 0.86959472f, -0.33333337f},
        {-0.17291149f, 0.64385240f, -0.74535609f}, {0.47113684f, 0.47167205f, 0.74535599f}, {0.11826704f, 0.93536185f, 0.33333335f},
        {0.17291155f, -0.64385245f, 0.74535603f}, {0.0f, 0.0f, 1.0f}, {-0.11826707f, -0.93536184f, -0.33333335f},
        {-0.36427169f, -0.86959473f, 0.33333337f}, {0.0f, 0.0f, -1.0f}, {-0.47113677f, -0.47167196f, -0.74535609f},
        {0.93522695f, 0.11932898f, 0.33333339f}, {0.75091367f, -0.57010306f, 0.33333340f}, {0.57095545f, -0.75026579f, -0.33333334f},
        {0.64404840f, -0.17218051f, -0.74535597f}, {0.86918063f, 0.36525872f, -0.33333332f}};
    float4 c0 = con(c, 0), c1 = con(c, 1), cm = con(c, 2), c4 = con(c, 4);
    int a = ti(cm.x), b = ti(cm.y);
    uint32_t ga = tu(c4.z), gb = tu(c4.w);
    f3 ng = gb < ga ? -xyz(c1) : xyz(c1);
    int rb = a >= 0 ? a : b;
    f3 org = mk3(0, 0, 0);
    if (rb >= 0) org = w.pose_c(rb) - qrot(w.pose_q(rb), w.body_com_local(rb));
    uint32_t bin = 0;
    float bd = -2.0f;
    for (uint32_t i = 0; i < 20; i++) { float dd = dot(ng, mk3(BINS[i][0], BINS[i][1], BINS[i][2])); if (dd > bd) { bd = dd; bin = i; } }
    f3 fn = mk3(BINS[bin][0], BINS[bin][1], BINS[bin][2]);
    f3 rv = fabsf(fn.y) < 0.9f ? mk3(0, 1, 0) : mk3(1, 0, 0);
    f3 u = nrm(rv - dot(rv, fn) * fn);
    f3 v = cross(fn, u);
    f3 x = xyz(c0) - org;
    f2 pos = mk2(dot(x, u), dot(x, v));
    float cell = w.hf(W_F_BP_CELL);
    uint32_t cx = (uint32_t)((int32_t)clampf_(floorf(pos.x / cell + 0.5f), -31.0f, 31.0f) + 31);
    uint32_t cy = (uint32_t)((int32_t)clampf_(floorf(pos.y / cell + 0.5f), -31.0f, 31.0f) + 31);
    uint32_t qx = (uint32_t)((int32_t)clampf_(floorf(pos.x * 2000.0f), -32768.0f, 32767.0f) + 32768);
    uint32_t qy = (uint32_t)((int32_t)clampf_(floorf(pos.y * 2000.0f), -8192.0f, 8191.0f) + 8192);
    BPC r;
    r.k0 = min(ga, gb); r.k1 = max(ga, gb); r.k2 = bin | (cx << 5) | (cy << 11);
    r.pos = pos; r.gap = c0.w; r.pk = (qx << 15) | (qy << 1) | 1u;
    return r;
  }
  // bp_value with the six directions' cosines and sines taken from trig[2k], trig[2k + 1] (computed once per launch
  // as cos_c / sin_c of k * 60 degrees, the values bp_value computes).
  static DEV uint2 bp_value_t(f2 pos, float gap, uint32_t pk, uint32_t q, const float* trig) {
    if (q == 6) return make_uint2(fflip_u(fbits(-gap)), pk);
    if (q > 6 && gap > 0.0f) return make_uint2(0, 0);
    uint32_t k = q > 6 ? q - 7 : q;
    return make_uint2(fflip_u(fbits(dot(pos, mk2(trig[2 * k], trig[2 * k + 1])))), pk);
  }
  DEV uint2 bp_value(const BPC& g, uint32_t q) const {  // wgsl.js:1498-1503
    if (q == 6) return make_uint2(fflip_u(fbits(-g.gap)), g.pk);
    if (q > 6 && g.gap > 0.0f) return make_uint2(0, 0);
    uint32_t k = q > 6 ? q - 7 : q;
    float ang = (float)k * 1.0471975511965976f;
    return make_uint2(fflip_u(fbits(dot(g.pos, mk2(cos_c(ang), sin_c(ang))))), g.pk);
  }
  // wgsl.js:1233-1246; M: per-body color masks (NCOLW words per body)
  static DEV uint32_t first_free_color(uint32_t* M, uint32_t NCOLW, uint32_t NCOL, int a, int b) {
    for (uint32_t k = 0; k < NCOLW; k++) {
      uint32_t used = 0;
      if (a >= 0) used |= M[(uint32_t)a * NCOLW + k];
      if (b >= 0) used |= M[(uint32_t)b * NCOLW + k];
      if (used != 0xffffffffu) {
        uint32_t bit = (uint32_t)__builtin_ctz(~used);
        if (a >= 0) M[(uint32_t)a * NCOLW + k] |= 1u << bit;
        if (b >= 0) M[(uint32_t)b * NCOLW + k] |= 1u << bit;
        return k * 32 + bit;
      }
    }
    return NCOL;
  }
  DEV uint32_t anchor_hash(uint32_t pair, uint32_t k) const { return max(hash_u32(pair * 2u + k), 1u); }
  DEV float4 row(uint32_t r, uint32_t k) const { const gv4f v = ((const AS_GLOBAL gv4f*)w.d.rows)[r * w.ROWS() + k]; return make_float4(v.x, v.y, v.z, v.w); }
  DEV void set_row(uint32_t r, uint32_t k, float4 v) const { ((AS_GLOBAL gv4f*)w.d.rows)[r * w.ROWS() + k] = gv4f{v.x, v.y, v.z, v.w}; }
  // Contact warm start (wgsl.js:2175-2190, 2421-2431): the previous substep's impulses stored under `key`, scaled by
  // the warm scale; the tangent part projected on (t1, t2) within the friction cone mu * ln.
  DEV void warm_lookup(uint32_t key, float mu, f3 t1, f3 t2, float& ln, f2& lt) const {
    ln = 0.0f; lt = mk2(0.0f, 0.0f);
    const uint32_t WH = w.h(W_WH);
    if (WH == 0) return;
    const float ws = w.hf(W_WARM_SCALE);
    f3 lf = mk3(0, 0, 0);
    uint32_t rbase = parity * WH, h = key % WH;
    for (uint32_t probe = 0; probe < 32; probe++) {
      uint32_t k = w.d.warm[rbase + h];
      if (k == 0) break;
      if (k == key) {
        const uint4 v = ((const uint4*)(w.d.warm + 2 * WH))[rbase + h];
        ln = ubits(v.x) * ws;
        lf = mk3(ubits(v.y), ubits(v.z), ubits(v.w)) * ws;
        break;
      }
      h = (h + 1) % WH;
    }
    lt = mk2(dot(lf, t1), dot(lf, t2));
    float lim = mu * ln, lm = len(lt);
    if (lm > lim) lt = lt * (lim / mx(lm, 1e-12f));
  }
  // Contact torsion (wgsl.js:1526-1557): one spin row per articulated unit about the normal of its deepest point,
  // over the members touching in that plane (normal within 0.995, offset along it within 0.002, penetrating),
  // with the largest member torsion radius r.  Header in the unit's first row: (Ma lin, mu), (Ma ang, 1 / K),
  // (Mb lin, spin impulse), (Mb ang, r), (n, first row's member flag); every row's TO + 4 .w is its member flag.
  DEVN void build_torsion(uint32_t sp) const {
    const uint32_t MAXC = w.MAXC(), TO = w.h(W_TO);
    const uint32_t st = w.sch(w.S.USTART)[sp];
    const uint32_t n = min(w.sch(w.S.USIZE)[sp], MAXC - min(st, MAXC));
    if (n == 0) return;
    const uint64_t pair = w.sch64(w.S.UKEY)[sp];
    const int a = key_a(pair), b = key_b(pair);
    const bool art_u = (a >= 0 && w.body_kind(a) >= 2) || (b >= 0 && w.body_kind(b) >= 2);
    uint32_t seed = st;
    for (uint32_t i = st; i < st + n; i++) if (row(i, 1).w < row(seed, 1).w) seed = i;
    const f3 ns = xyz(row(seed, 0)), ps = xyz(row(seed, TO + 4));
    float r = 0.0f;
    for (uint32_t i = st; i < st + n; i++) {
      const bool member = art_u && row(i, 6).w <= 0.0f && dot(xyz(row(i, 0)), ns) >= 0.995f && fabsf(dot(xyz(row(i, TO + 4)) - ps, ns)) <= 0.002f;
      if (member) r = mx(r, row(i, TO + 3).w);
      float4 t4 = row(i, TO + 4);
      t4.w = member ? 1.0f : 0.0f;
      set_row(i, TO + 4, t4);
    }
    const float mu = row(seed, 2).w;
    if (r <= 0.0f || mu <= 0.0f) { set_row(st, TO, make_float4(0.0f, 0.0f, 0.0f, 0.0f)); return; }
    f3 Ma[2], Mb[2];
    w.body_response(a, mk3(0, 0, 0), -ns, Ma);
    w.body_response(b, mk3(0, 0, 0), ns, Mb);
    const float k = dot(-ns, Ma[1]) + dot(ns, Mb[1]);
    const float lead = row(st, TO + 4).w;
    set_row(st, TO, F4(Ma[0], mu));
    set_row(st, TO + 1, F4(Ma[1], k > 1e-12f ? 1.0f / (k + w.hf(W_F_CFM)) : 0.0f));
    set_row(st, TO + 2, F4(Mb[0], 0.0f));
    set_row(st, TO + 3, F4(Mb[1], r));
    set_row(st, TO + 4, F4(ns, lead));
  }
  DEVN void build_patch(uint32_t sp) const {  // wgsl.js:1562-1725
    const uint32_t* US = w.sch(w.S.USTART);
    uint32_t MAXC = w.MAXC(), PO = w.PO(), AH = w.AH();
    uint32_t st = US[sp];
    uint32_t n = min(w.sch(w.S.USIZE)[sp], MAXC - min(st, MAXC));
    if (n == 0) return;
    const uint64_t key = w.sch64(w.S.UKEY)[sp];
    int a = key_a(key), b = key_b(key);
    // Anchor-table pair id: the reference's (a + 1) | (b + 1) << 16, taken modulo 2^32 past 65535 bodies.
    const uint32_t pair = (uint32_t)(a + 1) | ((uint32_t)(b + 1) << 16);
    f3 ca = mk3(0, 0, 0), cb = mk3(0, 0, 0);
    q4 qa = mk4(0, 0, 0, 1);
    if (a >= 0) { ca = w.pose_c(a); qa = w.pose_q(a); }
    if (b >= 0) cb = w.pose_c(b);
    float r = mx(mn(w.body_radius(a), w.body_radius(b)), 1e-4f);
    uint32_t seed = st;
    for (uint32_t i = st; i < st + n; i++) if (row(i, 1).w < row(seed, 1).w) seed = i;
    f3 ns = xyz(row(seed, 0));
    uint32_t kept = w.sch(w.S.OSTART)[sp] - st;
    const uint32_t* KEPT = w.sch(w.S.KEPT);
    float4 ms = con(KEPT[kept + seed], 5);
    uint32_t first = seed;
    for (uint32_t i = st; i < st + n; i++) {
      f3 dseed = xyz(row(i, PO + 2)) - xyz(row(seed, PO + 2));
      float4 mi = con(KEPT[kept + i], 5);
      if (dot(xyz(row(i, 0)), ns) < 0.995f || mi.x != ms.x || mi.y != ms.y || fabsf(dot(dseed, ns)) > 0.02f * r || len(dseed) > 2.0f * r) continue;
      float4 rp = row(i, PO); rp.w = 0.0f; set_row(i, PO, rp);
      f3 p = xyz(row(i, PO + 2)), pf = xyz(row(first, PO + 2));
      if (p.x < pf.x || (p.x == pf.x && (p.y < pf.y || (p.y == pf.y && p.z < pf.z)))) first = i;
    }
    f3 pf = xyz(row(first, PO + 2));
    uint32_t second = first;
    float far = 1e-8f * r * r;
    for (uint32_t i = st; i < st + n; i++) {
      if (row(i, PO).w != 0.0f) continue;
      f3 dd = xyz(row(i, PO + 2)) - pf;
      if (dot(dd, dd) > far) { far = dot(dd, dd); second = i; }
    }
    uint32_t na = second != first ? 2 : 1;
    f3 loc0 = xyz(row(first, PO + 2)), loc1 = xyz(row(second, PO + 2));
    if (na == 2) {
      f3 t0 = cross(ns, mk3(1, 0, 0));
      if (dot(t0, t0) < 1e-6f) t0 = cross(ns, mk3(0, 1, 0));
      t0 = nrm(t0);
      f3 t1 = nrm(cross(ns, t0));
      float sx = 0, sy = 0, sxx = 0, sxy = 0, syy = 0, cnt = 0;
      f3 cen = mk3(0, 0, 0);
      for (uint32_t i = st; i < st + n; i++) {
        if (row(i, PO).w != 0.0f) continue;
        f3 pc = xyz(row(i, PO + 2));
        f3 dd = pc - pf;
        float x = dot(dd, t0), y = dot(dd, t1);
        sx += x; sy += y; sxx += x * x; sxy += x * y; syy += y * y; cen += pc; cnt += 1.0f;
      }
      float xx = sxx - sx * sx / cnt, xy = sxy - sx * sy / cnt, yy = syy - sy * sy / cnt;
      float spread = sqrtf((xx - yy) * (xx - yy) + 4.0f * xy * xy);
      if (spread > 1e-6f * (xx + yy)) {
        float big = 0.5f * (xx + yy + spread);
        f2 a2 = mk2(xy, big - xx);
        if (xx >= yy) a2 = mk2(big - yy, xy);
        a2 = nrm(a2);
        f3 ax = a2.x * t0 + a2.y * t1;
        float lo_ = 1e30f, hi = -1e30f;
        for (uint32_t i = st; i < st + n; i++) {
          if (row(i, PO).w != 0.0f) continue;
          float pr = dot(xyz(row(i, PO + 2)) - pf, ax);
          if (pr < lo_) { lo_ = pr; first = i; }
          if (pr > hi) { hi = pr; second = i; }
        }
        f3 centroid = cen / cnt;
        loc0 = centroid + dot(xyz(row(first, PO + 2)) - centroid, ax) * ax;
        loc1 = centroid + dot(xyz(row(second, PO + 2)) - centroid, ax) * ax;
        f3 across = nrm(cross(ns, ax));
        float tol = mx(1e-5f * r, 0.05f * (hi - lo_));
        f3 e0 = mk3(0, 0, 0), e1 = mk3(0, 0, 0);
        float c0 = 0, c1 = 0;
        for (uint32_t i = st; i < st + n; i++) {
          if (row(i, PO).w != 0.0f) continue;
          f3 pc = xyz(row(i, PO + 2));
          float pr = dot(pc - pf, ax);
          if (pr <= lo_ + tol) { e0 += pc; c0 += 1.0f; }
          if (pr >= hi - tol) { e1 += pc; c1 += 1.0f; }
        }
        e0 = e0 / mx(c0, 1.0f); e1 = e1 / mx(c1, 1.0f);
        loc0 += dot(e0 - centroid, across) * across + dot(e0 - centroid, ns) * ns;
        loc1 += dot(e1 - centroid, across) * across + dot(e1 - centroid, ns) * ns;
      }
    }
    float load = 0.0f;
    for (uint32_t i = st; i < st + n; i++) {
      if (row(i, PO).w != 0.0f) continue;
      load += row(i, 6).x;
      if (i != first && !(na == 2 && i == second)) { float4 l6 = row(i, 6); set_row(i, 6, make_float4(l6.x, 0.0f, 0.0f, l6.w)); }
    }
    const float ws = w.hf(W_WARM_SCALE);
    uint32_t rbase = parity * AH;
    for (uint32_t k = 0; k < na; k++) {
      uint32_t i = k == 1 ? second : first;
      f3 loc = k == 1 ? loc1 : loc0;
      f3 nn = xyz(row(i, 0)), t1 = xyz(row(i, 1)), t2 = xyz(row(i, 2));
      f3 fra = loc - ca, frb = loc - cb;
      f3 Ma[2][2], Mb[2][2];
      f3 dirs[2] = {t1, t2};
      for (uint32_t q = 0; q < 2; q++) {
        w.body_response(a, -dirs[q], -cross(fra, dirs[q]), Ma[q]);
        w.body_response(b, dirs[q], cross(frb, dirs[q]), Mb[q]);
        set_row(i, 11 + q * 4, F4(Ma[q][0], 0.0f)); set_row(i, 12 + q * 4, F4(Ma[q][1], 0.0f));
        set_row(i, 13 + q * 4, F4(Mb[q][0], 0.0f)); set_row(i, 14 + q * 4, F4(Mb[q][1], 0.0f));
      }
      float K[4];
      for (uint32_t x = 0; x < 2; x++)
        for (uint32_t y = 0; y < 2; y++) {
          f3 dd = dirs[x];
          K[x * 2 + y] = dot(-dd, Ma[y][0]) + dot(-cross(fra, dd), Ma[y][1]) + dot(dd, Mb[y][0]) + dot(cross(frb, dd), Mb[y][1]);
        }
      float ks = mx(K[0] + K[3], 1e-30f);
      float k00 = K[0] + 1e-6f * ks, k01 = 0.5f * (K[1] + K[2]), k11 = K[3] + 1e-6f * ks;
      float det = k00 * k11 - k01 * k01;
      set_row(i, 5, make_float4(k11 / det, -k01 / det, k00 / det, row(i, 5).w));
      set_row(i, 2, F4(t2, row(i, 2).w / (float)na));
      f3 e = mk3(0, 0, 0), imp = mk3(0, 0, 0);
      bool carried = false;
      float best = 4.0f * r * r;
      for (uint32_t kk = 0; kk < 2; kk++) {
        uint32_t hk = anchor_hash(pair, kk);
        uint32_t h = hk % AH;
        for (uint32_t probe = 0; probe < 32; probe++) {
          uint32_t key = w.d.akey[rbase + h];
          if (key == 0) break;
          const uint32_t* av = w.d.aval + (size_t)(rbase + h) * AVN;
          if (key == hk && av[0] == pair && av[1] == kk) {
            f3 ll = mk3(ubits(av[2]), ubits(av[3]), ubits(av[4]));
            f3 el = mk3(ubits(av[5]), ubits(av[6]), ubits(av[7]));
            f3 prev = ca + qrot(qa, ll);
            f3 dl = prev - loc;
            if (dot(dl, dl) <= best && fabsf(dot(dl, nn)) <= 0.05f * r) {
              f3 ew = qrot(qa, el);
              f3 on = qrot(qa, mk3(ubits(av[11]), ubits(av[12]), ubits(av[13])));
              float cs = dot(on, nn);
              if (cs > 0.0f) { f3 ax = cross(on, nn); f3 ce = cross(ax, ew); ew += ce + cross(ax, ce) / (1.0f + cs); }
              ew -= nn * dot(nn, ew);
              if (dot(ew, ew) <= 0.01f * r * r) { e = ew; best = dot(dl, dl); carried = true; imp = qrot(qa, mk3(ubits(av[8]), ubits(av[9]), ubits(av[10]))); }
            }
            break;
          }
          h = (h + 1) % AH;
        }
      }
      if (ws > 0.0f) {  // the carried anchor impulse, else the contact's, within mu / na of the members' load
        float4 l6 = row(i, 6);
        f2 lw = mk2(l6.y, l6.z);
        if (carried) lw = mk2(dot(imp, t1), dot(imp, t2)) * ws;
        float rad = mx(row(i, 2).w * load, 0.0f);
        if (len(lw) > rad) lw = lw * (rad / len(lw));
        set_row(i, 6, make_float4(l6.x, lw.x, lw.y, l6.w));
      }
      set_row(i, PO, F4(fra, 2.0f));
      set_row(i, PO + 1, F4(frb, (float)k));
      set_row(i, PO + 2, F4(e, 0.0f));
      set_row(i, PO + 3, F4(qinvrot(qa, fra), ubits(pair)));
    }
  }
  DEVN void build_row(uint32_t idx, uint32_t nu) const {  // wgsl.js:2106-2269
    uint32_t sp = unit_of(idx, nu);
    uint32_t c = w.sch(w.S.KEPT)[w.sch(w.S.OSTART)[sp] + idx - w.sch(w.S.USTART)[sp]];
    float4 c0 = con(c, 0), c1 = con(c, 1), cmeta = con(c, 2);
    int a = ti(cmeta.x), b = ti(cmeta.y);
    uint32_t key = max(tu(cmeta.z) | (tu(cmeta.w) << 16), 1u);
    f3 p = xyz(c0);
    float phi = c0.w;
    f3 n = xyz(c1);
    float mu = phi <= w.hf(W_F_FGAP) ? c1.w * w.hf(W_F_FSCALE) : 0.0f;
    f3 t1 = cross(n, mk3(1, 0, 0));
    if (dot(t1, t1) < 0.1f) t1 = cross(n, mk3(0, 1, 0));
    t1 = nrm(t1);
    f3 t2 = cross(n, t1);
    f3 ca = mk3(0, 0, 0), cb = mk3(0, 0, 0);
    if (a >= 0) ca = w.pose_c(a);
    if (b >= 0) cb = w.pose_c(b);
    float half = 0.5f * phi;
    f3 ra = p - n * half - ca, rb = p + n * half - cb;
    f3 fa = ra, fb = rb;
    f3 dirs[3] = {n, t1, t2};
    f3 Ma[3][2], Mb[3][2];
    for (uint32_t r = 0; r < 3; r++) {
      f3 dd = dirs[r];
      f3 aa = r == 0 ? ra : fa, ab = r == 0 ? rb : fb;
      w.body_response(a, -dd, -cross(aa, dd), Ma[r]);
      w.body_response(b, dd, cross(ab, dd), Mb[r]);
      set_row(idx, 7 + r * 4, F4(Ma[r][0], 0.0f)); set_row(idx, 8 + r * 4, F4(Ma[r][1], 0.0f));
      set_row(idx, 9 + r * 4, F4(Mb[r][0], 0.0f)); set_row(idx, 10 + r * 4, F4(Mb[r][1], 0.0f));
    }
    float K[9];
    for (uint32_t r = 0; r < 3; r++)
      for (uint32_t s = 0; s < 3; s++) {
        f3 dd = dirs[r];
        f3 aa = r == 0 ? ra : fa, ab = r == 0 ? rb : fb;
        K[r * 3 + s] = dot(-dd, Ma[s][0]) + dot(-cross(aa, dd), Ma[s][1]) + dot(dd, Mb[s][0]) + dot(cross(ab, dd), Mb[s][1]);
      }
    uint32_t PO = w.PO();
    if (w.h(W_EXTRA)) { set_row(idx, PO, F4(fa, 1.0f)); set_row(idx, PO + 1, F4(fb, 0.0f)); set_row(idx, PO + 2, F4(p, 0.0f)); }
    if (w.h(W_TORS)) { set_row(idx, w.h(W_TO) + 3, make_float4(0.0f, 0.0f, 0.0f, con(c, 3).w)); set_row(idx, w.h(W_TO) + 4, F4(p, 0.0f)); }
    if (w.h(W_FMODE)) {
      set_row(idx, w.h(W_FMO), make_float4(K[0], 0.5f * (K[1] + K[3]), 0.5f * (K[2] + K[6]), 0.0f));
      set_row(idx, w.h(W_FMO) + 1, make_float4(K[4], 0.5f * (K[5] + K[7]), K[8], 0.0f));
    }
    float cfm = w.hf(W_F_CFM), dt = w.hf(W_DT);
    float kscale = mx(K[0] + K[4] + K[8], 1e-30f);
    float effn = 0.0f;
    if (K[0] > 1e-6f * kscale && K[0] > 1e-12f) effn = 1.0f / (K[0] + cfm);
    float k00 = K[4] + 1e-6f * kscale + cfm, k01 = 0.5f * (K[5] + K[7]), k11 = K[8] + 1e-6f * kscale + cfm;
    float det = k00 * k11 - k01 * k01;
    float rhs;
    if (phi < 0.0f) rhs = mx(w.hf(W_BETA) * phi / dt, -w.depen_cap(a, b)); else rhs = w.hf(W_F_SPEC) * phi / dt;
    float e = con(c, 3).x;
    bool bounce = false;
    if (e > 0.0f) {
      f3 ta0[2], tb0[2];
      pre_twist(a, ta0); pre_twist(b, tb0);
      float vn0 = dot(n, (tb0[0] + cross(tb0[1], rb)) - (ta0[0] + cross(ta0[1], ra)));
      bounce = restitution_fires(e, phi, vn0);
      if (bounce) rhs = e * vn0;
    }
    float effc = effn, rhsc = rhs;
    if (w.h(W_COMPL)) {
      // contact_compliance: a spring-damper normal; the pair's ke and kd split over its kept points (wgsl.js:2194-2207)
      const float4 kc4 = con(c, 4);
      const float nk = (float)max(w.sch(w.S.USIZE)[sp], 1u);
      const float kx = kc4.x / nk, ky = kc4.y / nk;
      float gam = 0.0f;
      if (kx > 0.0f) {
        const float ce = phi <= 0.0f ? ky : 0.0f;
        const float den = dt * kx + ce;
        gam = 1.0f / (dt * den);
        rhsc = kx * phi / den;
        if (effn > 0.0f) effc = 1.0f / (1.0f / effn + gam);
      }
      set_row(idx, w.h(W_CO), make_float4(gam, 0.0f, 0.0f, 0.0f));
    }
    set_row(idx, 0, F4(n, effc));
    set_row(idx, 1, F4(t1, rhsc));
    set_row(idx, 2, F4(t2, mu));
    set_row(idx, 3, F4(ra, (float)(key & 0xffffu)));
    set_row(idx, 4, F4(rb, (float)(key >> 16)));
    float wrow = row_weight(phi);
    if (wrow < 1.0f && ((a >= 0 && w.body_kind(a) == 2) || (b >= 0 && w.body_kind(b) == 2))) {
      f3 ua[2], ub[2];
      w.body_response_unsplit(a, -n, -cross(ra, n), ua);
      w.body_response_unsplit(b, n, cross(rb, n), ub);
      float du = dot(-n, ua[0]) + dot(-cross(ra, n), ua[1]) + dot(n, ub[0]) + dot(cross(rb, n), ub[1]);
      if (du > 0.0f && K[0] != du) wrow = wrow * K[0] / (wrow * K[0] + (1.0f - wrow) * du);
    }
    set_row(idx, 5, make_float4(k11 / det, -k01 / det, k00 / det, bounce ? 1.0f : wrow));
    float ln;
    f2 lt;
    warm_lookup(key, mu, t1, t2, ln, lt);
    set_row(idx, 6, make_float4(ln, lt.x, lt.y, bounce ? 1e30f : phi));
    if (w.h(W_ANG)) {
      // Spin (about n) and rolling (about t1, t2) rows: angular-only Jacobians (wgsl.js:2223-2243).
      const float4 mat = con(c, 3);
      f3 MA[3][2], MB[3][2];
      for (uint32_t r = 0; r < 3; r++) {
        const f3 dd = dirs[r];
        w.body_response(a, mk3(0, 0, 0), -dd, MA[r]);
        w.body_response(b, mk3(0, 0, 0), dd, MB[r]);
        set_row(idx, 22 + r * 4, F4(MA[r][0], 0.0f)); set_row(idx, 23 + r * 4, F4(MA[r][1], 0.0f));
        set_row(idx, 24 + r * 4, F4(MB[r][0], 0.0f)); set_row(idx, 25 + r * 4, F4(MB[r][1], 0.0f));
      }
      float KA[9];
      for (uint32_t r = 0; r < 3; r++)
        for (uint32_t s3 = 0; s3 < 3; s3++) KA[r * 3 + s3] = dot(-dirs[r], MA[s3][1]) + dot(dirs[r], MB[s3][1]);
      const float ks = mx(KA[0], 1e-12f);
      const float r00 = KA[4] + 1e-9f, r01 = 0.5f * (KA[5] + KA[7]), r11 = KA[8] + 1e-9f;
      const float rd = r00 * r11 - r01 * r01;
      set_row(idx, 19, make_float4(mat.y, mat.z, KA[0] > 1e-12f ? 1.0f / ks : 0.0f, 0.0f));
      set_row(idx, 20, make_float4(r11 / rd, -r01 / rd, r00 / rd, 0.0f));
      set_row(idx, 21, make_float4(0.0f, 0.0f, 0.0f, 0.0f));
      if (w.h(W_ANG_ART)) {
        // Raw 5x5 response H (t1, t2, spin, roll1, roll2), upper triangle, for the elliptic block (wgsl.js:2244-2267).
        f3 JA[5][2], JB[5][2], RA[5][2], RB[5][2];
        for (uint32_t q = 0; q < 2; q++) {
          const f3 dd = dirs[q + 1];
          JA[q][0] = -dd; JA[q][1] = -cross(fa, dd); JB[q][0] = dd; JB[q][1] = cross(fb, dd);
          RA[q][0] = Ma[q + 1][0]; RA[q][1] = Ma[q + 1][1]; RB[q][0] = Mb[q + 1][0]; RB[q][1] = Mb[q + 1][1];
        }
        for (uint32_t q = 0; q < 3; q++) {
          JA[q + 2][0] = mk3(0, 0, 0); JA[q + 2][1] = -dirs[q]; JB[q + 2][0] = mk3(0, 0, 0); JB[q + 2][1] = dirs[q];
          RA[q + 2][0] = MA[q][0]; RA[q + 2][1] = MA[q][1]; RB[q + 2][0] = MB[q][0]; RB[q + 2][1] = MB[q][1];
        }
        float Hu[16];
        uint32_t hi = 0;
        for (uint32_t x = 0; x < 5; x++)
          for (uint32_t y = x; y < 5; y++)
            Hu[hi++] = dot(JA[x][0], RA[y][0]) + dot(JA[x][1], RA[y][1]) + dot(JB[x][0], RB[y][0]) + dot(JB[x][1], RB[y][1]);
        Hu[15] = 0.0f;
        for (uint32_t q = 0; q < 4; q++) set_row(idx, w.h(W_EO) + q, make_float4(Hu[q * 4], Hu[q * 4 + 1], Hu[q * 4 + 2], Hu[q * 4 + 3]));
      }
    }
  }
  DEV float row_weight(float phi) const { float reg = w.hf(W_REG); if (phi > 0.0f || reg <= 0.0f) return 1.0f; return 1.0f / (1.0f + reg); }
  DEV bool restitution_fires(float e, float phi, float vn) const {
    if (w.h(W_REST) == 0 || e <= 0.0f || vn >= -w.hf(W_F_REST_SPEED)) return false;
    return phi <= 1e-6f || phi + w.hf(W_DT) * vn <= 1e-6f;
  }
  DEV void pre_twist(int b, f3 out[2]) const {
    out[0] = mk3(0, 0, 0); out[1] = mk3(0, 0, 0);
    if (b < 0) return;
    if (w.body_kind(b) == 1) { out[0] = xyz(w.d.vel[(uint32_t)b * 2]); out[1] = xyz(w.d.vel[(uint32_t)b * 2 + 1]); return; }
    if (w.body_kind(b) == 2) { uint32_t o = w.lo((uint32_t)w.body_link(b)) + L_V; out[0] = w.ld3(o + 3); out[1] = w.ld3(o); }
  }
  // Dense row of contact c, three lanes (base + r, r = 0..2): lane r solves row r's Y = H^-1 J_r^T, the
  // 3x3 K = J Y^T is gathered by shuffles, lane 0 writes the row header (wgsl.js:2370-2444).
  // Mimic row i in dense slot `slot`: qd_follower - c1 qd_leader, Baumgarte on q_follower - c1 q_leader - c0;
  // a bilateral block (J, Y = H^-1 J^T, 1 / (J Y), rhs) as in wgsl.js:1351-1369.
  // Bilateral dense block at art offset o: nb unclamped rows of J (3 NDA floats), Y = H^-1 J^T, 1 / (J Y) and the
  // rhs (wgsl.js:1351-1369).
  DEVN void put_bilateral(uint32_t o, uint32_t art_id, const float* J, uint32_t nb, f3 rhs) const {
    const uint4 ai = w.art_info(art_id);
    const uint32_t NDA = w.NDA(), DHDR = w.h(W_DHDR);
    AS_GLOBAL float* art_ = (AS_GLOBAL float*)w.d.art;
    for (uint32_t r = 0; r < nb; r++) {
      float tau[64], y[64];
      SVd ps[64], dvs[64];
      float uu[6 * 64];
      for (uint32_t k = 0; k < NDA; k++) { tau[k] = k < ai.w ? J[r * NDA + k] : 0.0f; y[k] = 0.0f; }
      w.art_solve_gen(art_id, tau, y, ps, dvs, uu);
      float kk = 0.0f;
      for (uint32_t k = 0; k < ai.w; k++) {
        art_[o + DHDR + r * NDA + k] = J[r * NDA + k];
        art_[o + DHDR + 3 * NDA + r * NDA + k] = y[k];
        kk += J[r * NDA + k] * y[k];
      }
      art_[o + 5 + r] = kk > 1e-12f ? 1.0f / kk : 0.0f;
    }
    art_[o] = (float)art_id; art_[o + 1] = 0.0f; art_[o + 2] = 0.0f;
    art_[o + 4] = rhs.x; art_[o + 22] = rhs.y; art_[o + 23] = rhs.z;
    art_[o + 9] = 0.0f; art_[o + 10] = 0.0f; art_[o + 11] = 0.0f;
    art_[o + 21] = (float)nb;
  }
  // Loop closure i in dense slot `slot` (and slot + 1 when revolute), wgsl.js:2445-2480: three point rows at the
  // anchors' midpoint, plus two hinge-alignment rows for a revolute closure; a world parent contributes no dofs.
  DEVN void loop_row(uint32_t i, uint32_t slot) const {
    const uint32_t SL = w.h(W_SF_LOOP) + i * 5;
    const float4 L0 = w.sf(SL);
    const int pl = ti(L0.x);
    const uint32_t cl = tu(L0.y), art_id = tu(L0.z);
    const bool rev = L0.w > 0.5f;
    f3 pa = xyz(w.sf(SL + 1));
    q4 rp = mk4(0, 0, 0, 1);
    if (pl >= 0) { rp = w.link_rot((uint32_t)pl); pa = w.ld3(w.lo((uint32_t)pl) + L_POS) + qrot(rp, pa); }
    const q4 rc = w.link_rot(cl);
    const f3 pc = w.ld3(w.lo(cl) + L_POS) + qrot(rc, xyz(w.sf(SL + 2)));
    const f3 pm = 0.5f * (pa + pc);
    const uint4 ai = w.art_info(art_id);
    const uint32_t NDA = w.NDA(), A_DENSE = w.h(W_A_DENSE), DPS = w.h(W_DPS);
    const float beta = w.hf(W_BETA), dt = w.hf(W_DT);
    const f3 ex = mk3(1, 0, 0), ey = mk3(0, 1, 0), ez = mk3(0, 0, 1);
    float J[3 * 64];
    for (uint32_t k = 0; k < 3 * NDA; k++) J[k] = 0.0f;
    w.add_gen_jac(cl, pm, 1.0f, ex, ey, ez, ai.z, J);
    if (pl >= 0) w.add_gen_jac((uint32_t)pl, pm, -1.0f, ex, ey, ez, ai.z, J);
    put_bilateral(A_DENSE + slot * DPS, art_id, J, 3, beta * (pc - pa) / dt);
    if (!rev) return;
    const f3 ap = qrot(rp, xyz(w.sf(SL + 3))), ac = qrot(rc, xyz(w.sf(SL + 4)));
    f3 e1 = cross(ap, ex);
    if (dot(e1, e1) < 1e-6f) e1 = cross(ap, ey);
    e1 = nrm(e1);
    const f3 e2 = cross(ap, e1);
    for (uint32_t k = 0; k < 3 * NDA; k++) J[k] = 0.0f;
    w.add_gen_ang(cl, 1.0f, e1, e2, ai.z, J);
    if (pl >= 0) w.add_gen_ang((uint32_t)pl, -1.0f, e1, e2, ai.z, J);
    const f3 err = cross(ap, ac);
    put_bilateral(A_DENSE + (slot + 1) * DPS, art_id, J, 2, beta * mk3(dot(e1, err), dot(e2, err), 0.0f) / dt);
  }
  DEVN void mimic_row(uint32_t i, uint32_t slot) const {
    const float4 M0 = w.sf(w.h(W_SF_MIMIC) + i * 2), M1 = w.sf(w.h(W_SF_MIMIC) + i * 2 + 1);
    const uint32_t art_id = tu(M0.x);
    const uint4 ai = w.art_info(art_id);
    const uint32_t NDA = w.NDA(), DHDR = w.h(W_DHDR);
    AS_GLOBAL float* art_ = (AS_GLOBAL float*)w.d.art;
    const uint32_t o = w.h(W_A_DENSE) + slot * w.h(W_DPS);
    float tau[64], y[64];
    SVd ps[64], dvs[64];
    float uu[6 * 64];
    for (uint32_t k = 0; k < NDA; k++) { tau[k] = 0.0f; y[k] = 0.0f; }
    if (tu(M0.y) < ai.w) tau[tu(M0.y)] = 1.0f;
    if (tu(M0.z) < ai.w) tau[tu(M0.z)] -= M0.w;
    w.art_solve_gen(art_id, tau, y, ps, dvs, uu);
    float kk = 0.0f;
    for (uint32_t k = 0; k < ai.w; k++) {
      art_[o + DHDR + k] = tau[k];
      art_[o + DHDR + 3 * NDA + k] = y[k];
      kk += tau[k] * y[k];
    }
    art_[o + 5] = kk > 1e-12f ? 1.0f / kk : 0.0f;
    const float phi = art_[w.h(W_A_Q) + tu(M1.x)] - M0.w * art_[w.h(W_A_Q) + tu(M1.y)] - M1.z;
    art_[o] = (float)art_id; art_[o + 1] = 0.0f; art_[o + 2] = 0.0f;
    art_[o + 4] = w.hf(W_BETA) * phi / w.hf(W_DT); art_[o + 22] = 0.0f; art_[o + 23] = 0.0f;
    art_[o + 9] = 0.0f; art_[o + 10] = 0.0f; art_[o + 11] = 0.0f;
    art_[o + 21] = 1.0f;
  }
  DEVN void dense_row(uint32_t c, uint32_t slot, uint32_t r, uint32_t base) const {
    float4 c0 = con(c, 0), c1 = con(c, 1), cmeta = con(c, 2);
    int a = ti(cmeta.x), b = ti(cmeta.y);
    uint32_t key = max(tu(cmeta.z) | (tu(cmeta.w) << 16), 1u);
    f3 p = xyz(c0);
    float phi = c0.w;
    f3 n = xyz(c1);
    float mu = phi <= w.hf(W_F_FGAP) ? c1.w * w.hf(W_F_FSCALE) : 0.0f;
    f3 t1 = cross(n, mk3(1, 0, 0));
    if (dot(t1, t1) < 0.1f) t1 = cross(n, mk3(0, 1, 0));
    t1 = nrm(t1);
    f3 t2 = cross(n, t1);
    uint32_t art_id = (uint32_t)w.body_art(a);
    uint4 ai = w.art_info(art_id);
    uint32_t NDA = w.NDA(), DHDR = w.h(W_DHDR);
    float J[3 * 64];
    for (uint32_t k = 0; k < 3 * NDA; k++) J[k] = 0.0f;
    w.add_gen_jac((uint32_t)w.body_link(b), p, 1.0f, n, t1, t2, ai.z, J);
    w.add_gen_jac((uint32_t)w.body_link(a), p, -1.0f, n, t1, t2, ai.z, J);
    AS_GLOBAL float* art_ = (AS_GLOBAL float*)w.d.art;
    uint32_t o = w.h(W_A_DENSE) + slot * w.h(W_DPS);
    SVd ps[64], dvs[64];
    float uu[6 * 64];
    float kc[3];
    {
      // Row r of J is the impulse (its entries past the articulation's dofs are zero and never read);
      // the solve writes every dof of y that is read.
      float y[64];
      w.art_solve_gen(art_id, J + r * NDA, y, ps, dvs, uu);
      for (uint32_t k = 0; k < ai.w; k++) { art_[o + DHDR + r * NDA + k] = J[r * NDA + k]; art_[o + DHDR + 3 * NDA + r * NDA + k] = y[k]; }
      for (uint32_t s2 = 0; s2 < 3; s2++) { float acc = 0.0f; for (uint32_t k = 0; k < ai.w; k++) acc += J[s2 * NDA + k] * y[k]; kc[s2] = acc; }
    }
    float K[9];  // K[s2 * 3 + rr] from lane base + rr
    for (uint32_t s2 = 0; s2 < 3; s2++)
      for (uint32_t rr = 0; rr < 3; rr++) K[s2 * 3 + rr] = __shfl(kc[s2], (int)(base + rr), 64);
    if (r != 0) return;
    float cfm = w.hf(W_F_CFM), dt = w.hf(W_DT);
    float kscale = mx(K[0] + K[4] + K[8], 1e-30f);
    float effn = 0.0f;
    if (K[0] > 1e-6f * kscale && K[0] > 1e-12f) effn = 1.0f / (K[0] + cfm);
    float k00 = K[4] + 1e-6f * kscale + cfm, k01 = 0.5f * (K[5] + K[7]), k11 = K[8] + 1e-6f * kscale + cfm;
    float det = k00 * k11 - k01 * k01;
    float rhs;
    if (phi < 0.0f) rhs = mx(w.hf(W_BETA) * phi / dt, -w.depen_cap(a, b)); else rhs = w.hf(W_F_SPEC) * phi / dt;
    float e = con(c, 3).x;
    bool bounce = false;
    if (e > 0.0f) {
      float vn0 = 0.0f;
      for (uint32_t k = 0; k < ai.w; k++) vn0 += J[k] * art_[w.h(W_A_QD) + ai.z + k];
      bounce = restitution_fires(e, phi, vn0);
      if (bounce) rhs = e * vn0;
    }
    float ln;
    f2 lt;
    warm_lookup(key, mu, t1, t2, ln, lt);
    art_[o] = (float)art_id; art_[o + 1] = (float)(key & 0xffffu); art_[o + 2] = (float)(key >> 16); art_[o + 3] = mu;
    art_[o + 4] = rhs; art_[o + 5] = effn;
    art_[o + 6] = k11 / det; art_[o + 7] = -k01 / det; art_[o + 8] = k00 / det;
    art_[o + 9] = ln; art_[o + 10] = lt.x; art_[o + 11] = lt.y;
    w.st3(o + 12, t1); w.st3(o + 15, t2); w.st3(o + 18, n);
    art_[o + 21] = 0.0f;
    art_[o + 24] = bounce ? 1.0f : row_weight(phi);
    art_[o + 22] = bounce ? 1e30f : phi;
  }
};

// LDS of prepare's coloring phase (words).
static constexpr uint32_t PREP_CH = 256, PREP_WC = 258;
// Walk entries each thread of the first workgroup holds per window when the color masks are in global memory.
static constexpr uint32_t PREP_CPER = 4;
// Walk entries up to which one workgroup whose color masks are in global memory colors them serially on one lane;
// past it the workgroup colors windows of entries in parallel rounds (the same colors: an entry is colored once no
// earlier pending entry shares one of its bodies).
static constexpr uint32_t PREP_SERIAL_COLOR = 256;

// LDS of prepare's phases, per workgroup.
template <int BS>
struct PrepLds {
  uint32_t wr[PREP_CH];              // coloring: row of each walk entry of a chunk
  uint64_t wk[PREP_CH];              // coloring: body-pair key of each walk entry of a chunk
  uint32_t wc[PREP_WC];              // coloring: per-color counts and offsets when NCOL + 2 fits
  uint32_t chg;                      // sleep islands: a contact joined two islands this pass
  uint32_t tree[(BS / 64) * 64];     // factor: link tree of each wave's articulation
  alignas(16) float pab[(BS / 64) * 512];  // factor: children's bias forces; then the link mirror (PrepMirror)
  float trig[12];                    // reduction: cos, sin of k * 60 degrees, k = 0..5
};

// Prepare of one world by a team.  cap: words of dynamic LDS (s_pdyn), which hold a block team's per-body color
// masks and first-pending slots when NB * (NCOLW + 1) fits (steel_prep_lds_words); a grid team keeps them in
// global memory.
// PART: 0 the whole prepare; 1 up to the forward kinematics and factor, 2 from there on (a large world's FK and
// factor run between them in k_art_factor, a workgroup per articulation over the whole device).
template <int BS, class TM, int PART = 0>
DEV void prepare_world(const TM& tm, const WorldDev wd, uint32_t* Hs, uint32_t cap, uint32_t* s_pdyn, PrepLds<BS>* pl) {
  uint32_t* s_wr = pl->wr;
  uint64_t* s_wk = pl->wk;
  uint32_t* s_wc = pl->wc;
  uint32_t* s_tree = pl->tree;
  float* s_pab = pl->pab;
  float* s_trig = pl->trig;
  if (threadIdx.x < 6) {
    const float ang = (float)threadIdx.x * 1.0471975511965976f;
    s_trig[2 * threadIdx.x] = cos_c(ang);
    s_trig[2 * threadIdx.x + 1] = sin_c(ang);
  }
  load_header(Hs, wd);
  Prep P;
  P.w = make_world(Hs, wd);
  const W& w = P.w;
  const SchedLayout& S = w.S;
  const uint32_t tid = tm.tid, NT = tm.NT;
  P.parity = wd.ctr[2] & 1u;
  const uint32_t NB = w.NB(), NA = w.NA(), NL = w.NL(), ND = w.ND(), MAXC = w.MAXC(), MAXU = w.MAXU(), UH = w.UH(), AH = w.AH();
  const uint32_t NCOL = w.NCOL(), NCOLW = w.h(W_NCOLW), PATCH = w.h(W_PATCH);
  uint32_t* sched = wd.sched;
  AS_GLOBAL float* art_ = (AS_GLOBAL float*)wd.art;
  if constexpr (PART != 2) {
  // ---- Phase 0
  for (uint32_t i = tid; i < UH; i += NT) { w.sch64(S.HKEY)[i] = 0; sched[S.HCNT + i] = 0; }
  if (PATCH) for (uint32_t i = tid; i < AH; i += NT) wd.akey[(1 - P.parity) * AH + i] = 0;
  if (const uint32_t WH = w.h(W_WH)) {  // the table this substep's solve writes, and its winners
    for (uint32_t i = tid; i < WH; i += NT) wd.warm[(1 - P.parity) * WH + i] = 0;
    for (uint32_t i = tid; i < 2 * WH; i += NT) sched[S.WWIN + i] = 0;
  }
  for (uint32_t i = tid; i < max(NB, 1u); i += NT) sched[S.LINKHIT + i] = 0;
  for (uint32_t i = tid; i < max(NB, 1u) * NCOLW; i += NT) sched[S.WMASK + i] = 0;
  for (uint32_t i = tid; i < NCOL + 2; i += NT) sched[S.WCOUNT + i] = 0;
  for (uint32_t i = tid; i < max(NL, 1u); i += NT) sched[S.JSPLIT + i] = 0;
  if (tid == 0) sched[S.META + 3] = 0;
  const float dt = w.hf(W_DT);
  if (w.h(W_SLEEP)) {
    // Sleep islands (wgsl.js:1370-1457): nodes joined by contacts, each island rooted at its lowest node (roots
    // hook under the smaller root with atomicMin and paths are compressed until no contact joins two islands,
    // the serial union-find's result); island flags on the root: 1 supported, 2 vetoed (moving or driven), 4 a
    // member still too young.  An island sleeps once all its members have been quiet for the quiet time.
    const uint32_t NN = NB + NA, nc0 = min(wd.ctr[1], MAXC);
    uint32_t* UF = sched + S.UF;
    uint32_t* UFF = sched + S.UFF;
    const float LIN = w.hf(W_F_SLEEP_LIN), ANG = w.hf(W_F_SLEEP_ANG), QUIET = w.hf(W_F_SLEEP_TIME);
    const uint32_t ASLP = w.h(W_A_SLP);
    auto find = [&](uint32_t n) {
      for (uint32_t k = 0; k < NN; k++) {
        uint32_t p = __atomic_load_n(&UF[n], __ATOMIC_RELAXED);
        if (p == n) break;
        n = p;
      }
      return n;
    };
    for (uint32_t n = tid; n < NN; n += NT) { UF[n] = n; UFF[n] = 0; }
    if (tid == 0) sched[S.META + 5] = 0;
    tm.sync();
    for (uint32_t c = tid; c < nc0; c += NT) {
      float4 m = P.con(c, 2);
      int na = w.body_node(ti(m.x)), nb = w.body_node(ti(m.y));
      if ((na < 0 || nb < 0) && max(na, nb) >= 0 && P.con(c, 0).w <= 1e-5f) atomicOr(&UFF[(uint32_t)max(na, nb)], 1u);
    }
    for (;;) {
      if (tid == 0) pl->chg = 0;
      tm.sync();
      for (uint32_t c = tid; c < nc0; c += NT) {
        float4 m = P.con(c, 2);
        int na = w.body_node(ti(m.x)), nb = w.body_node(ti(m.y));
        if (na < 0 || nb < 0) continue;
        uint32_t ra = find((uint32_t)na), rb = find((uint32_t)nb);
        if (ra != rb) { atomicMin(&UF[max(ra, rb)], min(ra, rb)); pl->chg = 1; }
      }
      tm.sync();
      for (uint32_t n = tid; n < NN; n += NT) atomicMin(&UF[n], find(n));
      tm.sync();
      if (pl->chg == 0) break;
      tm.sync();
    }
    for (uint32_t n = tid; n < NN; n += NT) {
      bool moving = false, fixed = false, veto = false;
      if (n < NB) {
        if (w.body_kind((int)n) != 1) continue;
        if (art_[ASLP + n * 4 + 1] < 0.5f) moving = len(xyz(wd.vel[n * 2])) > LIN || len(xyz(wd.vel[n * 2 + 1])) > ANG;
      } else {
        uint4 ai = w.art_info(n - NB);
        fixed = w.link_type(ai.x) != JT_FREE;
        for (uint32_t k = 0; k < ai.y; k++) {
          SVd v = w.ldsv(w.lo(ai.x + k) + L_V);
          if (art_[ASLP + n * 4 + 1] < 0.5f && (len(v.l) > LIN || len(v.a) > ANG)) moving = true;
        }
        for (uint32_t dd = ai.z; dd < ai.z + ai.w; dd++) { float4 dr = w.dof_drive(dd); if (dr.y > 0.0f || dr.z > 0.0f) veto = true; }
        if (w.su(w.h(W_SU_ACON) + n - NB) != 0) veto = true;  // owns a mimic
      }
      uint32_t f = (fixed ? 1u : 0u) | ((moving || veto) ? 2u : 0u) | (UFF[n] & 1u);
      atomicOr(&UFF[UF[n]], f);
    }
    tm.sync();
    for (uint32_t n = tid; n < NN; n += NT) {
      if (n < NB && w.body_kind((int)n) != 1) continue;
      uint32_t r = UF[n], f = __atomic_load_n(&UFF[r], __ATOMIC_RELAXED);
      float age = art_[ASLP + n * 4];
      if ((f & 2u) != 0 || (f & 1u) == 0) age = 0.0f; else age = mn(age + dt, QUIET);
      art_[ASLP + n * 4] = age;
      if (age < QUIET) atomicOr(&UFF[r], 4u);
    }
    tm.sync();
    for (uint32_t n = tid; n < NN; n += NT) {
      if (n < NB && w.body_kind((int)n) != 1) continue;
      bool s = (UFF[UF[n]] & 4u) == 0;
      art_[ASLP + n * 4 + 1] = s ? 1.0f : 0.0f;
      if (s) atomicAdd(&sched[S.META + 5], 1u);
    }
    tm.sync();
  }
  // Recorded target paths (60 Hz samples), interpolated at every substep at its midpoint half a frame back
  // (or held per frame), plus the integral term of the first dofs (wgsl.js:1800-1820).  Off once the user
  // set targets.
  {
    const uint32_t driven = (wd.ctr[3] & 1u) ? 0u : w.h(W_DRIVEN), frames = w.h(W_PATH_FRAMES);
    if (driven && frames) {
      const uint32_t sub = wd.ctr[2] % w.h(W_SUBSTEPS);
      const float phase = (float)__hiloint2double((int)wd.ctr[5], (int)wd.ctr[4]);
      const uint32_t PF = w.h(W_SF_PATH) * 4, NI = w.h(W_INTEGRAL_COUNT);
      const float* path = (const float*)wd.statF;
      for (uint32_t i = tid; i < driven; i += NT) {
        float ramp = (phase + ((float)sub + 0.5f) * dt) * 60.0f - 0.5f;
        float frame = clampf_(w.h(W_TGT_HELD) ? phase * 60.0f : ramp, 0.0f, (float)(frames - 1u));
        uint32_t fa = (uint32_t)frame, fb = min(fa + 1u, frames - 1u);
        float pa = path[PF + fa * driven + i], pb = path[PF + fb * driven + i];
        float t = frame - (float)fa;
        float value = pa * (1.0f - t) + pb * t;
        if (i < NI) {
          if (sub == 0) {
            float err = clampf_(value - art_[w.h(W_A_Q) + i], -0.08f, 0.08f);
            art_[w.h(W_A_INT) + i] = clampf_(art_[w.h(W_A_INT) + i] + 3.0f * w.hf(W_FRAME_DT) * err, -0.12f, 0.12f);
          }
          value += art_[w.h(W_A_INT) + i];
        }
        art_[w.h(W_A_TGT) + i] = value;
      }
      tm.sync();
    }
  }
  if (w.h(W_CTRL_KIND) == 1) {
    // Demo controller, every substep (update_control, wgsl.js:1821-1825): the fader sets the wheel's target speed.
    if (tid == 0) {
      const float4 cc = w.sf(w.h(W_SF_CTRL));
      art_[w.h(W_A_TGTV) + w.h(W_CTRL_B)] = cc.y * clampf_(art_[w.h(W_A_Q) + w.h(W_CTRL_A)] / cc.x, 0.0f, 1.0f);
    }
    tm.sync();
  }
  for (uint32_t a = tid; a < NA; a += NT) {
    uint4 ai = w.art_info(a);
    if (w.h(W_VLIM_ON)) {
      // Velocity-limit prescale: the articulation's scalar dof velocities scaled so the most exceeded limit is
      // met; then the dofs near their limit (activation fraction) and whether any is (wgsl.js:1828-1855).
      float ratio = 1.0f;
      for (uint32_t k = 0; k < ai.w; k++) {
        float vl = w.dof_passive(ai.z + k).z;
        if (vl < 0.0f || w.fused_dof(ai.z + k)) continue;
        float s = fabsf(art_[w.h(W_A_QD) + ai.z + k]);
        if (s > vl) ratio = mn(ratio, vl / s);
      }
      if (ratio < 1.0f)
        for (uint32_t k = 0; k < ai.w; k++)
          if (w.scalar_dof(ai.z + k) && !w.fused_dof(ai.z + k)) art_[w.h(W_A_QD) + ai.z + k] *= ratio;
      const float frac = w.hf(W_F_VLIM_FRAC);
      bool any = false;
      for (uint32_t k = 0; k < ai.w; k++) {
        uint32_t dd = ai.z + k;
        bool on = fabsf(art_[w.h(W_A_QD) + dd]) >= frac * w.dof_passive(dd).z;
        art_[w.h(W_A_VLG) + dd] = on ? 1.0f : 0.0f;
        any = any || (on && w.dof_passive(dd).z >= 0.0f && w.scalar_dof(dd));
      }
      art_[w.h(W_A_AVL) + a] = any ? 1.0f : 0.0f;
    }
    uint32_t n = 0;
    for (uint32_t k = 0; k < ai.w; k++) {
      uint32_t dd = ai.z + k;
      float4 lim = w.dof_limit(dd);
      bool on = false;
      if (lim.z > 0.5f) {
        uint32_t l = tu(lim.w);
        float q = art_[w.h(W_A_Q) + w.link_qs(l) + dd - w.link_dof(l)];
        float gg = w.hf(W_F_JL_GAP) + w.hf(W_F_JL_SPEED) * dt * fabsf(art_[w.h(W_A_QD) + dd]);
        on = q <= lim.x + gg || q >= lim.y - gg;
      }
      art_[w.h(W_A_LG) + dd] = on ? 1.0f : 0.0f;
      bool rowf = on || w.dof_passive(dd).y > 0.0f || w.h(W_JFL) != 0 || (w.h(W_DRIVE_ROWS) && w.driven_dof(dd));
      if (rowf && w.scalar_dof(dd)) { art_[w.h(W_A_RL) + ai.z + n] = (float)dd; n++; }
    }
    art_[w.h(W_A_RLN) + a] = (float)n;
  }
  }
  if constexpr (PART == 1) return;
  // FK and articulated-body factor, one wave per articulation, tree levels in parallel lanes.
  if constexpr (PART == 0)
  for (uint32_t base = 0; base < NA; base += NT / 64) {
    const uint32_t a = base + (tid >> 6), k = tid & 63;
    if (a < NA) {
      AS_LDS uint32_t* tree = (AS_LDS uint32_t*)(s_tree + (threadIdx.x >> 6) * 64);
      w.art_tree(a, tree);
      w.art_fk_lv(a, k < w.art_info(a).y ? (tree[k] & 0xffu) : 0xffu);
      w.art_factor_lv(a, tree, (AS_LDS float*)(s_pab + (threadIdx.x >> 6) * 512));
    }
  }
  tm.sync();
  // The factored link data the unit-impulse solves read (responses, dof_hinv, dense rows) in LDS when it fits the
  // factor's scratch (no later phase writes S, U, D^-1 or the link COMs); identical values, shorter loads.
  if constexpr (!TM::GRID) {
    const uint32_t cap = (BS / 64) * 512;
    uint32_t need = NL * 4;
    for (uint32_t l = 0; l < NL && need <= cap; l++) need += prep_mirror_floats(w.link_ndof(l));
    if (need <= cap) {
      AS_LDS float* mf = (AS_LDS float*)pl->pab;
      AS_LDS pm_u4* mm = (AS_LDS pm_u4*)pl->pab;
      for (uint32_t l = tid; l < NL; l += NT) {
        uint32_t off = NL * 4;
        for (uint32_t k = 0; k < l; k++) off += prep_mirror_floats(w.link_ndof(k));
        const uint32_t n = w.link_ndof(l);
        mm[l] = pm_u4{(uint32_t)w.link_parent(l), n, w.link_dof(l), off};
        for (uint32_t k = 0; k < n; k++) {
          const SVd sk = w.ldS(l, k), uk = w.ldU(l, k);
          mf[off + k * 6] = sk.a.x; mf[off + k * 6 + 1] = sk.a.y; mf[off + k * 6 + 2] = sk.a.z;
          mf[off + k * 6 + 3] = sk.l.x; mf[off + k * 6 + 4] = sk.l.y; mf[off + k * 6 + 5] = sk.l.z;
          mf[off + 6 * n + k * 6] = uk.a.x; mf[off + 6 * n + k * 6 + 1] = uk.a.y; mf[off + 6 * n + k * 6 + 2] = uk.a.z;
          mf[off + 6 * n + k * 6 + 3] = uk.l.x; mf[off + 6 * n + k * 6 + 4] = uk.l.y; mf[off + 6 * n + k * 6 + 5] = uk.l.z;
          for (uint32_t c = 0; c < n; c++) mf[off + 12 * n + k * n + c] = w.ldDI(l, k, c);
        }
        const f3 r = w.link_r_blocks(l);
        mf[off + 12 * n + n * n] = r.x; mf[off + 12 * n + n * n + 1] = r.y; mf[off + 12 * n + n * n + 2] = r.z;
      }
      tm.sync();
      P.w.pm = PrepMirror{(const AS_LDS float*)mf, (const AS_LDS pm_u4*)mm};
    }
  }
  if (w.h(W_VLIM_FREE)) {
    // Free-body velocity limits: per axis, whether the body's speed is near its limit (wgsl.js:1878-1887).
    const float frac = w.hf(W_F_VLIM_FRAC);
    for (uint32_t b = tid; b < NB; b += NT) {
      if (w.body_kind((int)b) != 1) continue;
      q4 lm = w.body_limits((int)b);
      f3 v = xyz(wd.vel[b * 2]), wv = xyz(wd.vel[b * 2 + 1]);
      const uint32_t g = w.h(W_A_VLG) + ND + b * 6;
      for (int k = 0; k < 3; k++) {
        art_[g + k] = fabsf(comp(v, k)) >= frac * lm.y ? 1.0f : 0.0f;
        art_[g + 3 + k] = fabsf(comp(wv, k)) >= frac * lm.z ? 1.0f : 0.0f;
      }
    }
  }
  tm.sync();
  // ---- Phase 1: classes, run heads and dense slots
  const uint32_t nc = min(wd.ctr[1], MAXC);
  for (uint32_t c = tid; c < nc; c += NT) {
    float4 m = P.con(c, 2);
    int a = ti(m.x), b = ti(m.y);
    const uint64_t pk = P.contact_pair(c);
    uint32_t cls = 0xfffffffeu, head = 0, dense = 0;  // no rows: both nodes asleep (or static)
    if (!w.pair_asleep(a, b)) {
      if (w.same_art(a, b)) {
        cls = 0xffffffffu;
        const bool rws = !w.node_asleep(w.body_node(a));
        dense = rws ? 1u : 0u;
        sched[S.CLOCAL + c] = dense;
      } else {
        cls = (c == 0 || P.contact_pair(c - 1) != pk) ? 1u : 0u;
        head = cls;
      }
    }
    sched[S.CSLOT + c] = cls;
    sched[S.HEADS + c] = head;
    sched[S.DSLOT + c] = dense;
  }
  tm.sync();
  const uint32_t nruns = tm.scan(sched + S.HEADS, nc);
  const uint32_t ndc = tm.scan(sched + S.DSLOT, nc);
  for (uint32_t c = tid; c < nc; c += NT) {
    uint32_t cls = sched[S.CSLOT + c];
    if (cls >= 0xfffffffeu) continue;
    uint32_t unit = sched[S.HEADS + c] + cls;  // inclusive count of heads
    uint32_t h = unit - 1;
    if (h >= UH) { sched[S.CSLOT + c] = 0xfffffffeu; continue; }
    const uint64_t pk = P.contact_pair(c);
    uint32_t run0 = c;
    while (run0 > 0 && P.contact_pair(run0 - 1) == pk) run0--;
    if (c == run0) w.sch64(S.HKEY)[h] = pk;
    atomicAdd(&sched[S.HCNT + h], 1u);
    sched[S.CSLOT + c] = h;
    sched[S.CLOCAL + c] = c - run0;
    float4 m = P.con(c, 2);
    int a = ti(m.x), b = ti(m.y);
    if (a >= 0 && w.body_kind(a) == 2) sched[S.LINKHIT + (uint32_t)a] = 1;
    if (b >= 0 && w.body_kind(b) == 2) sched[S.LINKHIT + (uint32_t)b] = 1;
  }
  tm.sync();
  // ---- Phase 2: units are the run slots in order (unit u = slot h); walk offsets
  const uint32_t nh = min(nruns, UH);
  for (uint32_t h = tid; h < nh; h += NT) sched[S.UWOFF + h] = sched[S.HCNT + h];
  tm.sync();
  const uint32_t nwalk = tm.scan(sched + S.UWOFF, nh);
  const uint32_t nu = min(nh, MAXU);
  for (uint32_t h = tid; h < nu; h += NT) { w.sch64(S.UKEY)[h] = w.sch64(S.HKEY)[h]; sched[S.USIZE + h] = sched[S.HCNT + h]; }
  for (uint32_t l = tid; l < NL; l += NT) {
    if (sched[S.LINKHIT + (uint32_t)w.link_body(l)] == 0) continue;
    int j = (int)l;
    while (j >= 0) { atomicAdd(&sched[S.JSPLIT + (uint32_t)j], 1u); j = w.link_parent((uint32_t)j); }
  }
  tm.sync();
  // Responses of contact-bearing links: six lanes per link (one per unit impulse), ten links per wave.
  for (uint32_t l0 = (tid >> 6) * 10; l0 < NL; l0 += (NT / 64) * 10) {
    const uint32_t lane = tid & 63, l = l0 + lane / 6, e = lane % 6;
    if (lane < 60 && l < NL && sched[S.LINKHIT + (uint32_t)w.link_body(l)] != 0) w.link_response_e(l, e, lane - e, sched + S.JSPLIT);
  }
  for (uint32_t dd = tid; dd < ND; dd += NT) {
    float4 lim = w.dof_limit(dd);
    if (lim.z > 0.5f || w.dof_passive(dd).y > 0.0f || w.dof_passive(dd).z >= 0.0f || w.sf(w.h(W_SF_DOF) + dd * 5 + 3).w > 0.0f ||
        (w.h(W_DRIVE_ROWS) && w.driven_dof(dd))) {
      uint32_t l = tu(lim.w);
      art_[w.h(W_A_HINV) + dd] = w.dof_hinv(l, dd - w.link_dof(l));
    }
  }
  for (uint32_t c = tid; c < nc; c += NT) {
    uint32_t h = sched[S.CSLOT + c];
    if (h >= UH) continue;
    sched[S.WALK + sched[S.UWOFF + h] + sched[S.CLOCAL + c]] = h >= MAXU ? 0xffffffffu : c;
  }
  if (tid == 0) { sched[S.META + 0] = nu; sched[S.META + 2] = nc; }
  for (uint32_t i = tid; i < nu; i += NT) sched[S.USTART + i] = sched[S.USIZE + i];
  tm.sync();
  uint32_t tot = tm.scan(sched + S.USTART, nu);
  if (tid == 0) sched[S.USTART + nu] = tot;
  tm.sync();
  // ---- Phase 5: listing and body-pair reduction
  for (uint32_t c = tid; c < nc; c += NT) {
    uint32_t slot = sched[S.CSLOT + c];
    if (slot >= 0xfffffffeu || slot >= MAXU) continue;
    sched[S.CORDER + sched[S.USTART + slot] + sched[S.CLOCAL + c]] = c;
  }
  tm.sync();
  const uint32_t nl = min(tot, MAXC);
  // Each listed contact's reduction inputs, once: group keys to RK0-2, (pos, gap, position identity) to the
  // solve's eigenbasis table (unused until the solve).
  float4* bpx = (float4*)(sched + S.FPX);
  for (uint32_t p = tid; p < nl; p += NT) {
    Prep::BPC g = P.bp_contact(sched[S.CORDER + p]);
    sched[S.RK0 + p] = g.k0; sched[S.RK1 + p] = g.k1; sched[S.RK2 + p] = g.k2;
    bpx[p] = make_float4(g.pos.x, g.pos.y, g.gap, ubits(g.pk));
  }
  for (uint32_t sp = tid; sp < nu; sp += NT) sched[S.OSTART + sp] = sched[S.USTART + sp];
  tm.sync();
  for (uint32_t p = tid; p < nl; p += NT) {
    uint32_t st = sched[S.USTART + P.unit_of(p, nu)];
    uint32_t k0 = sched[S.RK0 + p], k1 = sched[S.RK1 + p], k2 = sched[S.RK2 + p];
    uint32_t lead = p;
    for (uint32_t m = st; m < p; m++)
      if (sched[S.RK2 + m] == k2 && sched[S.RK0 + m] == k0 && sched[S.RK1 + m] == k1) { lead = m; break; }
    sched[S.LEAD + p] = lead;
    for (uint32_t q = 0; q < 26; q++) sched[S.SLOT + p * 26 + q] = 0;
  }
  tm.sync();
  for (uint32_t p = tid; p < nl; p += NT) {
    const float4 g = bpx[p];
    uint32_t sl = S.SLOT + sched[S.LEAD + p] * 26;
    for (uint32_t q = 0; q < 13; q++) {
      uint2 v = Prep::bp_value_t(mk2(g.x, g.y), g.z, fbits(g.w), q, s_trig);
      if (v.y != 0) atomicMax(&sched[sl + q * 2], v.x);
    }
  }
  tm.sync();
  for (uint32_t p = tid; p < nl; p += NT) {
    const float4 g = bpx[p];
    uint32_t sl = S.SLOT + sched[S.LEAD + p] * 26;
    for (uint32_t q = 0; q < 13; q++) {
      uint2 v = Prep::bp_value_t(mk2(g.x, g.y), g.z, fbits(g.w), q, s_trig);
      if (v.y != 0 && v.x == sched[sl + q * 2]) atomicMax(&sched[sl + q * 2 + 1], v.y);
    }
  }
  tm.sync();
  for (uint32_t p = tid; p < nl; p += NT) {
    const float4 g = bpx[p];
    uint32_t sl = S.SLOT + sched[S.LEAD + p] * 26;
    bool win = false;
    for (uint32_t q = 0; q < 13; q++) {
      uint2 v = Prep::bp_value_t(mk2(g.x, g.y), g.z, fbits(g.w), q, s_trig);
      if (v.y != 0 && v.x == sched[sl + q * 2] && v.y == sched[sl + q * 2 + 1]) win = true;
    }
    sched[S.HEADS + p] = (win || w.h(W_COMPL)) ? 1u : 0u;  // compliant contacts stay whole
  }
  tm.sync();
  for (uint32_t sp = tid; sp < nu; sp += NT) {
    uint32_t n = sched[S.USIZE + sp], st = sched[S.USTART + sp], kept = 0;
    for (uint32_t j = 0; j < n; j++) {
      if (st + j >= MAXC) break;
      if (sched[S.HEADS + st + j] != 0) { sched[S.KEPT + st + kept] = sched[S.CORDER + st + j]; kept++; }
    }
    sched[S.USIZE + sp] = kept;
  }
  tm.sync();
  for (uint32_t i = tid; i < nu; i += NT) sched[S.USTART + i] = sched[S.USIZE + i];
  tm.sync();
  tot = tm.scan(sched + S.USTART, nu);
  if (tid == 0) sched[S.USTART + nu] = tot;
  tm.sync();
  // ---- Phase 6: rows, friction patches
  const uint32_t total = min(tot, MAXC);
  for (uint32_t idx = tid; idx < total; idx += NT) P.build_row(idx, nu);
  tm.sync();
  if (w.h(W_TORS)) {
    for (uint32_t sp = tid; sp < nu; sp += NT) P.build_torsion(sp);
    tm.sync();
  }
  if (PATCH) {
    for (uint32_t sp = tid; sp < nu; sp += NT) P.build_patch(sp);
    tm.sync();
  }
  // ---- Phase 6c: first-fit coloring in walk order (serial by definition, wgsl.js:2282-2363)
  const uint32_t totp = total;
  for (uint32_t c = tid; c < nc; c += NT) sched[S.ROWOF + c] = 0xffffffffu;
  for (uint32_t sp = tid; sp < nu; sp += NT) { sched[S.UTC + sp] = 0; w.sch64(S.UTPW)[sp] = 0; }
  tm.sync();
  for (uint32_t r = tid; r < totp; r += NT) {
    uint32_t sp = P.unit_of(r, nu);
    uint32_t c = sched[S.KEPT + sched[S.OSTART + sp] + r - sched[S.USTART + sp]];
    sched[S.RUNIT + r] = sp; sched[S.RCON + r] = c;
    if (c < nc) sched[S.ROWOF + c] = r;
  }
  tm.sync();
  // First-fit in walk order.  With the masks in LDS (a block team whose world fits), lanes first gather each walk
  // entry's row and body pair into LDS and wave 0 colors 64 entries at a time in rounds: an entry is colored once no
  // earlier entry still pending shares one of its bodies (the lowest pending lane of each body wins an atomicMin), so
  // every entry sees exactly the masks the serial walk gives it and entries colored together touch disjoint bodies.
  // With the masks in global memory, a workgroup with up to PREP_SERIAL_COLOR entries colors them serially on lane 0;
  // a larger world's workgroup, or a grid team's first workgroup, does as wave 0 does over windows of PREP_CPER * BS
  // entries, each thread holding PREP_CPER of them, with the first-pending slots in LINKHIT (free after the link
  // responses).  Per-color counts, unit tags (sums and maxima) and the color-ordered row list (a stable scatter, or
  // for a grid team a stable sort by color) are then built in parallel.
  const uint32_t nbm = max(NB, 1u);
  const bool mask_lds = !TM::GRID && nbm * (NCOLW + 1) <= cap;
  uint32_t* LM = s_pdyn;                // nbm * NCOLW color masks
  uint32_t* OWN = s_pdyn + nbm * NCOLW; // nbm first-pending-lane slots
  if (mask_lds) for (uint32_t i = tid; i < nbm * NCOLW; i += NT) LM[i] = 0;
  const bool wc_lds = !TM::GRID && NCOL + 2 <= PREP_WC;
  uint32_t* WC = wc_lds ? s_wc : sched + S.WCOUNT;
  if (wc_lds) for (uint32_t i = tid; i < NCOL + 2; i += NT) s_wc[i] = 0;
  const uint32_t ncu = min(nwalk, MAXC);
  if (mask_lds) {
    for (uint32_t c0 = 0; c0 < ncu; c0 += PREP_CH) {
      const uint32_t n = min(PREP_CH, ncu - c0);
      tm.sync();
      for (uint32_t i = tid; i < n; i += NT) {
        uint32_t wc = sched[S.WALK + c0 + i];
        uint32_t r = wc < MAXC ? sched[S.ROWOF + min(wc, MAXC - 1)] : 0xffffffffu;
        s_wr[i] = r;
        s_wk[i] = r < totp ? w.sch64(S.UKEY)[sched[S.RUNIT + min(r, MAXC - 1)]] : 0ull;
      }
      tm.sync();
      if (tid < 64) {
        const uint32_t lane = tid;
        for (uint32_t i0 = 0; i0 < n; i0 += 64) {
          const uint32_t i = i0 + lane;
          const uint32_t r = i < n ? s_wr[i] : 0xffffffffu;
          bool pending = r < totp;
          const uint64_t key = pending ? s_wk[i] : 0ull;
          const int a = key_a(key), b = key_b(key);
          while (__ballot(pending) != 0) {
            if (pending) { if (a >= 0) OWN[a] = 64u; if (b >= 0) OWN[b] = 64u; }
            wave_sync();
            if (pending) { if (a >= 0) atomicMin(&OWN[a], lane); if (b >= 0) atomicMin(&OWN[b], lane); }
            wave_sync();
            if (pending && (a < 0 || OWN[a] == lane) && (b < 0 || OWN[b] == lane)) {
              sched[S.PCOL + r] = Prep::first_free_color(LM, NCOLW, NCOL, a, b);
              pending = false;
            }
            wave_sync();
          }
        }
      }
    }
  } else if (!TM::GRID && ncu <= PREP_SERIAL_COLOR) {
    // masks in global memory, one workgroup: lane 0 colors serially, a chunk of walk entries at a time
    for (uint32_t c0 = 0; c0 < ncu; c0 += PREP_CH) {
      const uint32_t n = min(PREP_CH, ncu - c0);
      tm.sync();
      for (uint32_t i = tid; i < n; i += NT) {
        uint32_t wc = sched[S.WALK + c0 + i];
        uint32_t r = wc < MAXC ? sched[S.ROWOF + min(wc, MAXC - 1)] : 0xffffffffu;
        s_wr[i] = r;
        s_wk[i] = r < totp ? w.sch64(S.UKEY)[sched[S.RUNIT + min(r, MAXC - 1)]] : 0ull;
      }
      tm.sync();
      if (tid == 0) {
        for (uint32_t i = 0; i < n; i++) {
          uint32_t r = s_wr[i];
          if (r >= totp) continue;
          const uint64_t key = s_wk[i];
          sched[S.PCOL + r] = Prep::first_free_color(sched + S.WMASK, NCOLW, NCOL, key_a(key), key_b(key));
        }
      }
    }
  } else {
    uint32_t* GO = sched + S.LINKHIT;   // first pending walk index of each body (0xffffffff: none)
    uint32_t* GM = sched + S.WMASK;
    for (uint32_t i = tid; i < nbm; i += NT) GO[i] = 0xffffffffu;
    tm.sync();
    if (!TM::GRID || blockIdx.x == 0) {
      constexpr uint32_t PER = PREP_CPER;
      const uint32_t lt = threadIdx.x;
      for (uint32_t c0 = 0; c0 < ncu; c0 += PER * BS) {
        uint32_t rr[PER];
        int aa[PER], bb[PER];
        uint32_t pend = 0;
        for (uint32_t k = 0; k < PER; k++) {
          const uint32_t i = c0 + k * BS + lt;
          const uint32_t wc = i < ncu ? sched[S.WALK + i] : 0xffffffffu;
          const uint32_t r = wc < MAXC ? sched[S.ROWOF + min(wc, MAXC - 1)] : 0xffffffffu;
          const uint64_t key = r < totp ? w.sch64(S.UKEY)[sched[S.RUNIT + min(r, MAXC - 1)]] : 0ull;
          rr[k] = r; aa[k] = key_a(key); bb[k] = key_b(key);
          if (r < totp) pend |= 1u << k;
        }
        while (__syncthreads_or(pend != 0)) {
          for (uint32_t k = 0; k < PER; k++)
            if ((pend >> k) & 1u) {
              const uint32_t i = c0 + k * BS + lt;
              if (aa[k] >= 0) atomicMin(&GO[aa[k]], i);
              if (bb[k] >= 0) atomicMin(&GO[bb[k]], i);
            }
          __syncthreads();
          uint32_t bid = pend;
          for (uint32_t k = 0; k < PER; k++)
            if ((pend >> k) & 1u) {
              const uint32_t i = c0 + k * BS + lt;
              if ((aa[k] < 0 || __hip_atomic_load(&GO[aa[k]], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == i) &&
                  (bb[k] < 0 || __hip_atomic_load(&GO[bb[k]], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == i)) {
                sched[S.PCOL + rr[k]] = Prep::first_free_color(GM, NCOLW, NCOL, aa[k], bb[k]);
                pend &= ~(1u << k);
              }
            }
          __syncthreads();
          for (uint32_t k = 0; k < PER; k++)
            if ((bid >> k) & 1u) {
              if (aa[k] >= 0) GO[aa[k]] = 0xffffffffu;
              if (bb[k] >= 0) GO[bb[k]] = 0xffffffffu;
            }
          __syncthreads();
        }
      }
    }
  }
  tm.sync();
  for (uint32_t i = tid; i < ncu; i += NT) {
    uint32_t wc = sched[S.WALK + i];
    uint32_t r = wc < MAXC ? sched[S.ROWOF + min(wc, MAXC - 1)] : 0xffffffffu;
    if (r >= totp) continue;
    uint32_t col = sched[S.PCOL + r];
    atomicAdd(&WC[col], 1u);
    atomicMax((unsigned long long*)(w.sch64(S.UTPW) + sched[S.RUNIT + r]), ((unsigned long long)(col + 1) << 32) | wc);
  }
  tm.sync();
  for (uint32_t sp = tid; sp < nu; sp += NT) {
    const uint64_t packed = w.sch64(S.UTPW)[sp];
    sched[S.UTC + sp] = (uint32_t)(packed >> 32);
    sched[S.UTP + sp] = packed != 0 ? sched[S.ROWOF + (uint32_t)packed] : 0xffffffffu;
  }
  if (tid == 0) {
    uint32_t acc = 0, ncol = 0;
    for (uint32_t c = 0; c <= NCOL; c++) {
      uint32_t n = WC[c];
      sched[S.CSTART + c] = acc; WC[c] = acc; acc += n;
      if (n > 0) ncol = c + 1;
    }
    sched[S.CSTART + NCOL + 1] = acc;
    sched[S.META + 1] = ncol;
  }
  tm.sync();
  // Rows in increasing order within each color.  A grid team sorts the rows by color (stable, so rows ascend
  // within a color); in a block team wave 0 takes 64 rows at a time and a lane's slot is its color's running
  // offset plus the number of lower lanes of the same color.
  if constexpr (TM::GRID) {
    for (uint32_t r = tid; r < totp; r += NT) { wd.skey[r] = sched[S.PCOL + r]; wd.sval[r] = r; }
    tm.sync();
    tm.sort(wd.skey, wd.sval, wd.skey + MAXC, wd.sval + MAXC, totp);
    for (uint32_t i = tid; i < totp; i += NT) sched[S.PORD + i] = wd.sval[i];
  } else if (tid < 64) {
    const uint32_t lane = tid;
    for (uint32_t r0 = 0; r0 < totp; r0 += 64) {
      const uint32_t r = r0 + lane;
      const bool valid = r < totp;
      const uint32_t col = valid ? sched[S.PCOL + r] : 0xffffffffu;
      uint32_t rank = 0, cnt = 0;
      bool last = true;
      for (uint32_t j = 0; j < 64; j++) {
        uint32_t cj = __shfl(col, j);
        if (cj == col) { cnt++; if (j < lane) rank++; if (j > lane) last = false; }
      }
      const uint32_t base = valid ? WC[col] : 0u;
      if (!wc_lds) __threadfence_block();
      if (valid) {
        uint32_t slot = base + rank;
        if (slot < MAXC) sched[S.PORD + slot] = r;
        if (last) WC[col] = base + cnt;
      }
      if (!wc_lds) __threadfence_block();
    }
  }
  tm.sync();
  if (tid == 0) {
    // overflow color by contact index (wgsl.js:1740-1780)
    uint32_t o0 = min(sched[S.CSTART + NCOL], MAXC), m = min(sched[S.CSTART + NCOL + 1], MAXC) - o0;
    for (uint32_t i = 1; i < m; i++) {
      uint32_t r = sched[S.PORD + o0 + i];
      uint32_t kr = sched[S.RCON + r];
      int j = (int)i - 1;
      while (j >= 0) {
        uint32_t rj = sched[S.PORD + o0 + j];
        uint32_t kj = sched[S.RCON + rj];
        if (kj < kr || (kj == kr && rj < r)) break;
        sched[S.PORD + o0 + j + 1] = rj;
        j--;
      }
      sched[S.PORD + o0 + j + 1] = r;
    }
    sched[S.META + 3] = ndc + w.loop_rows() + w.h(W_NMIMIC);
  }
  // ---- Phase 7: dense rows of same-articulation contacts (slot = rank in contact order), three lanes per
  // contact, 21 contacts per wave; RK0 (free after phase 5) maps slot -> contact.
  const uint32_t MAXDENSE = w.h(W_MAXDENSE);
  for (uint32_t c = tid; c < nc; c += NT) {
    if (sched[S.CSLOT + c] != 0xffffffffu || sched[S.CLOCAL + c] == 0) continue;
    uint32_t slot = sched[S.DSLOT + c];
    if (slot < MAXDENSE) sched[S.RK0 + slot] = c;
  }
  tm.sync();
  const uint32_t ndd = min(ndc, MAXDENSE);
  for (uint32_t s0 = (tid >> 6) * 21; s0 < ndd; s0 += (NT / 64) * 21) {
    const uint32_t lane = tid & 63, sl = s0 + lane / 3, r = lane % 3;
    if (lane < 63 && sl < ndd) P.dense_row(sched[S.RK0 + sl], sl, r, lane - r);
  }
  // Loop closures, then mimic rows, in the slots after the contacts' (wgsl.js:2445-2492), a lane each.
  const uint32_t nloop = w.h(W_NLOOP), nlr = w.loop_rows();
  for (uint32_t i = tid; i < nloop; i += NT) {
    uint32_t slot = ndc;
    for (uint32_t k = 0; k < i; k++) slot += w.sf(w.h(W_SF_LOOP) + k * 5).w > 0.5f ? 2u : 1u;
    if (slot + (w.sf(w.h(W_SF_LOOP) + i * 5).w > 0.5f ? 1u : 0u) < MAXDENSE) P.loop_row(i, slot);
  }
  for (uint32_t i = tid; i < w.h(W_NMIMIC); i += NT)
    if (ndc + nlr + i < MAXDENSE) P.mimic_row(i, ndc + nlr + i);
}


template <int BS, int PART = 0>
__global__ void __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(STEEL_WPE_PREPARE))) k_prepare(const WorldDev* __restrict__ worlds, uint32_t nworlds, uint32_t cap) {
  __shared__ uint32_t Hs[STEEL_WORLD_HEADER_WORDS];
  __shared__ TeamLds<BS> tl;
  __shared__ PrepLds<BS> pl;
  extern __shared__ uint32_t s_pdyn[];
  const uint32_t wid = blockIdx.x;
  if (wid >= nworlds) return;
  const BlockTeam<BS> tm{threadIdx.x, &tl, nullptr};
  prepare_world<BS, BlockTeam<BS>, PART>(tm, worlds[wid], Hs, cap, s_pdyn, &pl);
}

// Forward kinematics and the articulated-body factor of world 0's articulations, one 64-thread workgroup per
// articulation (a, a + gridDim.x, ...): the same per-articulation passes prepare runs on its waves, spread over the
// device for a world of many articulations.  Articulations share no links, so their order does not matter.
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(STEEL_WPE_PREPARE))) k_art_factor(const WorldDev* __restrict__ world) {
  __shared__ uint32_t Hs[STEEL_WORLD_HEADER_WORDS];
  __shared__ uint32_t tree[64];
  __shared__ float pab[512];
  const WorldDev wd = world[0];
  load_header(Hs, wd);
  const W w = make_world(Hs, wd);
  const uint32_t NA = w.NA(), k = threadIdx.x;
  for (uint32_t a = blockIdx.x; a < NA; a += gridDim.x) {
    w.art_tree(a, (AS_LDS uint32_t*)tree);
    w.art_fk_lv(a, k < w.art_info(a).y ? (tree[k] & 0xffu) : 0xffu);
    w.art_factor_lv(a, (const AS_LDS uint32_t*)tree, (AS_LDS float*)pab);
    __syncthreads();
  }
}

// One large world over a cooperative launch (g: grid team scratch).
template <int BS>
__global__ void __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(STEEL_WPE_PREPARE))) k_prepare_grid(const WorldDev* __restrict__ world, uint32_t* g) {
  __shared__ uint32_t Hs[STEEL_WORLD_HEADER_WORDS];
  __shared__ TeamLds<BS> tl;
  __shared__ PrepLds<BS> pl;
  __shared__ SortLds<BS> sl;
  const GridTeam<BS> tm{gridDim.x * BS, blockIdx.x * BS + threadIdx.x, &tl, &sl, g};
  prepare_world<BS>(tm, world[0], Hs, 0, nullptr, &pl);
}

// ------------------------------------------------------------------------------------------------
// k_solve (wgsl.js:2501-3891)
// ------------------------------------------------------------------------------------------------
// Per-link mirror of the fields the solve sweeps touch, in LDS for the whole kernel.  A link of n dofs takes
// mirror_floats(n) = n^2 + 14n + 21 floats: S (6n), U (6n), D^-1 (n x n), p, dv, v (6 each), u, tau (n each),
// COM (3).  Loaded from the link blocks (art[A_LINK + l*LB]) at kernel start; p, dv, v, u, tau are written
// back after the sweeps.
// Link-block offset of mirror field f of a link with n dofs (the world's link blocks use stride JM); LB: none.
__host__ __device__ inline uint32_t mirror_src(uint32_t f, uint32_t n, uint32_t JM) {
  if (f < 6 * n) return L_S + f;
  f -= 6 * n;
  if (f < 6 * n) return L_U + f;
  f -= 6 * n;
  if (f < n * n) return L_DI + (f / n) * JM + f % n;
  f -= n * n;
  if (f < 6) return L_P + f;
  f -= 6;
  if (f < 6) return L_DV + f;
  f -= 6;
  if (f < 6) return L_V + f;
  f -= 6;
  if (f < n) return L_UU + f;
  f -= n;
  if (f < n) return L_TAU + f;
  f -= n;
  if (f < 3) return L_COM + f;
  return LB;
}
// Mirror fields written back after the sweeps (p, dv, v, u, tau).
__host__ __device__ inline bool mirror_mutable(uint32_t f, uint32_t n) { return f >= 12 * n + n * n && f < 12 * n + n * n + 18 + 2 * n; }
struct Solver {
  W w;
  uint32_t parity;
  float* lm;      // LDS link mirror
  uint4* lmeta;   // LDS link records: (parent, ndof | mirror offset << 8, first dof, body)
  uint4* ltree;   // LDS link tree: (depth, highest-index child, next lower sibling, -)
  float* lcon;    // LDS per-link contribution to the parent's bias force (8 floats per link)
  float* laqd;    // LDS joint velocities (the reference's aqd)
  float4* tw;     // body twists (LDS, or the world's buffer for worlds too large for LDS)
  float4* imp;    // accumulated impulses on articulated bodies (same placement as tw)
  float* red;     // LDS partial sums of the momentum restart test (256 floats; Nesterov worlds only)
  bool vpass = false;  // velocity pass: rows without position bias or regularization
  DEV int mpar(uint32_t l) const { return (int)lmeta[l].x; }
  DEV uint32_t mnd(uint32_t l) const { return lmeta[l].y & 0xffu; }
  DEV uint32_t mdof(uint32_t l) const { return lmeta[l].z; }
  DEV uint32_t mbody(uint32_t l) const { return lmeta[l].w; }
  DEV float* ml(uint32_t l) const { return lm + (lmeta[l].y >> 8); }
  static DEV SVd msv(const float* p) { return mksv(mk3(p[0], p[1], p[2]), mk3(p[3], p[4], p[5])); }
  static DEV void mst(float* p, SVd v) { p[0] = v.a.x; p[1] = v.a.y; p[2] = v.a.z; p[3] = v.l.x; p[4] = v.l.y; p[5] = v.l.z; }
  DEV SVd mS(uint32_t l, uint32_t k) const { return msv(ml(l) + k * 6); }
  DEV SVd mU(uint32_t l, uint32_t k) const { return msv(ml(l) + 6 * mnd(l) + k * 6); }
  DEV float mDI(uint32_t l, uint32_t r, uint32_t c) const { uint32_t n = mnd(l); return ml(l)[12 * n + r * n + c]; }
  DEV float* mP(uint32_t l) const { uint32_t n = mnd(l); return ml(l) + 12 * n + n * n; }
  DEV float* mDV(uint32_t l) const { return mP(l) + 6; }
  DEV float* mV(uint32_t l) const { return mP(l) + 12; }
  DEV float* mUU(uint32_t l) const { return mP(l) + 18; }
  DEV float* mTAU(uint32_t l) const { return mP(l) + 18 + mnd(l); }
  DEV f3 mcom(uint32_t l) const { const float* p = mP(l) + 18 + 2 * mnd(l); return mk3(p[0], p[1], p[2]); }
  DEV f3 mr(uint32_t l) const {  // link_r from the mirror: COM(l) - COM(parent), zero at a root
    int par = mpar(l);
    if (par < 0) return mk3(0, 0, 0);
    return mcom(l) - mcom((uint32_t)par);
  }
  DEV float4 row(uint32_t r, uint32_t k) const { return w.d.rows[r * w.ROWS() + k]; }
  // Contact warm start (wgsl.js:3614-3631): row i's key (contact rows, then dense rows; 0xffffffff for a bilateral
  // row, which stores nothing), its value (normal impulse, then the tangent impulse as a world vector) and its order
  // among the rows sharing a key: the larger normal impulse (negative and -0 count as 0, and non-negative floats
  // order as their bits), then the lower row.
  DEV uint32_t warm_key(uint32_t i, uint32_t nc) const {
    if (i < nc) return (uint32_t)row(i, 3).w | ((uint32_t)row(i, 4).w << 16);
    const uint32_t o = w.h(W_A_DENSE) + (i - nc) * w.h(W_DPS);
    if (w.d.art[o + 21] > 0.5f) return 0xffffffffu;
    return (uint32_t)w.d.art[o + 1] | ((uint32_t)w.d.art[o + 2] << 16);
  }
  DEV float4 warm_value(uint32_t i, uint32_t nc) const {
    if (i < nc) {
      const float4 lam = row(i, 6);
      return F4(xyz(row(i, 1)) * lam.y + xyz(row(i, 2)) * lam.z, lam.x);  // (tangent, normal): reordered on store
    }
    const uint32_t o = w.h(W_A_DENSE) + (i - nc) * w.h(W_DPS);
    const AS_GLOBAL float* art_ = (const AS_GLOBAL float*)w.d.art;
    return F4(w.ld3(o + 12) * art_[o + 10] + w.ld3(o + 15) * art_[o + 11], art_[o + 9]);
  }
  DEV unsigned long long warm_order(float x, uint32_t i) const {
    return ((unsigned long long)(x > 0.0f ? fbits(x) + 1u : 1u) << 32) | (0xffffffffu - i);
  }
  // Twists of the velocity pass's snapshot (wgsl.js:2930-2934).
  DEV void snap_tw(int b, f3 out[2]) const {
    if (b < 0) { out[0] = mk3(0, 0, 0); out[1] = mk3(0, 0, 0); return; }
    const uint32_t o = w.h(W_A_SNAP) + (uint32_t)b * 8;
    out[0] = w.ld3(o); out[1] = w.ld3(o + 4);
  }
  DEV void load_tw(int b, f3 out[2]) const {
    if (b < 0) { out[0] = mk3(0, 0, 0); out[1] = mk3(0, 0, 0); return; }
    out[0] = xyz(tw[(uint32_t)b * 2]); out[1] = xyz(tw[(uint32_t)b * 2 + 1]);
  }
  // Nesterov momentum on contact row r (wgsl.js:3499-3523): y = lam + beta (lam - lam_last), the change e = y - lam
  // applied through the row's responses; rows[MO] keeps lam, rows[MO + 1] keeps e; init records lam only.
  DEVN void momentum_unit(uint32_t r, int a, int b, float beta, bool init) const {
    float4* R = w.d.rows + (size_t)r * w.ROWS();
    const uint32_t MO = w.h(W_MO), PO = w.PO();
    const float4 lam = R[6];
    const f3 l3 = xyz(lam);
    if (init) { R[MO] = F4(l3, 0.0f); R[MO + 1] = make_float4(0, 0, 0, 0); return; }
    const f3 y = l3 + beta * (l3 - xyz(R[MO]));
    const f3 e = y - l3;
    R[MO] = F4(l3, 0.0f); R[MO + 1] = F4(e, 0.0f);
    if (e.x == 0.0f && e.y == 0.0f && e.z == 0.0f) return;
    R[6] = F4(y, lam.w);
    f3 ta[2], tb[2];
    load_tw(a, ta); load_tw(b, tb);
    ta[0] += xyz(R[7]) * e.x + xyz(R[11]) * e.y + xyz(R[15]) * e.z;
    ta[1] += xyz(R[8]) * e.x + xyz(R[12]) * e.y + xyz(R[16]) * e.z;
    tb[0] += xyz(R[9]) * e.x + xyz(R[13]) * e.y + xyz(R[17]) * e.z;
    tb[1] += xyz(R[10]) * e.x + xyz(R[14]) * e.y + xyz(R[18]) * e.z;
    const f3 ra = xyz(R[3]), rb = xyz(R[4]);
    f3 fa = ra, fb = rb;
    if (w.h(W_EXTRA)) { fa = xyz(R[PO]); fb = xyz(R[PO + 1]); }
    const f3 fn = xyz(R[0]) * e.x, ft = xyz(R[1]) * e.y + xyz(R[2]) * e.z;
    add_imp(a, -fn - ft, -cross(ra, fn) - cross(fa, ft));
    add_imp(b, fn + ft, cross(rb, fn) + cross(fb, ft));
    if (a >= 0) { tw[(uint32_t)a * 2] = F4(ta[0], 0.0f); tw[(uint32_t)a * 2 + 1] = F4(ta[1], 0.0f); }
    if (b >= 0) { tw[(uint32_t)b * 2] = F4(tb[0], 0.0f); tw[(uint32_t)b * 2 + 1] = F4(tb[1], 0.0f); }
  }
  DEV void add_imp(int b, f3 f, f3 t) const {
    if (b < 0) return;
    if (w.body_kind(b) != 2) return;
    float4 i0 = imp[(uint32_t)b * 2], i1 = imp[(uint32_t)b * 2 + 1];
    imp[(uint32_t)b * 2] = make_float4(i0.x + f.x, i0.y + f.y, i0.z + f.z, i0.w + 0.0f);
    imp[(uint32_t)b * 2 + 1] = make_float4(i1.x + t.x, i1.y + t.y, i1.z + t.z, i1.w + 0.0f);
  }
  // ---- Steel's joint torsional/rolling cone for articulated contacts (PropagationAngularFrictionBlock, elliptic cone;
  // wgsl.js:2655-2925): rows (t1, t2, spin, roll1, roll2) minimize their quadratic over the ball |lambda / mu| <= N,
  // by a More-Sorensen trust region (proximal fallback), relaxed by omega and line-searched from the current
  // impulses moved into the cone.
  struct Ball { float y[5]; bool ok, interior; };
  static DEV float cone_norm(const float* y) { float s = 0; for (uint32_t k = 0; k < 5; k++) s += y[k] * y[k]; return sqrtf(s); }
  static DEV void cone_project(float* y, float load) {
    float n = cone_norm(y);
    if (n > load) for (uint32_t k = 0; k < 5; k++) y[k] *= load / n;
  }
  static DEV bool live5(const float* H, const bool* act, uint32_t k) { return act[k] && H[k * 5 + k] > 0.0f; }
  // Cholesky of H + alpha I over the live rows; fails when a pivot keeps under 1e-6 of its diagonal.
  static DEV bool factor5(const float* H, const bool* act, float alpha, float* L) {
    for (uint32_t k = 0; k < 5; k++)
      for (uint32_t l = 0; l <= k; l++) {
        float v = (live5(H, act, k) && live5(H, act, l)) ? H[k * 5 + l] : 0.0f;
        if (k == l) v = live5(H, act, k) ? v + alpha : 1.0f;
        for (uint32_t m = 0; m < l; m++) v -= L[k * 5 + m] * L[l * 5 + m];
        if (k == l) {
          if (!(v > 1e-6f * (live5(H, act, k) ? H[k * 5 + k] + alpha : 1.0f))) return false;
          L[k * 5 + k] = sqrtf(v);
        } else {
          L[k * 5 + l] = v / L[l * 5 + l];
        }
      }
    return true;
  }
  static DEV void lower5(const float* L, float* x) {
    for (uint32_t k = 0; k < 5; k++) {
      for (uint32_t m = 0; m < k; m++) x[k] -= L[k * 5 + m] * x[m];
      x[k] /= L[k * 5 + k];
    }
  }
  static DEV void upper5(const float* L, float* x) {
    for (uint32_t kk = 0; kk < 5; kk++) {
      uint32_t k = 4 - kk;
      for (uint32_t m = k + 1; m < 5; m++) x[k] -= L[m * 5 + k] * x[m];
      x[k] /= L[k * 5 + k];
    }
  }
  // y = -(H + alpha I)^-1 c with alpha by Newton on 1/|y(alpha)| inside a bracket; ok when it met the 1e-5 boundary
  // residual or found an interior minimizer.
  static DEVN Ball cholesky_ball(const float* H, const float* c, const bool* act, float load) {
    Ball r;
    float b[5], cn = 0;
    for (uint32_t k = 0; k < 5; k++) { b[k] = live5(H, act, k) ? c[k] : 0.0f; cn += b[k] * b[k]; r.y[k] = 0; }
    cn = sqrtf(cn);
    r.interior = !(cn > 0.0f);
    r.ok = true;
    if (!(load > 0.0f && cn > 0.0f)) return r;
    r.ok = false;
    float lo = 0, hi = cn / load, alpha = 0;
    for (uint32_t it = 0; it < 40; it++) {
      float nx = -1;
      float L[25] = {};
      if (factor5(H, act, alpha, L)) {
        float yy[5];
        for (uint32_t k = 0; k < 5; k++) yy[k] = -b[k];
        lower5(L, yy); upper5(L, yy);
        for (uint32_t k = 0; k < 5; k++) r.y[k] = yy[k];
        float n = cone_norm(r.y);
        if (alpha == 0.0f && n <= load) { r.interior = true; r.ok = true; return r; }
        if (fabsf(n - load) <= 1e-5f * load) { r.ok = true; break; }
        if (n < load) hi = alpha; else lo = alpha;
        float wv[5];
        for (uint32_t k = 0; k < 5; k++) wv[k] = r.y[k];
        lower5(L, wv);
        float wn = 0;
        for (uint32_t k = 0; k < 5; k++) wn += wv[k] * wv[k];
        nx = alpha + n * n / wn * (n - load) / load;
      } else {
        lo = alpha;
      }
      if (hi - lo <= 1e-4f * hi) break;
      alpha = (nx > lo && nx < hi) ? nx : (lo > 0.0f ? sqrtf(lo * hi) : 1e-3f * hi);
    }
    float n = cone_norm(r.y);
    if (n > load) for (uint32_t k = 0; k < 5; k++) r.y[k] *= load / n;
    return r;
  }
  // Minimize 0.5 y'Hy + c'y over |y| <= load; when dependent rows defeat Cholesky, a proximal term
  // 0.5 rho (y - center/scale)' diag(H) (y - center/scale), rho raised tenfold until the ball solve converges.
  static DEVN Ball trust_region(const float* H, const float* c, const bool* act, float load, const float* center, const float* scale) {
    Ball first = cholesky_ball(H, c, act, load);
    if (first.ok) return first;
    float ce[5];
    for (uint32_t k = 0; k < 5; k++) ce[k] = (act[k] && scale[k] > 0.0f) ? center[k] / scale[k] : 0.0f;
    float rho = 1e-5f;
    for (uint32_t it = 0; it < 7; it++) {
      float K[25], g[5];
      for (uint32_t i = 0; i < 25; i++) K[i] = H[i];
      for (uint32_t k = 0; k < 5; k++) {
        float wt = live5(H, act, k) ? rho * H[k * 5 + k] : 0.0f;
        K[k * 5 + k] += wt;
        g[k] = c[k] - wt * ce[k];
      }
      rho *= 10.0f;
      Ball t = cholesky_ball(K, g, act, load);
      if (!t.ok) continue;
      bool fin = true;
      for (uint32_t k = 0; k < 5; k++) fin = fin && fabsf(t.y[k]) < 3.0e38f;
      if (fin) return t;
    }
    return first;
  }
  DEVN void ell_block(const float* H0, const float* g0, const float* mu_in, const float* lam, float load_in, float* out) const {
    float load = mx(load_in, 0.0f);
    float mu[5], H[25], gr[5], big = 0;
    for (uint32_t i = 0; i < 25; i++) H[i] = H0[i];
    for (uint32_t k = 0; k < 5; k++) { gr[k] = g0[k]; mu[k] = mx(mu_in[k], 0.0f); big = mx(big, H[k * 5 + k]); }
    bool act[5];
    float comp[5];
    for (uint32_t k = 0; k < 5; k++) {
      act[k] = mu[k] > 0.0f && H[k * 5 + k] > 1e-12f * big;
      comp[k] = (k >= 2 && act[k] && load > 0.0f) ? w.hf(W_F_CREEP) / (mu[k] * mu[k] * load) : 0.0f;
      H[k * 5 + k] += comp[k]; gr[k] += comp[k] * lam[k];
    }
    float y0[5], c[5], y[5];
    for (uint32_t k = 0; k < 5; k++) {
      y0[k] = act[k] ? lam[k] / mu[k] : 0.0f;
      gr[k] = act[k] ? gr[k] * mu[k] : 0.0f;
      for (uint32_t l = 0; l < 5; l++) H[k * 5 + l] *= mu[k] * mu[l];
    }
    for (uint32_t k = 0; k < 5; k++) {
      float v = gr[k];
      for (uint32_t l = 0; l < 5; l++) v -= H[k * 5 + l] * y0[l];
      c[k] = act[k] ? v : 0.0f;
    }
    Ball tr = trust_region(H, c, act, load, lam, mu_in);
    for (uint32_t k = 0; k < 5; k++) y[k] = tr.y[k];
    bool stick = tr.interior && cone_norm(y) <= load;
    if (!stick) cone_project(y, load);
    float stp[5], shift[5];
    for (uint32_t k = 0; k < 5; k++) stp[k] = lam[k] + w.hf(W_F_OMEGA) * (mu[k] * y[k] - lam[k]);
    if (w.hf(W_F_OMEGA) != 1.0f) {
      for (uint32_t k = 0; k < 5; k++) y[k] = act[k] ? stp[k] / mu_in[k] : 0.0f;
      cone_project(y, load);
      for (uint32_t k = 0; k < 5; k++) stp[k] = mu[k] * y[k];
    }
    for (uint32_t k = 0; k < 5; k++) y[k] = act[k] ? lam[k] / mu_in[k] : 0.0f;
    cone_project(y, load);
    for (uint32_t k = 0; k < 5; k++) { float b = mu[k] * y[k]; shift[k] = b - lam[k]; stp[k] -= b; }
    float slope = 0, curv = 0;
    for (uint32_t k = 0; k < 5; k++) {
      float hs = 0, hp = 0;
      for (uint32_t l = 0; l < 5; l++) { hs += H0[k * 5 + l] * shift[l]; hp += H0[k * 5 + l] * stp[l]; }
      slope += stp[k] * (g0[k] + hs + comp[k] * (lam[k] + shift[k]));
      curv += stp[k] * (hp + comp[k] * stp[k]);
    }
    float t = slope < 0.0f ? (curv > -slope ? -slope / curv : 1.0f) : 0.0f;
    for (uint32_t k = 0; k < 5; k++) out[k] = lam[k] + shift[k] + (t > 0.0f ? t * stp[k] : 0.0f);
  }
  // ---- Steel's coupled contact solves (friction_mode bisection, bisection_desaxce, coulomb_newton; wgsl.js:2564-2653):
  // u = live (normal, t1, t2) velocities, G the 3x3 Delassus block (nn, nt1, nt2 | t1t1, t1t2, t2t2).
  static DEV f2 tangent_step(f2 ut, float4 g1) {
    const float det = g1.x * g1.z - g1.y * g1.y;
    if (fabsf(det) <= 1e-20f) return mk2(0.0f, 0.0f);
    return mk2((-ut.x * g1.z + ut.y * g1.y) / det, (ut.x * g1.y - ut.y * g1.x) / det);
  }
  static DEV f2 clip_cone(f2 t, float lim) {
    const float m = len(t);
    if (m > lim && m > 1e-20f) return t * (lim / m);
    return t;
  }
  static DEVN f3 fixed_normal_pair(f2 ut0, float4 g1, float mu, f3 lam) {
    const f2 f = friction_pair_solve(g1.x, g1.y, g1.z, ut0.x, ut0.y, lam.y, lam.z, mx(mu * lam.x, 0.0f), 1.0f);
    return mk3(lam.x, f.x, f.y);
  }
  static DEVN f3 contact_block(uint32_t mode, float un, f2 ut0, float4 g0, float4 g1, float vtarget0, float mu, f3 lam) {
    const float gnn = g0.x;
    const f2 gnt = mk2(g0.y, g0.z);
    const float trace = g1.x + g1.z;
    if (g1.x * g1.z - g1.y * g1.y <= 1e-6f * trace * trace) return fixed_normal_pair(ut0, g1, mu, lam);
    if (gnn < 1e-20f) return lam;
    const f2 lt = mk2(lam.y, lam.z);
    float vt_n = vtarget0;
    if (mode == 2) vt_n += mu * len(ut0);
    if (mode == 3) {
      const float un_f = un - (gnn * lam.x + dot(gnt, lt));
      const f2 ut_f = ut0 - mk2(gnt.x * lam.x + g1.x * lt.x + g1.y * lt.y, gnt.y * lam.x + g1.y * lt.x + g1.z * lt.y);
      if (un_f >= vt_n) return mk3(0, 0, 0);
      const float bN = un_f - vt_n;
      const float A00 = g1.x - gnt.x * gnt.x / gnn, A01 = g1.y - gnt.x * gnt.y / gnn, A11 = g1.z - gnt.y * gnt.y / gnn;
      if (A00 * A11 - A01 * A01 <= 1e-6f * trace * trace) return fixed_normal_pair(ut0, g1, mu, lam);
      const f2 c = ut_f - (bN / gnn) * gnt;
      const float id0 = 1.0f / (A00 * A11 - A01 * A01);
      f2 s = mk2(A11 * c.x - A01 * c.y, A00 * c.y - A01 * c.x) * id0;
      const float phi0 = len(s) - mu * (dot(gnt, s) - bN) / gnn;
      if (phi0 > 0.0f) {
        float hi = 1.0f;
        for (uint32_t e = 0; e < 30; e++) {
          const float a = A00 + hi, d = A11 + hi, idt = 1.0f / (a * d - A01 * A01);
          const f2 sh = mk2(d * c.x - A01 * c.y, a * c.y - A01 * c.x) * idt;
          if (len(sh) - mu * (dot(gnt, sh) - bN) / gnn < 0.0f) break;
          hi *= 2.0f;
        }
        float lo = 0.0f, x = 0.5f * hi;
        const float tol = 1e-6f + 1e-6f * phi0;
        for (uint32_t k = 0; k < 50; k++) {
          const float a = A00 + x, d = A11 + x, idt = 1.0f / (a * d - A01 * A01);
          const f2 sx = mk2(d * c.x - A01 * c.y, a * c.y - A01 * c.x) * idt;
          const f2 tx = mk2(d * sx.x - A01 * sx.y, a * sx.y - A01 * sx.x) * idt;
          const float ns = len(sx);
          const float fx = ns - mu * (dot(gnt, sx) - bN) / gnn;
          const float dfx = -dot(sx, tx) / ns + mu * dot(gnt, tx) / gnn;
          s = sx;
          if (fabsf(fx) < tol || fabsf(hi - lo) < 1e-6f * (1.0f + hi)) break;
          if (fx > 0.0f) lo = x; else hi = x;
          float xn = 0.5f * (lo + hi);
          if (dfx != 0.0f) { const float xt = x - fx / dfx; if (xt > lo && xt < hi) xn = xt; }
          x = xn;
        }
      }
      const f2 rt = -1.0f * s;
      const float ln = mx(-(dot(gnt, rt) + bN) / gnn, 0.0f);
      const float lim = mx(mu * ln, 0.0f);
      const float m = len(rt);
      const f2 r2 = m > lim ? rt * (lim / m) : rt;
      return mk3(ln, r2.x, r2.y);
    }
    // Bisection on the normal impulse with the tangent pair re-solved and clipped at each trial.
    const float u0 = un - gnn * lam.x - dot(gnt, lt);
    if (u0 >= vt_n) return mk3(0, 0, 0);
    float lo = 0.0f;
    float hi = mx(mx(2.0f * lam.x, (vt_n - un) / gnn + lam.x), 1.0f);
    float flo = u0 - vt_n, fhi = -1.0f;
    for (uint32_t k = 0; k < 20; k++) {
      const float mid = 0.5f * (lo + hi);
      const float dn = mid - lam.x;
      const f2 tr = clip_cone(lt + tangent_step(ut0 + gnt * dn, g1), mu * mid);
      const float u = un + gnn * dn + dot(gnt, tr - lt);
      if (u < vt_n) { lo = mid; flo = u - vt_n; } else { hi = mid; fhi = u - vt_n; }
    }
    float ln = 0.5f * (lo + hi);
    if (fhi >= 0.0f) ln = lo - flo * (hi - lo) / (fhi - flo);
    const f2 tf = clip_cone(lt + tangent_step(ut0 + gnt * (ln - lam.x), g1), mu * ln);
    return mk3(ln, tf.x, tf.y);
  }
  // The part of the friction pair solve that depends on the row's constants only: the scaled 2x2 response's
  // eigenvalues and eigenvector.
  struct FricConst { float scale, largest, smallest, vx, vy; };
  static DEV FricConst friction_const(float a0, float c0, float d0) {
    FricConst k;
    k.scale = mx(mx(a0, d0), 1e-20f);
    float a = a0 / k.scale, c = c0 / k.scale, d = d0 / k.scale;
    k.largest = 0.5f * (a + d + sqrtf((a - d) * (a - d) + 4.0f * c * c));
    k.smallest = k.largest > 0.0f ? mx((a * d - c * c) / k.largest, 0.0f) : 0.0f;
    float vx = c, vy = k.largest - a;
    if (a >= d) { vx = k.largest - d; vy = c; }
    float vn = sqrtf(vx * vx + vy * vy);
    if (vn > 0.0f) { vx /= vn; vy /= vn; } else { vx = 1.0f; vy = 0.0f; }
    k.vx = vx; k.vy = vy;
    return k;
  }
  static DEV f2 friction_pair_solve(float a0, float c0, float d0, float r0, float r1, float old0, float old1, float radius, float omega) {
    if (radius <= 0.0f) return mk2(0, 0);
    return friction_pair_apply(friction_const(a0, c0, d0), r0, r1, old0, old1, radius, omega);
  }
  static DEV f2 friction_pair_apply(const FricConst& fk, float r0, float r1, float old0, float old1, float radius, float omega) {
    if (radius <= 0.0f) return mk2(0, 0);
    const float scale = fk.scale, largest = fk.largest, smallest = fk.smallest, vx = fk.vx, vy = fk.vy;
    float r0r = (vx * r0 + vy * r1) / scale;
    float r1r = (-vy * r0 + vx * r1) / scale;
    f2 res = mk2(old0, old1);
    bool sticking = false;
    if (largest > 0.0f && (smallest > 0.0f || r1r == 0.0f)) {
      float dx = r0r / largest;
      float dy = smallest > 0.0f ? r1r / smallest : 0.0f;
      res = mk2(old0 - (vx * dx - vy * dy), old1 - (vy * dx + vx * dy));
      sticking = len(res) <= radius;
    }
    if (!sticking) {
      float on = len(mk2(old0, old1)), rn = len(mk2(r0, r1));
      if (on > 0.0f && on <= radius && radius - on <= 2e-7f * radius && r0 * old0 + r1 * old1 <= 0.0f &&
          fabsf(r0 * old1 - r1 * old0) <= 2e-7f * on * rn)
        return mk2(old0, old1);
      float ox = vx * old0 + vy * old1, oy = -vy * old0 + vx * old1;
      float b0 = largest * ox - r0r, b1 = smallest * oy - r1r;
      float x = 0.0f, y = 0.0f;
      float hi = sqrtf(b0 * b0 + b1 * b1) / radius;
      if (hi > 0.0f) {
        float lo_ = mx(mx(hi - largest, fabsf(b1) / radius - smallest), 0.0f);
        float alpha = lo_;
        bool done = false;
        for (uint32_t it = 0; it < 8; it++) {
          float i0 = 1.0f / (largest + alpha);
          float i1 = smallest + alpha > 0.0f ? 1.0f / (smallest + alpha) : 0.0f;
          float tx = ox - (r0r + alpha * ox) * i0, ty = oy - (r1r + alpha * oy) * i1;
          float nrmv = sqrtf(tx * tx + ty * ty);
          if (fabsf(nrmv - radius) <= 2e-7f * radius || (alpha == 0.0f && nrmv <= radius)) { x = tx; y = ty; done = true; break; }
          if (nrmv > radius) lo_ = alpha; else hi = alpha;
          float slope = tx * tx * i0 + ty * ty * i1;
          float cand = alpha;
          if (slope > 0.0f) cand = alpha + (nrmv / radius - 1.0f) * nrmv * nrmv / slope;
          alpha = (cand > lo_ && cand < hi) ? cand : 0.5f * (lo_ + hi);
        }
        if (!done) {
          for (uint32_t it = 0; it < 24; it++) {
            alpha = 0.5f * (lo_ + hi);
            if (len(mk2(b0 / (largest + alpha), b1 / (smallest + alpha))) > radius) lo_ = alpha; else hi = alpha;
          }
          x = b0 / (largest + hi); y = b1 / (smallest + hi);
        }
      }
      res = mk2(old0 + vx * (x - ox) - vy * (y - oy), old1 + vy * (x - ox) + vx * (y - oy));
    }
    f2 oldv = mk2(old0, old1);
    f2 relaxed = oldv + omega * (res - oldv);
    float m = len(relaxed);
    return m > radius ? relaxed * (radius / m) : relaxed;
  }
  // EXT: the off-default row kinds (coupled friction modes, spin/rolling rows, contact torsion, compliance) are
  // compiled in; the default path instantiates without them, so it makes no calls and keeps its registers.
  template <bool EXT>
  DEVN void solve_unit(uint32_t us, uint32_t r, int a, int b, uint32_t mode) const {  // wgsl.js:3033-3231
    const uint32_t PO = w.PO();
    const bool EXTRA = w.h(W_EXTRA) != 0, PATCH = w.h(W_PATCH) != 0;
    AS_GLOBAL float* art_ = (AS_GLOBAL float*)w.d.art;
    f3 ta[2], tb[2];
    load_tw(a, ta); load_tw(b, tb);
    float4 r0 = row(r, 0), r1 = row(r, 1), r2 = row(r, 2);
    f3 ra = xyz(row(r, 3)), rb = xyz(row(r, 4));
    f3 n = xyz(r0), t1 = xyz(r1), t2 = xyz(r2);
    float4 lam = row(r, 6);
    float dn = 0.0f, dt1 = 0.0f, dt2 = 0.0f;
    float4* R = w.d.rows + (size_t)r * w.ROWS();
    if (mode == 0) {
      dn = lam.x; dt1 = lam.y; dt2 = lam.z;
      for (uint32_t k = 0; k < 3; k++) {
        float dd = k == 0 ? dn : (k == 1 ? dt1 : dt2);
        ta[0] += xyz(R[7 + k * 4]) * dd; ta[1] += xyz(R[8 + k * 4]) * dd;
        tb[0] += xyz(R[9 + k * 4]) * dd; tb[1] += xyz(R[10 + k * 4]) * dd;
      }
      f3 fra = ra, frb = rb;
      if (EXTRA) { fra = xyz(R[PO]); frb = xyz(R[PO + 1]); }
      f3 fn = n * dn, ft = t1 * dt1 + t2 * dt2;
      add_imp(a, -fn - ft, -cross(ra, fn) - cross(fra, ft));
      add_imp(b, fn + ft, cross(rb, fn) + cross(frb, ft));
    } else {
      f3 vrel = (tb[0] + cross(tb[1], rb)) - (ta[0] + cross(ta[1], ra));
      float jv = dot(n, vrel);
      float wgt = vpass ? 1.0f : R[5].w;
      float res = jv + r1.w;
      if (EXT && w.h(W_COMPL)) res += R[w.h(W_CO)].x * lam.x;
      float omega = w.hf(W_F_OMEGA);
      float nl = mx(lam.x + omega * (-res * r0.w * wgt - (1.0f - wgt) * lam.x), 0.0f);
      dn = nl - lam.x;
      lam.x = nl;
      ta[0] += xyz(R[7]) * dn; ta[1] += xyz(R[8]) * dn;
      tb[0] += xyz(R[9]) * dn; tb[1] += xyz(R[10]) * dn;
      add_imp(a, -n * dn, -cross(ra, n * dn));
      add_imp(b, n * dn, cross(rb, n * dn));
      f3 fra = ra, frb = rb;
      float lim = r2.w * nl, pooled = nl;
      f2 bias = mk2(0, 0);
      float role = 1.0f;
      if (EXTRA) {
        role = R[PO].w;
        if (PATCH && role != 1.0f) art_[w.h(W_A_ULOAD) + us] += dn;
        fra = xyz(R[PO]); frb = xyz(R[PO + 1]);
        if (PATCH && role == 2.0f) {
          float load = art_[w.h(W_A_ULOAD) + us];
          lim = r2.w * load;
          pooled = load;
          f3 e = xyz(R[PO + 2]);
          float beta = w.hf(W_F_ANCHOR_BETA), dt = w.hf(W_DT);
          if (!vpass) bias = mk2(beta * dot(t1, e) / dt, beta * dot(t2, e) / dt);
          float4 ex = R[PO + 2]; ex.w = lim; R[PO + 2] = ex;
        }
      }
      if (const uint32_t FM = EXT ? w.h(W_FMODE) : 0u) {
        // Coupled (normal, t1, t2) re-solve; overwrites all three impulses (wgsl.js:3081-3097).
        const f3 vn = (tb[0] + cross(tb[1], rb)) - (ta[0] + cross(ta[1], ra));
        const f3 vt = (tb[0] + cross(tb[1], frb)) - (ta[0] + cross(ta[1], fra));
        const uint32_t FMO = w.h(W_FMO);
        const f3 nlam = contact_block(FM, dot(n, vn), mk2(dot(t1, vt), dot(t2, vt)), R[FMO], R[FMO + 1], -r1.w, r2.w, xyz(lam));
        const f3 d = nlam - xyz(lam);
        if (PATCH && role != 1.0f) art_[w.h(W_A_ULOAD) + us] += d.x;
        ta[0] += xyz(R[7]) * d.x + xyz(R[11]) * d.y + xyz(R[15]) * d.z;
        ta[1] += xyz(R[8]) * d.x + xyz(R[12]) * d.y + xyz(R[16]) * d.z;
        tb[0] += xyz(R[9]) * d.x + xyz(R[13]) * d.y + xyz(R[17]) * d.z;
        tb[1] += xyz(R[10]) * d.x + xyz(R[14]) * d.y + xyz(R[18]) * d.z;
        const f3 fn_ = n * d.x, ft = t1 * d.y + t2 * d.z;
        add_imp(a, -fn_ - ft, -cross(ra, fn_) - cross(fra, ft));
        add_imp(b, fn_ + ft, cross(rb, fn_) + cross(frb, ft));
        lam = F4(nlam, lam.w);
      } else if (role != 0.0f) {
        f3 vr2 = (tb[0] + cross(tb[1], frb)) - (ta[0] + cross(ta[1], fra));
        float4 ki = R[5];
        float v1 = dot(t1, vr2) + bias.x, v2 = dot(t2, vr2) + bias.y;
        float kd = ki.x * ki.z - ki.y * ki.y;
        f2 lt = friction_pair_solve(ki.z / kd, -ki.y / kd, ki.x / kd, v1, v2, lam.y, lam.z, lim, omega);
        dt1 = lt.x - lam.y; dt2 = lt.y - lam.z;
        lam.y = lt.x; lam.z = lt.y;
        ta[0] += xyz(R[11]) * dt1 + xyz(R[15]) * dt2;
        ta[1] += xyz(R[12]) * dt1 + xyz(R[16]) * dt2;
        tb[0] += xyz(R[13]) * dt1 + xyz(R[17]) * dt2;
        tb[1] += xyz(R[14]) * dt1 + xyz(R[18]) * dt2;
        f3 f = t1 * dt1 + t2 * dt2;
        add_imp(a, -f, -cross(fra, f));
        add_imp(b, f, cross(frb, f));
      }
      bool ell = false;
      if (EXT && w.h(W_ANG_ART)) {
        const bool art_u = (a >= 0 && w.body_kind(a) >= 2) || (b >= 0 && w.body_kind(b) >= 2);
        ell = w.h(W_ANG) && art_u && (R[19].x > 0.0f || R[19].y > 0.0f);
      }
      if (ell) {
        // Articulated spin/rolling: after the sliding pair, Steel's block over sliding, spin and rolling (wgsl.js:3113-3141).
        float H5[25];
        uint32_t hidx = 0;
        const uint32_t EO = w.h(W_EO);
        for (uint32_t x = 0; x < 5; x++)
          for (uint32_t y2 = x; y2 < 5; y2++) {
            const float4 hv4 = R[EO + hidx / 4];
            const uint32_t c4 = hidx % 4;
            const float hv = c4 == 0 ? hv4.x : (c4 == 1 ? hv4.y : (c4 == 2 ? hv4.z : hv4.w));
            H5[x * 5 + y2] = hv; H5[y2 * 5 + x] = hv;
            hidx++;
          }
        const f3 vt = (tb[0] + cross(tb[1], frb)) - (ta[0] + cross(ta[1], fra));
        const f3 wr = tb[1] - ta[1];
        const float4 la0 = R[21];
        const float g5[5] = {dot(t1, vt), dot(t2, vt), dot(n, wr), dot(t1, wr), dot(t2, wr)};
        const float mu5[5] = {r2.w, r2.w, R[19].x, R[19].y, R[19].y};
        const float l5[5] = {lam.y, lam.z, la0.x, la0.y, la0.z};
        float nl5[5], d[5];
        ell_block(H5, g5, mu5, l5, pooled, nl5);
        for (uint32_t k = 0; k < 5; k++) d[k] = nl5[k] - l5[k];
        ta[0] += xyz(R[11]) * d[0] + xyz(R[15]) * d[1] + xyz(R[22]) * d[2] + xyz(R[26]) * d[3] + xyz(R[30]) * d[4];
        ta[1] += xyz(R[12]) * d[0] + xyz(R[16]) * d[1] + xyz(R[23]) * d[2] + xyz(R[27]) * d[3] + xyz(R[31]) * d[4];
        tb[0] += xyz(R[13]) * d[0] + xyz(R[17]) * d[1] + xyz(R[24]) * d[2] + xyz(R[28]) * d[3] + xyz(R[32]) * d[4];
        tb[1] += xyz(R[14]) * d[0] + xyz(R[18]) * d[1] + xyz(R[25]) * d[2] + xyz(R[29]) * d[3] + xyz(R[33]) * d[4];
        const f3 f = t1 * d[0] + t2 * d[1];
        const f3 tq = n * d[2] + t1 * d[3] + t2 * d[4];
        add_imp(a, -f, -cross(fra, f) - tq);
        add_imp(b, f, cross(frb, f) + tq);
        lam.y = nl5[0]; lam.z = nl5[1];
        R[21] = make_float4(nl5[2], nl5[3], nl5[4], 0.0f);
      }
      R[6] = lam;
      if (EXT && w.h(W_ANG)) {
        // Spin within mu_t N, rolling within a disk of mu_r N; under friction patches only the region's first
        // anchor carries them, with the pooled load (wgsl.js:3144-3173).
        const float4 g = R[19];
        const bool aon = !ell && (!PATCH || role == 1.0f || (role == 2.0f && R[PO + 1].w == 0.0f));
        if (aon && (g.x > 0.0f || g.y > 0.0f)) {
          const float4 la = R[21];
          const f3 wrel = tb[1] - ta[1];
          const float ns = clampf_(la.x - dot(n, wrel) * g.z, -g.x * pooled, g.x * pooled);
          const float ds = ns - la.x;
          ta[0] += xyz(R[22]) * ds; ta[1] += xyz(R[23]) * ds;
          tb[0] += xyz(R[24]) * ds; tb[1] += xyz(R[25]) * ds;
          const f3 w2 = tb[1] - ta[1];
          const float4 ri = R[20];
          const float u1 = dot(t1, w2), u2 = dot(t2, w2);
          f2 lr = mk2(la.y, la.z) - mk2(ri.x * u1 + ri.y * u2, ri.y * u1 + ri.z * u2);
          const float rl = g.y * pooled;
          if (len(lr) > rl) lr = lr * (rl / mx(len(lr), 1e-12f));
          const float d1 = lr.x - la.y, d2 = lr.y - la.z;
          ta[0] += xyz(R[26]) * d1 + xyz(R[30]) * d2; ta[1] += xyz(R[27]) * d1 + xyz(R[31]) * d2;
          tb[0] += xyz(R[28]) * d1 + xyz(R[32]) * d2; tb[1] += xyz(R[29]) * d1 + xyz(R[33]) * d2;
          R[21] = make_float4(ns, lr.x, lr.y, 0.0f);
          const f3 tq = n * ds + t1 * d1 + t2 * d2;
          add_imp(a, mk3(0, 0, 0), -tq);
          add_imp(b, mk3(0, 0, 0), tq);
        }
      }
      if (EXT && w.h(W_TORS) && r == w.d.sched[w.S.UTP + us]) unit_torsion(us, a, b, ta, tb);
    }
    if (a >= 0) { tw[(uint32_t)a * 2] = F4(ta[0], 0.0f); tw[(uint32_t)a * 2 + 1] = F4(ta[1], 0.0f); }
    if (b >= 0) { tw[(uint32_t)b * 2] = F4(tb[0], 0.0f); tw[(uint32_t)b * 2 + 1] = F4(tb[1], 0.0f); }
  }
  // Contact torsion after the unit's last colored row (wgsl.js:3185-3228): sliding and spin share the members'
  // Coulomb budget mu * N; when sliding exceeds it (later normals lowered the load) the members' tangent pairs are
  // scaled onto it, then the spin row is relaxed within r * (budget - sliding).  The unit's rows were all solved
  // in earlier colors, by this thread or before the last barrier.
  DEVN void unit_torsion(uint32_t us, int a, int b, f3 ta[2], f3 tb[2]) const {
    const uint32_t* sched = w.d.sched;
    const uint32_t ROWS = w.ROWS(), TO = w.h(W_TO), PO = w.PO(), MAXC = w.MAXC();
    const bool EXTRA = w.h(W_EXTRA) != 0;
    const uint32_t s0 = sched[w.S.USTART + us], s1 = min(s0 + sched[w.S.USIZE + us], MAXC);
    float4* L = w.d.rows + (size_t)s0 * ROWS;
    if (!(L[TO].w > 0.0f)) return;
    const float mu = L[TO].w, omega = w.hf(W_F_OMEGA);
    float load = 0.0f, slide = 0.0f;
    for (uint32_t m = s0; m < s1; m++) {
      const float4* M = w.d.rows + (size_t)m * ROWS;
      if (M[TO + 4].w < 0.5f) continue;
      const float4 lm = M[6];
      load += mx(lm.x, 0.0f);
      if (!EXTRA || M[PO].w != 0.0f) slide += len(mk2(lm.y, lm.z));
    }
    const float budget = mu * load;
    if (slide > budget && slide > 0.0f) {
      const float scale = budget / slide;
      for (uint32_t m = s0; m < s1; m++) {
        float4* M = w.d.rows + (size_t)m * ROWS;
        if (M[TO + 4].w < 0.5f || (EXTRA && M[PO].w == 0.0f)) continue;
        const float4 lm = M[6];
        const float d1 = lm.y * (scale - 1.0f), d2 = lm.z * (scale - 1.0f);
        ta[0] += xyz(M[11]) * d1 + xyz(M[15]) * d2; ta[1] += xyz(M[12]) * d1 + xyz(M[16]) * d2;
        tb[0] += xyz(M[13]) * d1 + xyz(M[17]) * d2; tb[1] += xyz(M[14]) * d1 + xyz(M[18]) * d2;
        f3 fa = xyz(M[3]), fb = xyz(M[4]);
        if (EXTRA) { fa = xyz(M[PO]); fb = xyz(M[PO + 1]); }
        const f3 f = xyz(M[1]) * d1 + xyz(M[2]) * d2;
        add_imp(a, -f, -cross(fa, f));
        add_imp(b, f, cross(fb, f));
        M[6] = make_float4(lm.x, lm.y * scale, lm.z * scale, lm.w);
      }
      slide = budget;
    }
    const float bound = L[TO + 3].w * mx(budget - slide, 0.0f);
    const f3 ng = xyz(L[TO + 4]);
    const float ls = L[TO + 2].w;
    const float nls = clampf_(ls - omega * dot(ng, tb[1] - ta[1]) * L[TO + 1].w, -bound, bound);
    const float ds = nls - ls;
    if (ds != 0.0f) {
      ta[0] += xyz(L[TO]) * ds; ta[1] += xyz(L[TO + 1]) * ds;
      tb[0] += xyz(L[TO + 2]) * ds; tb[1] += xyz(L[TO + 3]) * ds;
      float4 l2 = L[TO + 2];
      l2.w = nls;
      L[TO + 2] = l2;
      add_imp(a, mk3(0, 0, 0), -ng * ds);
      add_imp(b, mk3(0, 0, 0), ng * ds);
    }
  }
  DEV void set_link_twist(uint32_t l, SVd v) const {
    uint32_t b = mbody(l);
    tw[b * 2] = F4(v.l, 0.0f); tw[b * 2 + 1] = F4(v.a, 0.0f);
  }
  // Orders LDS and memory between the lanes of one wave (the tree passes below run a wave per articulation).
  static DEV void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  }
  // Tree passes over an articulation, one wave, lane k holding link ai.x + k, depth levels in turn.  Each
  // link's arithmetic is the serial pass's; where the serial backward pass adds children into a parent in
  // decreasing link order, the parent here sums its children's contributions in that same order, so the
  // results are bit for bit those of the serial passes (wgsl.js:3240-3299).
  DEVN void art_twists(uint32_t a, uint32_t lane) const {  // wgsl.js:3240-3251
    uint4 ai = w.art_info(a);
    const uint32_t MP = w.h(W_MAXPATH);
    const bool has = lane < ai.y;
    const uint32_t l = ai.x + lane, dl = has ? ltree[l].x : 0xffffffffu;
    wave_sync();
    for (uint32_t d = 0; d < MP; d++) {
      if (dl == d) {
        int par = mpar(l);
        SVd v = sv0();
        if (par >= 0) v = mshift(msv(mV((uint32_t)par)), mr(l));
        for (uint32_t r = 0; r < mnd(l); r++) v = sadd(v, sscale(mS(l, r), laqd[mdof(l) + r]));
        mst(mV(l), v);
        set_link_twist(l, v);
      }
      wave_sync();
    }
  }
  DEVN void art_apply(uint32_t a, uint32_t lane) const {  // wgsl.js:3254-3299
    uint4 ai = w.art_info(a);
    const uint32_t MP = w.h(W_MAXPATH);
    const bool has = lane < ai.y;
    const uint32_t l = ai.x + lane, dl = has ? ltree[l].x : 0xffffffffu;
    wave_sync();
    for (int d = (int)MP - 1; d >= 0; d--) {
      if (dl == (uint32_t)d) {
        uint32_t n = mnd(l);
        float *UU = mUU(l), *TAU = mTAU(l);
        SVd p = msv(mP(l));
        float u[6];
        for (uint32_t r = 0; r < n; r++) { u[r] = TAU[r] - sdot(mS(l, r), p); UU[r] = u[r]; }
        if (mpar(l) >= 0) {
          SVd pp = p;
          for (uint32_t r = 0; r < n; r++) {
            float du = 0.0f;
            for (uint32_t c = 0; c < n; c++) du += mDI(l, r, c) * u[c];
            pp = sadd(pp, sscale(mU(l, r), du));
          }
          mst(lcon + l * 8, fshift(pp, mr(l)));
        }
      }
      wave_sync();
      if (d > 0 && dl == (uint32_t)(d - 1)) {
        uint32_t c = ltree[l].y;
        if (c != 0xffffffffu) {
          SVd P = msv(mP(l));
          for (; c != 0xffffffffu; c = ltree[c].z) P = sadd(P, msv(lcon + c * 8));
          mst(mP(l), P);
        }
      }
      wave_sync();
    }
    for (uint32_t d = 0; d < MP; d++) {
      if (dl == d) {
        int par = mpar(l);
        SVd dv = sv0();
        if (par >= 0) dv = mshift(msv(mDV((uint32_t)par)), mr(l));
        uint32_t n = mnd(l);
        const float* UU = mUU(l);
        float dq[6];
        for (uint32_t r = 0; r < n; r++) {
          dq[r] = 0.0f;
          for (uint32_t c = 0; c < n; c++) dq[r] += mDI(l, r, c) * (UU[c] - sdot(mU(l, c), dv));
        }
        for (uint32_t r = 0; r < n; r++) { dv = sadd(dv, sscale(mS(l, r), dq[r])); laqd[mdof(l) + r] += dq[r]; }
        mst(mDV(l), dv);
        SVd v = sadd(msv(mV(l)), dv);
        mst(mV(l), v);
        set_link_twist(l, v);
        mst(mP(l), sv0());
        float* TAU = mTAU(l);
        for (uint32_t r = 0; r < n; r++) TAU[r] = 0.0f;
      }
      wave_sync();
    }
  }
  // subgroupAdd of the reference's 32-lane path (wgsl.js:3305), order fixed as an xor butterfly over strides
  // 16, 8, 4, 2, 1 within each 32-lane half of the wave; every lane ends with the same bits (x + y == y + x).
  // All strides stay in VALU: v_permlane16_swap exchanges the 16-lane rows of each half (the sum of the two
  // swapped registers is x + partner in every lane), DPP row_ror:8 gives lane ^ 8, row_half_mirror then a
  // reversed quad gives lane ^ 4 (7 - (i ^ 3) within 8 lanes), and quad permutes give lane ^ 2 and lane ^ 1.
  static DEV float lane_sum32(float x) {
    const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(x), __float_as_uint(x), false, false);
    x = __uint_as_float(r[0]) + __uint_as_float(r[1]);
    x = x + __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0x128, 0xF, 0xF, false));
    x = x + __int_as_float(__builtin_amdgcn_mov_dpp(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0x141, 0xF, 0xF, false), 0x1B, 0xF, 0xF, false));
    x = x + __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0x4E, 0xF, 0xF, false));
    x = x + __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0xB1, 0xF, 0xF, false));
    return x;
  }
  // A dense row's uniform fields and this lane's entries of its three J and Y rows (dofs lane and lane + 32).
  struct DenseRec {
    float bil, mu, rhs, meff, k6, k7, k8, lx, ly, lz, wgt;
    float j0[6], j1[6];  // J rows 0-2 then Y rows 0-2
  };
  // Dense same-articulation rows, one wave per articulation (wgsl.js:3306-3364, subgroup form): lane l < 32
  // owns dofs l and l + 32 and keeps their joint velocities in registers for the whole pass; rows run in
  // order (Gauss-Seidel).  The wave finds the articulation's rows 64 at a time (a ballot over their owners) and
  // loads each row's record while the row before it is solved.  Called by all 64 lanes of the wave with the same
  // (a, mode, nd).
  DEVN bool art_dense(uint32_t a, uint32_t mode, uint32_t nd, uint32_t lane) const {
    const uint4 ai = w.art_info(a);
    const AS_GLOBAL float* ar = (const AS_GLOBAL float*)w.d.art;
    AS_GLOBAL float* aw = (AS_GLOBAL float*)w.d.art;
    float* aqd = laqd;
    const uint32_t NDA = w.NDA(), DHDR = w.h(W_DHDR), DPS = w.h(W_DPS), AD = w.h(W_A_DENSE);
    const float omega = w.hf(W_F_OMEGA);
    const uint32_t n = ai.w;
    const bool own0 = lane < 32 && lane < n, own1 = lane < 32 && lane + 32 < n;
    float q0 = own0 ? aqd[ai.z + lane] : 0.0f, q1 = own1 ? aqd[ai.z + lane + 32] : 0.0f;
    bool touched = false;
    auto fetch = [&](uint32_t i, DenseRec& r) {
      const uint32_t o = AD + i * DPS, jo = o + DHDR, yo = o + DHDR + 3 * NDA;
      r.bil = ar[o + 21]; r.mu = ar[o + 3]; r.rhs = ar[o + 4]; r.meff = ar[o + 5];
      r.k6 = ar[o + 6]; r.k7 = ar[o + 7]; r.k8 = ar[o + 8];
      r.lx = ar[o + 9]; r.ly = ar[o + 10]; r.lz = ar[o + 11];
      r.wgt = vpass ? 1.0f : ar[o + 24];
      for (uint32_t k = 0; k < 3; k++) {
        r.j0[k] = own0 ? ar[jo + k * NDA + lane] : 0.0f;
        r.j0[3 + k] = own0 ? ar[yo + k * NDA + lane] : 0.0f;
        r.j1[k] = own1 ? ar[jo + k * NDA + lane + 32] : 0.0f;
        r.j1[3 + k] = own1 ? ar[yo + k * NDA + lane + 32] : 0.0f;
      }
    };
    auto solve = [&](uint32_t i, const DenseRec& r, const FricConst& fl, uint32_t base) {
      const uint32_t o = AD + i * DPS;
      if (r.bil > 0.5f) {
        // Bilateral rows (mimics): unclamped, one at a time; no warm impulse (wgsl.js:3319-3332).
        if (mode == 0) return;
        float nlam[3] = {r.lx, r.ly, r.lz};
        const float rhs3[3] = {r.rhs, ar[o + 22], ar[o + 23]}, m3_[3] = {r.meff, r.k6, r.k7};
        const uint32_t nb = tu(r.bil);
#pragma unroll
        for (uint32_t k = 0; k < 3; k++) {
          if (k >= nb) break;
          float part = 0.0f;
          if (own0) part += r.j0[k] * q0;
          if (own1) part += r.j1[k] * q1;
          const float jv = lane_sum32(part);
          const float rhs = vpass ? 0.0f : rhs3[k];
          const float dl = -omega * (jv + rhs) * m3_[k];
          nlam[k] += dl;
          if (own0) q0 += r.j0[3 + k] * dl;
          if (own1) q1 += r.j1[3 + k] * dl;
        }
        if (lane == 0) { aw[o + 9] = nlam[0]; aw[o + 10] = nlam[1]; aw[o + 11] = nlam[2]; }
        return;
      }
      if (mode == 0) {
        if (own0) q0 += r.j0[3] * r.lx + r.j0[4] * r.ly + r.j0[5] * r.lz;
        if (own1) q1 += r.j1[3] * r.lx + r.j1[4] * r.ly + r.j1[5] * r.lz;
        return;
      }
      float part = 0.0f;
      if (own0) part += r.j0[0] * q0;
      if (own1) part += r.j1[0] * q1;
      const float jv = lane_sum32(part);
      const float nl = mx(r.lx + omega * (-(jv + r.rhs) * r.meff * r.wgt - (1.0f - r.wgt) * r.lx), 0.0f);
      const float dn = nl - r.lx;
      float p1 = 0.0f, p2 = 0.0f;
      if (own0) { q0 += r.j0[3] * dn; p1 += r.j0[1] * q0; p2 += r.j0[2] * q0; }
      if (own1) { q1 += r.j1[3] * dn; p1 += r.j1[1] * q1; p2 += r.j1[2] * q1; }
      const float v1 = lane_sum32(p1), v2 = lane_sum32(p2);
      const float lim = r.mu * nl;
      FricConst fk;  // the row's constants, computed by lane i - base of the chunk
      const int src = (int)(i - base);
      fk.scale = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(fl.scale), src));
      fk.largest = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(fl.largest), src));
      fk.smallest = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(fl.smallest), src));
      fk.vx = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(fl.vx), src));
      fk.vy = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(fl.vy), src));
      f2 lt = friction_pair_apply(fk, v1, v2, r.ly, r.lz, lim, omega);
      const float d1 = lt.x - r.ly, d2 = lt.y - r.lz;
      if (own0) q0 += r.j0[4] * d1 + r.j0[5] * d2;
      if (own1) q1 += r.j1[4] * d1 + r.j1[5] * d2;
      if (lane == 0) { aw[o + 9] = nl; aw[o + 10] = lt.x; aw[o + 11] = lt.y; }
    };
    DenseRec cur, nxt;
    for (uint32_t base = 0; base < nd; base += 64) {
      const uint32_t ii = base + lane;
      const uint32_t owner = ii < nd ? tu(ar[AD + ii * DPS]) : 0xffffffffu;
      uint64_t m = __ballot(owner == a);
      if (m == 0) continue;
      touched = true;
      // Each lane computes the friction constants of its row of the chunk (they depend on the row's K only),
      // the same expressions the serial row solve would evaluate, once per pass instead of once per row.
      FricConst fl = {1.0f, 0.0f, 0.0f, 1.0f, 0.0f};
      if (owner == a) {
        const uint32_t o = AD + ii * DPS;
        const float k6 = ar[o + 6], k7 = ar[o + 7], k8 = ar[o + 8];
        const float kd = k6 * k8 - k7 * k7;
        fl = friction_const(k8 / kd, -k7 / kd, k6 / kd);
      }
      uint32_t i = base + (uint32_t)__builtin_ctzll(m);
      m &= m - 1;
      fetch(i, cur);
      for (;;) {
        const bool more = m != 0;
        const uint32_t inext = more ? base + (uint32_t)__builtin_ctzll(m) : 0u;
        if (more) { m &= m - 1; fetch(inext, nxt); }
        solve(i, cur, fl, base);
        if (!more) break;
        i = inext;
        cur = nxt;
      }
    }
    if (touched) {
      if (own0) aqd[ai.z + lane] = q0;
      if (own1) aqd[ai.z + lane + 32] = q1;
    }
    return touched;
  }
  DEVN void art_contact_impulses(uint32_t a, uint32_t lane) const {  // wgsl.js:3380-3394
    uint4 ai = w.art_info(a);
    bool hit = false;
    if (lane < ai.y) {
      uint32_t l = ai.x + lane, b = mbody(l);
      f3 f = xyz(imp[b * 2]), t = xyz(imp[b * 2 + 1]);
      if (vne0(f) || vne0(t)) {
        hit = true;
        mst(mP(l), mksv(-t, -f));
        imp[b * 2] = make_float4(0, 0, 0, 0); imp[b * 2 + 1] = make_float4(0, 0, 0, 0);
      }
    }
    if (__ballot(hit) != 0) art_apply(a, lane);
  }
  // Joint-limit rows (wgsl.js:3404-3461), one wave: every lane computes the same row update (all inputs are
  // uniform) and writes the same values; lane 0 sets the link's tau before the wave's art_apply.
  DEVN void art_propagate_rows(uint32_t a, uint32_t lane) const {
    uint4 ai = w.art_info(a);
    AS_GLOBAL float* art_ = (AS_GLOBAL float*)w.d.art;
    uint32_t n = tu(art_[w.h(W_A_RLN) + a]);
    float dt = w.hf(W_DT), beta = w.hf(W_BETA), maxd = w.hf(W_MAX_DEPEN), omega = w.hf(W_F_OMEGA);
    for (uint32_t i = 0; i < n; i++) {
      uint32_t dd = tu(art_[w.h(W_A_RL) + ai.z + i]);
      float4 lim = w.dof_limit(dd);
      uint32_t l = tu(lim.w), r = dd - w.link_dof(l);
      float hinv = mx(art_[w.h(W_A_HINV) + dd], 1e-12f);
      const float4 dr = w.dof_drive(dd);
      if (w.h(W_DRIVE_ROWS) && (dr.y > 0.0f || dr.z > 0.0f) && !(vpass && !w.h(W_VEL_DRIVE_ACTIVE))) {
        // drive_mode rows (wgsl.js:3425-3438): a PhysX-style drive row, cold-started each step, not relaxed.
        const float q = art_[w.h(W_A_Q) + w.link_qs(l) + r];
        const float pd = vpass ? dt * art_[w.h(W_A_SNAP) + w.NB() * 8 + dd] : 0.0f;
        const float ca = dt * (dt * dr.y + dr.z);
        const float x = 1.0f / (1.0f + ca * hinv);
        const float bias = x * dt * dr.z * art_[w.h(W_A_TGTV) + dd] + dr.y * x * dt * (art_[w.h(W_A_TGT) + dd] - q - pd);
        const float cap = dr.w > 0.0f ? dr.w * dt : 1e20f;
        const float lam = art_[w.h(W_A_DRV) + dd];
        const float nl = clampf_(lam * (1.0f - x) - x * ca * laqd[dd] + bias, -cap, cap);
        if (nl != lam) {
          art_[w.h(W_A_DRV) + dd] = nl;
          if (lane == 0) mTAU(l)[r] = nl - lam;
          art_apply(a, lane);
        }
      }
      // Joint friction: a velocity row with |impulse| <= friction * dt (plus the load term).
      float fr = w.dof_passive(dd).y;
      if (w.h(W_JFL)) fr += art_[w.h(W_A_JFL) + dd];
      if (fr > 0.0f) {
        float lam = art_[w.h(W_A_FRIC) + dd];
        float nl = clampf_(lam - omega * laqd[dd] / hinv, -fr * dt, fr * dt);
        if (nl != lam) {
          art_[w.h(W_A_FRIC) + dd] = nl;
          if (lane == 0) mTAU(l)[r] = nl - lam;
          art_apply(a, lane);
        }
      }
      if (art_[w.h(W_A_LG) + dd] > 0.5f) {
        float q = art_[w.h(W_A_Q) + w.link_qs(l) + r];
        for (uint32_t side = 0; side < 2; side++) {
          float sg = side == 0 ? 1.0f : -1.0f;
          float gap = side == 0 ? q - lim.x : lim.y - q;
          float lam = art_[w.h(W_A_LIM) + 2 * dd + side];
          float rhs = gap / dt;
          if (gap < 0.0f) rhs = vpass ? 0.0f : mx(beta * gap / dt, -maxd);
          float nl = mx(lam - omega * (sg * laqd[dd] + rhs) / hinv, 0.0f);
          if (nl == lam) continue;
          art_[w.h(W_A_LIM) + 2 * dd + side] = nl;
          if (lane == 0) mTAU(l)[r] = sg * (nl - lam);
          art_apply(a, lane);
        }
      }
    }
  }
  // Velocity limits, last in a sweep: a stateless push of each active scalar dof back onto |qd| <= limit
  // through the articulation (wgsl.js:3413-3419, 3462-3467); a wave per articulation, as the limit rows.
  DEVN void art_velocity_limits(uint32_t a, uint32_t lane) const {
    AS_GLOBAL float* art_ = (AS_GLOBAL float*)w.d.art;
    if (!(w.h(W_VLIM_ON) && art_[w.h(W_A_AVL) + a] > 0.5f)) return;
    uint4 ai = w.art_info(a);
    for (uint32_t k = 0; k < ai.y; k++) {
      uint32_t l = ai.x + k, t = w.link_type(l);
      if (t != JT_REVOLUTE && t != JT_PRISMATIC && t != JT_D6) continue;
      for (uint32_t r = 0; r < w.link_ndof(l); r++) {
        uint32_t dd = w.link_dof(l) + r;
        float hinv = mx(art_[w.h(W_A_HINV) + dd], 1e-12f);
        float4 pv = w.dof_passive(dd);
        if (pv.z >= 0.0f && art_[w.h(W_A_VLG) + dd] > 0.5f) {
          float over = laqd[dd] - clampf_(laqd[dd], -pv.z, pv.z);
          if (over != 0.0f) {
            if (lane == 0) mTAU(l)[r] = -over / hinv;
            art_apply(a, lane);
          }
        }
      }
    }
  }
  // Free-body velocity limits: a per-axis box on v (clamped) and w (pushed back through the body's inverse
  // inertia), stateless (wgsl.js:3586-3610).
  DEV void body_velocity_limits(uint32_t b) const {
    if (w.body_kind((int)b) != 1 || (w.body_flags((int)b) & 1) != 0) return;
    const AS_GLOBAL float* art_ = (const AS_GLOBAL float*)w.d.art;
    q4 lm = w.body_limits((int)b);
    const uint32_t g = w.h(W_A_VLG) + w.ND() + b * 6;
    if (lm.y > 0.0f) {
      f3 v = xyz(tw[b * 2]);
      for (int k = 0; k < 3; k++) if (art_[g + k] > 0.5f) v = with_comp(v, k, clampf_(comp(v, k), -lm.y, lm.y));
      tw[b * 2] = F4(v, 0.0f);
    }
    if (lm.z > 0.0f) {
      f3 wv = xyz(tw[b * 2 + 1]);
      for (int k = 0; k < 3; k++) {
        float over = comp(wv, k) - clampf_(comp(wv, k), -lm.z, lm.z);
        if (over == 0.0f || art_[g + 3 + k] < 0.5f) continue;
        f3 e = with_comp(mk3(0, 0, 0), k, 1.0f);
        m3 R = qmat(w.pose_q((int)b));
        f3 iw = mul(R, mul(w.body_inv_inertia((int)b), mul(tr(R), e)));
        wv = wv - iw * (over / comp(iw, k));
      }
      tw[b * 2 + 1] = F4(wv, 0.0f);
    }
  }
  DEVN f3 gyroscopic_velocity(const m3& inertia, f3 u0, float h0) const {  // wgsl.js:2543-2562
    auto inverse3 = [](const m3& m) {
      f3 a = m.c[0], b = m.c[1], c = m.c[2];
      f3 r0 = cross(b, c), r1 = cross(c, a), r2 = cross(a, b);
      float det = dot(a, r0);
      m3 t; t.c[0] = r0; t.c[1] = r1; t.c[2] = r2;
      return mscale(tr(t), 1.0f / det);
    };
    m3 a = mscale(inertia, 1.0f / mx(inertia.c[0].x, mx(inertia.c[1].y, inertia.c[2].z)));
    m3 inv = inverse3(a);
    float speed = sqrtf(mx(dot(u0, mul(a, u0)) * (inv.c[0].x + inv.c[1].y + inv.c[2].z), 0.0f));
    uint32_t n = (uint32_t)clampf_(ceilf(2.0f * fabsf(h0) * speed), 1.0f, 32.0f);
    float h = h0 / (float)n;
    f3 wv = u0;
    for (uint32_t m = 0; m < n; m++) {
      f3 u = wv;
      float energy = dot(u, mul(a, u));
      for (uint32_t it = 0; it < 8; it++) {
        m3 Sm = mscale(0.5f * h, skew3(mul(a, 0.5f * (u + wv))));
        f3 c = mul(inverse3(msub(a, Sm)), mul(madd(a, Sm), u));
        f3 dd = c - wv;
        wv = c;
        if (dot(dd, mul(a, dd)) <= 1e-12f * energy) break;
      }
    }
    return wv;
  }
  DEV f3 body_motion(int b, f3 x) const {  // wgsl.js:3472-3489
    if (b < 0) return mk3(0, 0, 0);
    uint32_t PP = w.h(W_P_PREV), PB = w.h(W_P_BODY);
    f3 p0 = xyz(w.d.poses[PP + (uint32_t)b * 2]);
    q4 q0 = Q(w.d.poses[PP + (uint32_t)b * 2 + 1]);
    f3 p1 = xyz(w.d.poses[PB + (uint32_t)b * 2]);
    q4 q1 = Q(w.d.poses[PB + (uint32_t)b * 2 + 1]);
    q4 r = qmul(q1, qconj(q0));
    if (r.w < 0.0f) r = mk4(-r.x, -r.y, -r.z, -r.w);
    float sn = len(xyz(r));
    f3 ang = 2.0f * xyz(r);
    f3 trv = p1 - p0;
    f3 lin = trv;
    if (sn > 1e-6f) {
      float half = atan2_c(sn, r.w);
      f3 axis = xyz(r) / sn;
      ang = (2.0f * half) * axis;
      float sc = half * r.w / sn;
      lin = sc * trv + ((1.0f - sc) * dot(axis, trv)) * axis;
    }
    return lin + cross(ang, x - 0.5f * (p0 + p1));
  }
};

// Dense rows of every articulation, one wave each in rounds, then the twists of the articulations they moved
// (wgsl.js:3366-3377 dense_phase).
// ---- The default-path row solve with explicit address spaces.  Solver::solve_unit<false> reached its pointers
// through the Solver object (on the stack, since the object's address escapes to out-of-line members) and used
// flat memory operations, so every access waited for every outstanding load; here the row pointers are global,
// the twists and impulses LDS (worlds whose twists fit) or global, all carried in registers, and the row's
// records are loaded before anything is stored.  Same arithmetic, in the same order, as solve_unit<false>.
typedef float v4f __attribute__((ext_vector_type(4)));
typedef uint32_t v4u __attribute__((ext_vector_type(4)));
template <bool TWL> struct TwSpace { typedef AS_GLOBAL v4f T; };
template <> struct TwSpace<true> { typedef AS_LDS v4f T; };
DEV float4 to4(v4f v) { return make_float4(v.x, v.y, v.z, v.w); }
DEV v4f tov(float4 v) { v4f r = {v.x, v.y, v.z, v.w}; return r; }
DEV f3 xyzv(v4f v) { return mk3(v.x, v.y, v.z); }

template <bool TWL>
struct RowCtx {
  const AS_GLOBAL v4f* rows;
  typename TwSpace<TWL>::T* tw;
  typename TwSpace<TWL>::T* imp;
  AS_GLOBAL float* uload;   // art + A_ULOAD
  const AS_GLOBAL v4u* slots;
  uint32_t ROWS, PO;
  float omega, beta_anchor, dt;
  bool extra, patch, vpass;
};

template <bool TWL>
DEV RowCtx<TWL> row_ctx(const Solver& sv) {
  const W& w = sv.w;
  RowCtx<TWL> c;
  c.rows = (const AS_GLOBAL v4f*)w.d.rows;
  c.tw = (typename TwSpace<TWL>::T*)sv.tw;
  c.imp = (typename TwSpace<TWL>::T*)sv.imp;
  c.uload = (AS_GLOBAL float*)(w.d.art + w.h(W_A_ULOAD));
  c.slots = (const AS_GLOBAL v4u*)(w.d.sched + w.S.SLOTT);
  c.ROWS = w.ROWS();
  c.PO = w.PO();
  c.omega = w.hf(W_F_OMEGA);
  c.beta_anchor = w.hf(W_F_ANCHOR_BETA);
  c.dt = w.hf(W_DT);
  c.extra = w.h(W_EXTRA) != 0;
  c.patch = w.h(W_PATCH) != 0;
  c.vpass = sv.vpass;
  return c;
}

// Accumulated link impulses.  Free bodies' entries are written too: nothing reads them (only link bodies' are
// propagated), which saves the body-kind lookup.
template <bool TWL>
DEV void row_add_imp(const RowCtx<TWL>& c, int b, f3 f, f3 t) {
  if (b < 0) return;
  const v4f i0 = c.imp[(uint32_t)b * 2], i1 = c.imp[(uint32_t)b * 2 + 1];
  c.imp[(uint32_t)b * 2] = tov(make_float4(i0.x + f.x, i0.y + f.y, i0.z + f.z, i0.w + 0.0f));
  c.imp[(uint32_t)b * 2 + 1] = tov(make_float4(i1.x + t.x, i1.y + t.y, i1.z + t.z, i1.w + 0.0f));
}

template <bool TWL>
DEV void solve_row(const RowCtx<TWL>& c, uint32_t us, uint32_t r, int a, int b, uint32_t mode) {  // wgsl.js:3033-3231
  const AS_GLOBAL v4f* R = c.rows + (size_t)r * c.ROWS;
  const uint32_t PO = c.PO;
  // Every record the row reads, first.
  const float4 r0 = to4(R[0]), r1 = to4(R[1]), r2 = to4(R[2]), r5 = to4(R[5]);
  const f3 ra = xyzv(R[3]), rb = xyzv(R[4]);
  float4 lam = to4(R[6]);
  const f3 na0 = xyzv(R[7]), na1 = xyzv(R[8]), nb0 = xyzv(R[9]), nb1 = xyzv(R[10]);
  const f3 t1a0 = xyzv(R[11]), t1a1 = xyzv(R[12]), t1b0 = xyzv(R[13]), t1b1 = xyzv(R[14]);
  const f3 t2a0 = xyzv(R[15]), t2a1 = xyzv(R[16]), t2b0 = xyzv(R[17]), t2b1 = xyzv(R[18]);
  float4 x0 = make_float4(0, 0, 0, 0), x1 = x0, x2 = x0;
  if (c.extra) { x0 = to4(R[PO]); x1 = to4(R[PO + 1]); x2 = to4(R[PO + 2]); }
  const f3 n = xyz(r0), t1 = xyz(r1), t2 = xyz(r2);
  f3 ta[2], tb[2];
  if (a >= 0) { ta[0] = xyzv(c.tw[(uint32_t)a * 2]); ta[1] = xyzv(c.tw[(uint32_t)a * 2 + 1]); } else { ta[0] = mk3(0, 0, 0); ta[1] = ta[0]; }
  if (b >= 0) { tb[0] = xyzv(c.tw[(uint32_t)b * 2]); tb[1] = xyzv(c.tw[(uint32_t)b * 2 + 1]); } else { tb[0] = mk3(0, 0, 0); tb[1] = tb[0]; }
  const f3 fra = c.extra ? xyz(x0) : ra, frb = c.extra ? xyz(x1) : rb;
  AS_GLOBAL v4f* Rw = (AS_GLOBAL v4f*)R;
  if (mode == 0) {
    const float dn = lam.x, dt1 = lam.y, dt2 = lam.z;
    ta[0] += na0 * dn; ta[1] += na1 * dn; tb[0] += nb0 * dn; tb[1] += nb1 * dn;
    ta[0] += t1a0 * dt1; ta[1] += t1a1 * dt1; tb[0] += t1b0 * dt1; tb[1] += t1b1 * dt1;
    ta[0] += t2a0 * dt2; ta[1] += t2a1 * dt2; tb[0] += t2b0 * dt2; tb[1] += t2b1 * dt2;
    const f3 fn = n * dn, ft = t1 * dt1 + t2 * dt2;
    row_add_imp(c, a, -fn - ft, -cross(ra, fn) - cross(fra, ft));
    row_add_imp(c, b, fn + ft, cross(rb, fn) + cross(frb, ft));
  } else {
    const bool anchor = c.extra && c.patch && x0.w == 2.0f;
    const float uload0 = (c.extra && c.patch && x0.w != 1.0f) ? c.uload[us] : 0.0f;
    const f3 vrel = (tb[0] + cross(tb[1], rb)) - (ta[0] + cross(ta[1], ra));
    const float jv = dot(n, vrel);
    const float wgt = c.vpass ? 1.0f : r5.w;
    const float res = jv + r1.w;
    const float nl = mx(lam.x + c.omega * (-res * r0.w * wgt - (1.0f - wgt) * lam.x), 0.0f);
    const float dn = nl - lam.x;
    lam.x = nl;
    ta[0] += na0 * dn; ta[1] += na1 * dn;
    tb[0] += nb0 * dn; tb[1] += nb1 * dn;
    row_add_imp(c, a, -n * dn, -cross(ra, n * dn));
    row_add_imp(c, b, n * dn, cross(rb, n * dn));
    float lim = r2.w * nl;
    f2 bias = mk2(0, 0);
    const float role = c.extra ? x0.w : 1.0f;
    if (c.extra && c.patch && role != 1.0f) {
      const float load = uload0 + dn;
      c.uload[us] = load;
      if (anchor) {
        lim = r2.w * load;
        const f3 e = xyz(x2);
        if (!c.vpass) bias = mk2(c.beta_anchor * dot(t1, e) / c.dt, c.beta_anchor * dot(t2, e) / c.dt);
        Rw[PO + 2] = tov(make_float4(x2.x, x2.y, x2.z, lim));
      }
    }
    if (role != 0.0f) {
      const f3 vr2 = (tb[0] + cross(tb[1], frb)) - (ta[0] + cross(ta[1], fra));
      const float4 ki = r5;
      const float v1 = dot(t1, vr2) + bias.x, v2 = dot(t2, vr2) + bias.y;
      const float kd = ki.x * ki.z - ki.y * ki.y;
      const f2 lt = Solver::friction_pair_solve(ki.z / kd, -ki.y / kd, ki.x / kd, v1, v2, lam.y, lam.z, lim, c.omega);
      const float dt1 = lt.x - lam.y, dt2 = lt.y - lam.z;
      lam.y = lt.x; lam.z = lt.y;
      ta[0] += t1a0 * dt1 + t2a0 * dt2;
      ta[1] += t1a1 * dt1 + t2a1 * dt2;
      tb[0] += t1b0 * dt1 + t2b0 * dt2;
      tb[1] += t1b1 * dt1 + t2b1 * dt2;
      const f3 f = t1 * dt1 + t2 * dt2;
      row_add_imp(c, a, -f, -cross(fra, f));
      row_add_imp(c, b, f, cross(frb, f));
    }
    Rw[6] = tov(lam);
  }
  if (a >= 0) { c.tw[(uint32_t)a * 2] = tov(F4(ta[0], 0.0f)); c.tw[(uint32_t)a * 2 + 1] = tov(F4(ta[1], 0.0f)); }
  if (b >= 0) { c.tw[(uint32_t)b * 2] = tov(F4(tb[0], 0.0f)); c.tw[(uint32_t)b * 2 + 1] = tov(F4(tb[1], 0.0f)); }
}

// ---- Articulation propagation of the sweep with the link mirror addressed as LDS (the Solver's members are generic
// pointers, so its out-of-line passes used flat accesses that wait for all outstanding memory).  Same arithmetic, in
// the same order, as Solver::art_contact_impulses and Solver::art_apply.
template <bool TWL>
struct ArtLds {
  AS_LDS float* lm;
  const AS_LDS v4u* lmeta;
  const AS_LDS v4u* ltree;
  AS_LDS float* lcon;
  AS_LDS float* laqd;
  typename TwSpace<TWL>::T* tw;
  typename TwSpace<TWL>::T* imp;
  const AS_GLOBAL v4u* art_info;   // statU + SU_ART
  uint32_t MP;
  DEV int mpar(uint32_t l) const { return (int)lmeta[l].x; }
  DEV uint32_t mnd(uint32_t l) const { return lmeta[l].y & 0xffu; }
  DEV uint32_t mdof(uint32_t l) const { return lmeta[l].z; }
  DEV uint32_t mbody(uint32_t l) const { return lmeta[l].w; }
  DEV AS_LDS float* ml(uint32_t l) const { return lm + (lmeta[l].y >> 8); }
  static DEV SVd msv(const AS_LDS float* p) { return mksv(mk3(p[0], p[1], p[2]), mk3(p[3], p[4], p[5])); }
  static DEV void mst(AS_LDS float* p, SVd v) { p[0] = v.a.x; p[1] = v.a.y; p[2] = v.a.z; p[3] = v.l.x; p[4] = v.l.y; p[5] = v.l.z; }
  DEV SVd mS(uint32_t l, uint32_t k) const { return msv(ml(l) + k * 6); }
  DEV SVd mU(uint32_t l, uint32_t k) const { return msv(ml(l) + 6 * mnd(l) + k * 6); }
  DEV float mDI(uint32_t l, uint32_t r, uint32_t c) const { uint32_t n = mnd(l); return ml(l)[12 * n + r * n + c]; }
  DEV AS_LDS float* mP(uint32_t l) const { uint32_t n = mnd(l); return ml(l) + 12 * n + n * n; }
  DEV AS_LDS float* mDV(uint32_t l) const { return mP(l) + 6; }
  DEV AS_LDS float* mV(uint32_t l) const { return mP(l) + 12; }
  DEV AS_LDS float* mUU(uint32_t l) const { return mP(l) + 18; }
  DEV AS_LDS float* mTAU(uint32_t l) const { return mP(l) + 18 + mnd(l); }
  DEV f3 mcom(uint32_t l) const { const AS_LDS float* p = mP(l) + 18 + 2 * mnd(l); return mk3(p[0], p[1], p[2]); }
  DEV f3 mr(uint32_t l) const {
    int par = mpar(l);
    if (par < 0) return mk3(0, 0, 0);
    return mcom(l) - mcom((uint32_t)par);
  }
  DEV void set_link_twist(uint32_t l, SVd v) const {
    uint32_t b = mbody(l);
    tw[b * 2] = tov(F4(v.l, 0.0f)); tw[b * 2 + 1] = tov(F4(v.a, 0.0f));
  }
  // Joint-limit, friction and drive rows (Solver::art_propagate_rows with LDS-typed accesses).
  DEV void propagate_rows(const Solver& sv, uint32_t a, uint32_t lane) const {
    const W& w = sv.w;
    const bool vpass = sv.vpass;
    const v4u ai = art_info[a];
    AS_GLOBAL float* art_ = (AS_GLOBAL float*)w.d.art;
    uint32_t n = tu(art_[w.h(W_A_RLN) + a]);
    float dt = w.hf(W_DT), beta = w.hf(W_BETA), maxd = w.hf(W_MAX_DEPEN), omega = w.hf(W_F_OMEGA);
    for (uint32_t i = 0; i < n; i++) {
      uint32_t dd = tu(art_[w.h(W_A_RL) + ai.z + i]);
      float4 lim = w.dof_limit(dd);
      uint32_t l = tu(lim.w), r = dd - w.link_dof(l);
      float hinv = mx(art_[w.h(W_A_HINV) + dd], 1e-12f);
      const float4 dr = w.dof_drive(dd);
      if (w.h(W_DRIVE_ROWS) && (dr.y > 0.0f || dr.z > 0.0f) && !(vpass && !w.h(W_VEL_DRIVE_ACTIVE))) {
        // drive_mode rows (wgsl.js:3425-3438): a PhysX-style drive row, cold-started each step, not relaxed.
        const float q = art_[w.h(W_A_Q) + w.link_qs(l) + r];
        const float pd = vpass ? dt * art_[w.h(W_A_SNAP) + w.NB() * 8 + dd] : 0.0f;
        const float ca = dt * (dt * dr.y + dr.z);
        const float x = 1.0f / (1.0f + ca * hinv);
        const float bias = x * dt * dr.z * art_[w.h(W_A_TGTV) + dd] + dr.y * x * dt * (art_[w.h(W_A_TGT) + dd] - q - pd);
        const float cap = dr.w > 0.0f ? dr.w * dt : 1e20f;
        const float lam = art_[w.h(W_A_DRV) + dd];
        const float nl = clampf_(lam * (1.0f - x) - x * ca * laqd[dd] + bias, -cap, cap);
        if (nl != lam) {
          art_[w.h(W_A_DRV) + dd] = nl;
          if (lane == 0) mTAU(l)[r] = nl - lam;
          apply(a, lane);
        }
      }
      // Joint friction: a velocity row with |impulse| <= friction * dt (plus the load term).
      float fr = w.dof_passive(dd).y;
      if (w.h(W_JFL)) fr += art_[w.h(W_A_JFL) + dd];
      if (fr > 0.0f) {
        float lam = art_[w.h(W_A_FRIC) + dd];
        float nl = clampf_(lam - omega * laqd[dd] / hinv, -fr * dt, fr * dt);
        if (nl != lam) {
          art_[w.h(W_A_FRIC) + dd] = nl;
          if (lane == 0) mTAU(l)[r] = nl - lam;
          apply(a, lane);
        }
      }
      if (art_[w.h(W_A_LG) + dd] > 0.5f) {
        float q = art_[w.h(W_A_Q) + w.link_qs(l) + r];
        for (uint32_t side = 0; side < 2; side++) {
          float sg = side == 0 ? 1.0f : -1.0f;
          float gap = side == 0 ? q - lim.x : lim.y - q;
          float lam = art_[w.h(W_A_LIM) + 2 * dd + side];
          float rhs = gap / dt;
          if (gap < 0.0f) rhs = vpass ? 0.0f : mx(beta * gap / dt, -maxd);
          float nl = mx(lam - omega * (sg * laqd[dd] + rhs) / hinv, 0.0f);
          if (nl == lam) continue;
          art_[w.h(W_A_LIM) + 2 * dd + side] = nl;
          if (lane == 0) mTAU(l)[r] = sg * (nl - lam);
          apply(a, lane);
        }
      }
    }
  }
  // Link twists from the joint velocities, root first (Solver::art_twists with LDS-typed accesses).
  DEV void twists(uint32_t a, uint32_t lane) const {  // wgsl.js:3240-3251
    const v4u ai = art_info[a];
    const bool has = lane < ai.y;
    const uint32_t l = ai.x + lane, dl = has ? ltree[l].x : 0xffffffffu;
    Solver::wave_sync();
    for (uint32_t d = 0; d < MP; d++) {
      if (dl == d) {
        const int par = mpar(l);
        SVd v = sv0();
        if (par >= 0) v = mshift(msv(mV((uint32_t)par)), mr(l));
        const uint32_t n = mnd(l), dof = mdof(l);
        for (uint32_t r = 0; r < n; r++) v = sadd(v, sscale(mS(l, r), laqd[dof + r]));
        mst(mV(l), v);
        set_link_twist(l, v);
      }
      Solver::wave_sync();
    }
  }
  DEV void apply(uint32_t a, uint32_t lane) const {  // wgsl.js:3254-3299
    const v4u ai = art_info[a];
    const bool has = lane < ai.y;
    const uint32_t l = ai.x + lane, dl = has ? ltree[l].x : 0xffffffffu;
    Solver::wave_sync();
    for (int d = (int)MP - 1; d >= 0; d--) {
      if (dl == (uint32_t)d) {
        uint32_t n = mnd(l);
        AS_LDS float *UU = mUU(l), *TAU = mTAU(l);
        SVd p = msv(mP(l));
        float u[6];
        for (uint32_t r = 0; r < n; r++) { u[r] = TAU[r] - sdot(mS(l, r), p); UU[r] = u[r]; }
        if (mpar(l) >= 0) {
          SVd pp = p;
          for (uint32_t r = 0; r < n; r++) {
            float du = 0.0f;
            for (uint32_t c = 0; c < n; c++) du += mDI(l, r, c) * u[c];
            pp = sadd(pp, sscale(mU(l, r), du));
          }
          mst(lcon + l * 8, fshift(pp, mr(l)));
        }
      }
      Solver::wave_sync();
      if (d > 0 && dl == (uint32_t)(d - 1)) {
        uint32_t c = ltree[l].y;
        if (c != 0xffffffffu) {
          SVd P = msv(mP(l));
          for (; c != 0xffffffffu; c = ltree[c].z) P = sadd(P, msv(lcon + c * 8));
          mst(mP(l), P);
        }
      }
      Solver::wave_sync();
    }
    for (uint32_t d = 0; d < MP; d++) {
      if (dl == d) {
        int par = mpar(l);
        SVd dv = sv0();
        if (par >= 0) dv = mshift(msv(mDV((uint32_t)par)), mr(l));
        uint32_t n = mnd(l);
        const AS_LDS float* UU = mUU(l);
        float dq[6];
        for (uint32_t r = 0; r < n; r++) {
          dq[r] = 0.0f;
          for (uint32_t c = 0; c < n; c++) dq[r] += mDI(l, r, c) * (UU[c] - sdot(mU(l, c), dv));
        }
        for (uint32_t r = 0; r < n; r++) { dv = sadd(dv, sscale(mS(l, r), dq[r])); laqd[mdof(l) + r] += dq[r]; }
        mst(mDV(l), dv);
        SVd v = sadd(msv(mV(l)), dv);
        mst(mV(l), v);
        set_link_twist(l, v);
        mst(mP(l), sv0());
        AS_LDS float* TAU = mTAU(l);
        for (uint32_t r = 0; r < n; r++) TAU[r] = 0.0f;
      }
      Solver::wave_sync();
    }
  }
  DEV void contact_impulses(uint32_t a, uint32_t lane) const {  // wgsl.js:3380-3394
    const v4u ai = art_info[a];
    bool hit = false;
    if (lane < ai.y) {
      uint32_t l = ai.x + lane, b = mbody(l);
      const v4f i0 = imp[b * 2], i1 = imp[b * 2 + 1];
      f3 f = xyzv(i0), t = xyzv(i1);
      if (vne0(f) || vne0(t)) {
        hit = true;
        mst(mP(l), mksv(-t, -f));
        imp[b * 2] = tov(make_float4(0, 0, 0, 0)); imp[b * 2 + 1] = tov(make_float4(0, 0, 0, 0));
      }
    }
    if (__ballot(hit) != 0) apply(a, lane);
  }
};

template <bool TWL>
DEV ArtLds<TWL> art_lds(const Solver& sv) {
  ArtLds<TWL> m;
  m.lm = (AS_LDS float*)sv.lm;
  m.lmeta = (const AS_LDS v4u*)sv.lmeta;
  m.ltree = (const AS_LDS v4u*)sv.ltree;
  m.lcon = (AS_LDS float*)sv.lcon;
  m.laqd = (AS_LDS float*)sv.laqd;
  m.tw = (typename TwSpace<TWL>::T*)sv.tw;
  m.imp = (typename TwSpace<TWL>::T*)sv.imp;
  m.art_info = (const AS_GLOBAL v4u*)(sv.w.d.statU + sv.w.h(W_SU_ART));
  m.MP = sv.w.h(W_MAXPATH);
  return m;
}

template <int BS, bool FAST = false, bool TWL = false, class TM>
DEV void dense_phase(const TM& tm, const Solver& sv, uint32_t mode, uint32_t nd) {
  const uint32_t tid = tm.tid, NT = tm.NT, lane = tid & 63, wave = tid >> 6, NA = sv.w.NA();
  for (uint32_t base = 0; base < NA; base += NT / 64) {
    const uint32_t a = base + wave;
    bool touched = false;
    if (a < NA) touched = sv.art_dense(a, mode, nd, lane);
    tm.sync();
    if (a < NA && touched) {
      if constexpr (FAST) art_lds<TWL>(sv).twists(a, lane);
      else sv.art_twists(a, lane);
    }
    tm.sync();
  }
}

// FAST: the default-path rows go through solve_row with the twists in LDS (TWL) or global memory.
template <int BS, bool FAST = false, bool TWL = false, class TM>
DEV void sweep(const TM& tm, const Solver& sv, uint32_t mode, uint32_t ncol, uint32_t nu, uint32_t nd) {
  const W& w = sv.w;
  const uint32_t tid = tm.tid, NT = tm.NT, NA = w.NA(), NCOL = w.NCOL(), PO = w.PO();
  const uint32_t* sched = w.d.sched;
  const SchedLayout& S = w.S;
  const uint32_t lane = tid & 63, wave = tid >> 6;
  const bool ext = w.h(W_FMODE) || w.h(W_ANG) || w.h(W_TORS) || w.h(W_COMPL);
  // Out-of-line calls spill the caller's live registers: call only for articulations with joint rows.
  for (uint32_t a = wave; a < NA; a += NT / 64)
    if (!FAST) sv.art_propagate_rows(a, lane);
    else if (w.d.art[w.h(W_A_RLN) + a] != 0.0f) art_lds<TWL>(sv).propagate_rows(sv, a, lane);
  if (w.h(W_PATCH) && mode == 1) {
    if (FAST) {
      const AS_GLOBAL uint32_t* ust = (const AS_GLOBAL uint32_t*)(sched + S.USTART);
      const AS_GLOBAL uint32_t* usz = (const AS_GLOBAL uint32_t*)(sched + S.USIZE);
      const AS_GLOBAL float* rw = (const AS_GLOBAL float*)w.d.rows;
      AS_GLOBAL float* ul = (AS_GLOBAL float*)(w.d.art + w.h(W_A_ULOAD));
      const uint32_t RW = w.ROWS() * 4;
      for (uint32_t us = tid; us < nu; us += NT) {
        const uint32_t s0 = ust[us], n = usz[us];
        float load = 0.0f;
        for (uint32_t m = s0; m < s0 + n; m++) {
          const float role = rw[(size_t)m * RW + PO * 4 + 3], ln = rw[(size_t)m * RW + 6 * 4];
          if (role != 1.0f) load += ln;
        }
        ul[us] = load;
      }
    } else
    for (uint32_t us = tid; us < nu; us += NT) {
      uint32_t s0 = sched[S.USTART + us];
      float load = 0.0f;
      for (uint32_t m = s0; m < s0 + sched[S.USIZE + us]; m++) if (sv.row(m, PO).w != 1.0f) load += sv.row(m, 6).x;
      w.d.art[w.h(W_A_ULOAD) + us] = load;
    }
  }
  tm.sync();
  if (!FAST || nd) dense_phase<BS, FAST, TWL>(tm, sv, 1, nd);
  const uint4* slots = (const uint4*)(sched + S.SLOTT);
  if (FAST && !ext) {
    const RowCtx<TWL> rc = row_ctx<TWL>(sv);
    const AS_GLOBAL uint32_t* cst = (const AS_GLOBAL uint32_t*)(sched + S.CSTART);
    const uint32_t nc = min(NCOL, ncol);
    uint32_t s = cst[0];
    for (uint32_t c = 0; c < nc; c++) {
      const uint32_t e = cst[c + 1];
      for (uint32_t u = s + tid; u < e; u += NT) { const v4u sl = rc.slots[u]; solve_row<TWL>(rc, sl.y, sl.x, (int)sl.z, (int)sl.w, mode); }
      s = e;
      tm.sync();
    }
  } else
  for (uint32_t c = 0; c < NCOL && c < ncol; c++) {
    uint32_t s = sched[S.CSTART + c], e = sched[S.CSTART + c + 1];
    if (ext)
      for (uint32_t u = s + tid; u < e; u += NT) { uint4 sl = slots[u]; sv.solve_unit<true>(sl.y, sl.x, (int)sl.z, (int)sl.w, mode); }
    else
      for (uint32_t u = s + tid; u < e; u += NT) { uint4 sl = slots[u]; sv.solve_unit<false>(sl.y, sl.x, (int)sl.z, (int)sl.w, mode); }
    tm.sync();
  }
  if (tid == 0) {
    uint32_t MAXC = w.MAXC();
    for (uint32_t u = sched[S.CSTART + NCOL]; u < sched[S.CSTART + NCOL + 1]; u++) {
      uint4 sl = slots[min(u, MAXC - 1)];
      if (ext) sv.solve_unit<true>(sl.y, sl.x, (int)sl.z, (int)sl.w, mode);
      else sv.solve_unit<false>(sl.y, sl.x, (int)sl.z, (int)sl.w, mode);
    }
  }
  tm.sync();
  if (FAST) {
    const ArtLds<TWL> am = art_lds<TWL>(sv);
    const bool vlim = w.h(W_VLIM_ON) != 0;
    for (uint32_t a = wave; a < NA; a += NT / 64) {
      am.contact_impulses(a, lane);
      if (vlim && w.d.art[w.h(W_A_AVL) + a] > 0.5f) sv.art_velocity_limits(a, lane);
    }
  } else
  for (uint32_t a = wave; a < NA; a += NT / 64) {
    sv.art_contact_impulses(a, lane);
    sv.art_velocity_limits(a, lane);
  }
  if (w.h(W_VLIM_FREE)) {
    tm.sync();
    for (uint32_t b = tid; b < w.NB(); b += NT) sv.body_velocity_limits(b);
  }
  tm.sync();
}

// One Nesterov momentum step before position sweep `it` (wgsl.js:3527-3557).  Restart test: the sum over the
// colored rows of (lam - lam_last - e) . e, as 256 strided partials (partial j over rows j, j + 256, ...) added in
// order, whatever the block size; below zero restarts (k = 0), else k + 1 and beta = (k - 1) / (k + 2) for k > 1.
// Then the rows extrapolate color by color (the overflow color serially) and the articulations take their impulses.
// kb: (k, beta) of the previous step, returned updated.  Not inlined: inlined, it pushed k_solve into spilling.
template <int BS>
__device__ __noinline__ f2 momentum_step(const Solver& sv, uint32_t ncol, uint32_t it, f2 kb) {
  float mk = kb.x, mbeta = kb.y;
  const W& w = sv.w;
  const uint32_t tid = threadIdx.x, NA = w.NA(), NCOL = w.NCOL(), ROWS = w.ROWS(), MO = w.h(W_MO), MAXC = w.MAXC();
  const uint32_t lane = tid & 63, wave = tid >> 6;
  const uint32_t* sched = w.d.sched;
  const SchedLayout& S = w.S;
  const uint4* slots = (const uint4*)(sched + S.SLOTT);
  if (it > 0) {
    const uint32_t nrows = min(sched[S.CSTART + NCOL + 1], MAXC);
    for (uint32_t j = tid; j < 256; j += BS) {
      float p = 0.0f;
      for (uint32_t u = j; u < nrows; u += 256) {
        const float4* R = w.d.rows + (size_t)slots[u].x * ROWS;
        const f3 e = xyz(R[MO + 1]);
        p += dot(xyz(R[6]) - xyz(R[MO]) - e, e);
      }
      sv.red[j] = p;
    }
    __syncthreads();
    float total = 0.0f;
    for (uint32_t j = 0; j < 256; j++) total += sv.red[j];
    mk = total < 0.0f ? 0.0f : mk + 1.0f;
    mbeta = mk > 1.0f ? (mk - 1.0f) / (mk + 2.0f) : 0.0f;
    __syncthreads();
  } else {
    mk = 0.0f; mbeta = 0.0f;
  }
  for (uint32_t c = 0; c < NCOL && c < ncol; c++) {
    const uint32_t s = sched[S.CSTART + c], e = sched[S.CSTART + c + 1];
    for (uint32_t u = s + tid; u < e; u += BS) { const uint4 sl = slots[u]; sv.momentum_unit(sl.x, (int)sl.z, (int)sl.w, mbeta, it == 0); }
    __syncthreads();
  }
  if (tid == 0) {
    for (uint32_t u = sched[S.CSTART + NCOL]; u < sched[S.CSTART + NCOL + 1]; u++) {
      const uint4 sl = slots[min(u, MAXC - 1)];
      sv.momentum_unit(sl.x, (int)sl.z, (int)sl.w, mbeta, it == 0);
    }
  }
  __syncthreads();
  for (uint32_t a = wave; a < NA; a += BS / 64) {
    sv.art_contact_impulses(a, lane);
    sv.art_velocity_limits(a, lane);
  }
  __syncthreads();
  return mk2(mk, mbeta);
}

// Impulses for the next substep's warm start (wgsl.js:3737-3767): every row's key is inserted by linear probing
// (compare-and-swap, at most 32 probes) and each key keeps the value of its largest row order, so the table repeats
// run to run.  The winners were cleared with this table in prepare.  Not inlined, to keep k_solve's registers for
// the common path.
template <int BS>
__device__ __noinline__ void warm_store(const Solver& sv, uint32_t nrows) {
  const W& w = sv.w;
  const WorldDev& wd = w.d;
  const SchedLayout& S = w.S;
  const uint32_t tid = threadIdx.x, MAXC = w.MAXC(), WH = w.h(W_WH);
  uint32_t* sched = wd.sched;
  const uint32_t nc = min(nrows, MAXC), nw = nc + min(sched[S.META + 3], w.h(W_MAXDENSE)), wb = (1 - sv.parity) * WH;
  unsigned long long* win = (unsigned long long*)(sched + S.WWIN);
  for (uint32_t i = tid; i < nw; i += BS) {
    sched[S.WSLOT + i] = 0xffffffffu;
    const uint32_t key = sv.warm_key(i, nc);
    if (key == 0xffffffffu) continue;
    uint32_t h = key % WH;
    for (uint32_t probe = 0; probe < 32; probe++) {
      const uint32_t old = atomicCAS(&wd.warm[wb + h], 0u, key);
      if (old == 0u || old == key) { atomicMax(&win[h], sv.warm_order(sv.warm_value(i, nc).w, i)); sched[S.WSLOT + i] = h; break; }
      h = (h + 1) % WH;
    }
  }
  __syncthreads();
  for (uint32_t i = tid; i < nw; i += BS) {
    const uint32_t h = sched[S.WSLOT + i];  // written by this thread above
    if (h >= WH) continue;
    const float4 v = sv.warm_value(i, nc);
    if (__hip_atomic_load(&win[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != sv.warm_order(v.w, i)) continue;
    ((uint4*)(wd.warm + 2 * WH))[wb + h] = make_uint4(fbits(v.w), fbits(v.x), fbits(v.y), fbits(v.z));
  }
  __syncthreads();
}

// Velocity pass (velocity_iterations, wgsl.js:3700-3736): positions integrate with the velocities reached by the
// position sweeps (kept as a snapshot); bias-free sweeps from the accumulated impulses then set the output
// velocities.  Contact right-hand sides become 0, or a speculative gap still open at the snapshot velocity.
// Not inlined, to keep k_solve's registers for the common path.
template <int BS>
__device__ __noinline__ void velocity_pass(const Solver& sv, uint32_t ncol, uint32_t nu, uint32_t nd) {
  const W& w = sv.w;
  const WorldDev& wd = w.d;
  const SchedLayout& S = w.S;
  const uint32_t tid = threadIdx.x, NB = w.NB(), ND = w.ND(), MAXC = w.MAXC();
  const uint32_t* sched = wd.sched;
  AS_GLOBAL float* art_ = (AS_GLOBAL float*)wd.art;
  const uint32_t VI = w.h(W_VELITERS);
  const uint32_t SN = w.h(W_A_SNAP), ROWS = w.ROWS();
  const float dt = w.hf(W_DT);
  for (uint32_t b = tid; b < NB; b += BS) { w.st3(SN + b * 8, xyz(sv.tw[b * 2])); w.st3(SN + b * 8 + 4, xyz(sv.tw[b * 2 + 1])); }
  for (uint32_t i = tid; i < ND; i += BS) art_[SN + NB * 8 + i] = sv.laqd[i];
  for (uint32_t us = tid; us < nu; us += BS) {
    const uint64_t key = *(const uint64_t*)(sched + S.UKEY + 2 * us);
    const int a = key_a(key), b = key_b(key);
    f3 ta[2], tb[2];
    sv.load_tw(a, ta); sv.load_tw(b, tb);
    const uint32_t s0 = sched[S.USTART + us];
    for (uint32_t i = s0; i < min(s0 + sched[S.USIZE + us], MAXC); i++) {
      float4* R = wd.rows + (size_t)i * ROWS;
      const float phi = R[6].w;
      if (phi >= 1e29f) continue;
      const float jv = dot(xyz(R[0]), (tb[0] + cross(tb[1], xyz(R[4]))) - (ta[0] + cross(ta[1], xyz(R[3]))));
      R[1].w = (phi > 0.0f && phi + dt * jv > 1e-6f) ? phi / dt : 0.0f;
    }
  }
  for (uint32_t i = tid; i < nd; i += BS) {
    const uint32_t o = w.h(W_A_DENSE) + i * w.h(W_DPS);
    if (art_[o + 21] > 0.5f || art_[o + 22] >= 1e29f) continue;
    const uint4 ai = w.art_info(tu(art_[o]));
    float jv = 0.0f;
    for (uint32_t k = 0; k < ai.w; k++) jv += art_[o + w.h(W_DHDR) + k] * sv.laqd[ai.z + k];
    const float phi = art_[o + 22];
    art_[o + 4] = (phi > 0.0f && phi + dt * jv > 1e-6f) ? phi / dt : 0.0f;
  }
  __syncthreads();
  Solver vs = sv;  // the bias-free copy; k_solve's solver is never written after set-up
  vs.vpass = true;
  const BlockTeam<BS> tm{threadIdx.x, nullptr, nullptr};
  for (uint32_t it = 0; it < VI; it++) sweep<BS>(tm, vs, 1, ncol, nu, nd);
}

// MGL: the launch holds a world whose link records and mirror live in global memory (WorldDev::mscr); every world of
// the launch then runs the generic-pointer sweep, the others keeping their mirror in LDS.
// Predictor of body b (wgsl.js solve prologue): impulses and twists start at zero; a free body's twist is its
// velocity with gravity, applied forces and the gyroscopic term.
DEV void predict_body(const Solver& sv, const WorldDev& wd, uint32_t b, f3 g, float dt) {
  const W& w = sv.w;
  uint32_t kind = w.body_kind((int)b);
  sv.imp[b * 2] = make_float4(0, 0, 0, 0); sv.imp[b * 2 + 1] = make_float4(0, 0, 0, 0);
  // Bodies other than free ones: zero (static bodies keep it; link bodies get theirs from art_twists below).
  sv.tw[b * 2] = make_float4(0, 0, 0, 0); sv.tw[b * 2 + 1] = make_float4(0, 0, 0, 0);
  if (kind != 1) return;
  uint32_t fl = w.body_flags((int)b);
  f3 v = xyz(wd.vel[b * 2]) + ((w.node_asleep((int)b) || fl != 0) ? mk3(0, 0, 0) : g * dt);
  f3 wv = xyz(wd.vel[b * 2 + 1]);
  uint32_t bfo = w.h(W_A_BF) + b * 8;
  if ((fl & 1) == 0 && (vne0(w.ld3(bfo)) || vne0(w.ld3(bfo + 4)))) {
    m3 R = qmat(w.pose_q((int)b));
    v += w.ld3(bfo) * w.body_inv_mass((int)b) * dt;
    wv += mul(R, mul(w.body_inv_inertia((int)b), mul(tr(R), w.ld3(bfo + 4)))) * dt;
  }
  m3 I = w.body_inertia((int)b);
  bool iso = I.c[0].x == I.c[1].y && I.c[1].y == I.c[2].z && I.c[0].y == 0.0f && I.c[0].z == 0.0f && I.c[1].z == 0.0f;
  if ((fl & 1) == 0 && !iso && vne0(wv)) {
    m3 R = qmat(w.pose_q((int)b));
    wv = mul(R, sv.gyroscopic_velocity(I, mul(tr(R), wv), dt));
  }
  sv.tw[b * 2] = F4(v, 0.0f); sv.tw[b * 2 + 1] = F4(wv, 0.0f);
}

// PRED: the body predictor runs here (false: k_solve_pred ran it over the device first).
template <int BS, bool TWL, bool MGL, class TM, bool PRED = true>
DEV void solve_world(const TM& tm, const WorldDev wd, uint32_t* Hs, float4* s_dyn4) {
  load_header(Hs, wd);
  Solver sv;
  sv.w = make_world(Hs, wd);
  const W& w = sv.w;
  const SchedLayout& S = w.S;
  const uint32_t tid = tm.tid, NT = tm.NT;
  sv.parity = wd.ctr[2] & 1u;
  const uint32_t NB = w.NB(), NA = w.NA(), NL = w.NL(), ND = w.ND(), MAXC = w.MAXC(), NCOL = w.NCOL(), AH = w.AH(), PO = w.PO();
  uint32_t* sched = wd.sched;
  AS_GLOBAL float* art_ = (AS_GLOBAL float*)wd.art;
  const float dt = w.hf(W_DT);
  const f3 g = mk3(w.hf(W_GX), w.hf(W_GY), w.hf(W_GZ));
  const uint32_t nu = sched[S.META], ncol = min(sched[S.META + 1], NCOL);
  const uint32_t nd = min(sched[S.META + 3], w.h(W_MAXDENSE));
  // Dynamic LDS (steel_solve_lds_bytes of the launch's largest world): link records, joint velocities,
  // with TWL the body twists and impulses, then the link mirror.
  char* lds = (MGL && wd.mscr) ? (char*)wd.mscr : (char*)s_dyn4;
  sv.lmeta = (uint4*)lds;
  lds += (size_t)NL * 16;
  sv.ltree = (uint4*)lds;
  lds += (size_t)NL * 16;
  sv.lcon = (float*)lds;
  lds += (size_t)NL * 32;
  sv.laqd = (float*)lds;
  lds += ((size_t)ND * 4 + 15) & ~(size_t)15;
  if (TWL) {
    sv.tw = (float4*)lds;
    sv.imp = sv.tw + 2 * (size_t)max(NB, 1u);
    lds += (size_t)max(NB, 1u) * 64;
  } else {
    sv.tw = wd.tw;
    sv.imp = wd.imp;
  }
  const bool NEST = w.h(W_NEST) != 0;
  sv.red = (float*)lds;
  if (NEST) lds += 1024;
  sv.lm = (float*)lds;
  // Link records and mirror offsets (exclusive prefix of mirror_floats(ndof) over links).
  uint32_t* msz = (uint32_t*)sv.lm;
  for (uint32_t l = tid; l < NL; l += NT) {
    uint32_t n = w.link_ndof(l);
    sv.lmeta[l] = make_uint4((uint32_t)w.link_parent(l), n, w.link_dof(l), (uint32_t)w.link_body(l));
    msz[l] = mirror_floats(n);
  }
  tm.sync();
  const uint32_t MF = tm.scan(msz, NL);
  for (uint32_t l = tid; l < NL; l += NT) sv.lmeta[l].y |= msz[l] << 8;
  tm.sync();
  // Link tree: depth by walking to the root; children lists in decreasing link order (built in link order).
  for (uint32_t l = tid; l < NL; l += NT) {
    uint32_t d = 0;
    for (int p = sv.mpar(l); p >= 0; p = sv.mpar((uint32_t)p)) d++;
    sv.ltree[l] = make_uint4(d, 0xffffffffu, 0xffffffffu, 0u);
  }
  tm.sync();
  // Children lists, a thread per articulation (a link's parent is in its articulation), links in order.
  for (uint32_t a = tid; a < NA; a += NT) {
    const uint4 ai = w.art_info(a);
    for (uint32_t l = ai.x; l < ai.x + ai.y; l++) {
      int p = sv.mpar(l);
      if (p >= 0) { sv.ltree[l].z = sv.ltree[p].y; sv.ltree[p].y = l; }
    }
  }
  tm.sync();
  const uint32_t JMw = w.JM();
  auto mirror_link = [&](uint32_t i) {  // link holding mirror float i (largest l with offset <= i)
    uint32_t lo = 0, hi = NL - 1;
    while (lo < hi) {
      uint32_t mid = (lo + hi + 1) >> 1;
      if ((sv.lmeta[mid].y >> 8) <= i) lo = mid; else hi = mid - 1;
    }
    return lo;
  };
  for (uint32_t i = tid; i < MF; i += NT) {
    uint32_t l = mirror_link(i), n = sv.mnd(l), src = mirror_src(i - (sv.lmeta[l].y >> 8), n, JMw);
    sv.lm[i] = src < LB ? art_[w.lo(l) + src] : 0.0f;
  }
  for (uint32_t i = tid; i < ND; i += NT) sv.laqd[i] = art_[w.h(W_A_QDH) + i];
  if (w.h(W_SLEEP)) {  // asleep articulations start at rest (wgsl.js:3640-3646)
    tm.sync();
    for (uint32_t a = tid; a < NA; a += NT) {
      if (!w.node_asleep((int)(NB + a))) continue;
      uint4 ai = w.art_info(a);
      for (uint32_t dd = ai.z; dd < ai.z + ai.w; dd++) sv.laqd[dd] = 0.0f;
    }
  }
  if (PRED)
    for (uint32_t b = tid; b < NB; b += NT) predict_body(sv, wd, b, g, dt);
  for (uint32_t i = tid; i < 2 * ND; i += NT) art_[w.h(W_A_LIM) + i] = 0.0f;
  for (uint32_t i = tid; i < ND; i += NT) { art_[w.h(W_A_FRIC) + i] = 0.0f; art_[w.h(W_A_DRV) + i] = 0.0f; }
  const uint32_t nslot = min(sched[S.CSTART + NCOL + 1], MAXC);
  uint4* slots = (uint4*)(sched + S.SLOTT);
  for (uint32_t u = tid; u < nslot; u += NT) {
    const uint32_t r = sched[S.PORD + u], us = sched[S.RUNIT + r];
    const uint64_t key = w.sch64(S.UKEY)[us];
    slots[u] = make_uint4(r, us, (uint32_t)key_a(key), (uint32_t)key_b(key));
  }
  tm.sync();
  for (uint32_t a = tid >> 6; a < NA; a += NT / 64) {
    if constexpr (!MGL) art_lds<TWL>(sv).twists(a, tid & 63);
    else sv.art_twists(a, tid & 63);
  }
  tm.sync();
  dense_phase<BS, !MGL, TWL>(tm, sv, 0, nd);
  sweep<BS, !MGL, TWL>(tm, sv, 0, ncol, nu, nd);
  const uint32_t iters = w.h(W_ITERS);
  f2 kb = mk2(0.0f, 0.0f);
  for (uint32_t it = 0; it < iters; it++) {
    if constexpr (!TM::GRID) if (NEST) kb = momentum_step<BS>(sv, ncol, it, kb);
    sweep<BS, !MGL, TWL>(tm, sv, 1, ncol, nu, nd);
  }
  if constexpr (!TM::GRID) if (w.h(W_VELITERS)) velocity_pass<BS>(sv, ncol, nu, nd);
  // Mutable mirror fields back to the link blocks (p, dv, v, u, tau; S, U, D^-1, COM are unchanged).
  for (uint32_t i = tid; i < MF; i += NT) {
    uint32_t l = mirror_link(i), n = sv.mnd(l), f = i - (sv.lmeta[l].y >> 8);
    if (mirror_mutable(f, n)) art_[w.lo(l) + mirror_src(f, n, JMw)] = sv.lm[i];
  }
  tm.sync();
  const uint32_t nrows = sched[S.USTART + nu];
  if constexpr (!TM::GRID) if (w.h(W_WH)) warm_store<BS>(sv, nrows);
  const bool PATCH = w.h(W_PATCH) != 0;
  const uint32_t PB = w.h(W_P_BODY), PP = w.h(W_P_PREV);
  if (PATCH) {
    for (uint32_t b = tid; b < NB * 2; b += NT) wd.poses[PP + b] = wd.poses[PB + b];
    tm.sync();
  }
  for (uint32_t b = tid; b < NB; b += NT) {
    if (w.body_kind((int)b) != 1) continue;
    if (w.node_asleep((int)b)) { wd.vel[b * 2] = make_float4(0, 0, 0, 0); wd.vel[b * 2 + 1] = make_float4(0, 0, 0, 0); continue; }
    f3 v = xyz(sv.tw[b * 2]), wv = xyz(sv.tw[b * 2 + 1]);
    float damp = (w.body_flags((int)b) & 1) != 0 ? 1.0f : 1.0f - w.body_limits((int)b).x * dt;
    wd.vel[b * 2] = F4(v, 0.0f); wd.vel[b * 2 + 1] = F4(wv * damp, 0.0f);
    if (w.h(W_VELITERS)) { v = w.ld3(w.h(W_A_SNAP) + b * 8); wv = w.ld3(w.h(W_A_SNAP) + b * 8 + 4); }
    f3 c = xyz(wd.poses[PB + b * 2]) + v * dt;
    q4 q = Q(wd.poses[PB + b * 2 + 1]);
    q = nrm4(q + (0.5f * dt) * qmul(mk4(wv, 0.0f), q));
    wd.poses[PB + b * 2] = F4(c, 0.0f);
    wd.poses[PB + b * 2 + 1] = F4(q);
  }
  for (uint32_t i = tid; i < ND; i += NT) {
    art_[w.h(W_A_QD) + i] = sv.laqd[i];
    if (w.h(W_VELITERS)) sv.laqd[i] = art_[w.h(W_A_SNAP) + NB * 8 + i];
  }
  tm.sync();
  for (uint32_t l = tid; l < NL; l += NT) {
    if (w.link_type(l) != JT_FREE || w.link_parent(l) >= 0) continue;
    float k = 1.0f - w.body_limits(w.link_body(l)).x * dt;
    uint32_t dd = w.h(W_A_QD) + w.link_dof(l);
    art_[dd + 3] *= k; art_[dd + 4] *= k; art_[dd + 5] *= k;
  }
  // Joint positions (per link), then FK and body poses, one wave per articulation.
  for (uint32_t base = 0; base < NA; base += NT / 64) {
    const uint32_t a = base + (tid >> 6), k = tid & 63;
    if (a >= NA) continue;
    const uint4 ai = w.art_info(a);
    const uint32_t AQ = w.h(W_A_Q), l = ai.x + k;
    if (k < ai.y) {
      uint32_t qs = w.link_qs(l), dd = w.link_dof(l), t = w.link_type(l);
      if (t == JT_REVOLUTE || t == JT_PRISMATIC || t == JT_D6)
        for (uint32_t r = 0; r < w.link_ndof(l); r++) art_[AQ + qs + r] += sv.laqd[dd + r] * dt;
      if (t == JT_BALL) {
        f3 wv = mk3(sv.laqd[dd], sv.laqd[dd + 1], sv.laqd[dd + 2]);
        q4 q = mk4(art_[AQ + qs], art_[AQ + qs + 1], art_[AQ + qs + 2], art_[AQ + qs + 3]);
        q = nrm4(q + (0.5f * dt) * qmul(mk4(wv, 0.0f), q));
        art_[AQ + qs] = q.x; art_[AQ + qs + 1] = q.y; art_[AQ + qs + 2] = q.z; art_[AQ + qs + 3] = q.w;
      }
      if (t == JT_FREE) {
        for (uint32_t c = 0; c < 3; c++) art_[AQ + qs + c] += sv.laqd[dd + c] * dt;
        f3 wv = mk3(sv.laqd[dd + 3], sv.laqd[dd + 4], sv.laqd[dd + 5]);
        q4 q = mk4(art_[AQ + qs + 3], art_[AQ + qs + 4], art_[AQ + qs + 5], art_[AQ + qs + 6]);
        q = nrm4(q + (0.5f * dt) * qmul(mk4(wv, 0.0f), q));
        art_[AQ + qs + 3] = q.x; art_[AQ + qs + 4] = q.y; art_[AQ + qs + 5] = q.z; art_[AQ + qs + 6] = q.w;
      }
    }
    w.art_fk_lv(a, k < ai.y ? sv.ltree[l].x : 0xffffffffu);
    if (k < ai.y) {
      uint32_t b = (uint32_t)w.link_body(l);
      wd.poses[PB + b * 2] = F4(w.ld3(w.lo(l) + L_COM), 0.0f);
      wd.poses[PB + b * 2 + 1] = F4(w.link_rot(l));
    }
  }
  if (PATCH) {
    tm.sync();
    uint32_t ab = (1 - sv.parity) * AH;
    for (uint32_t i = tid; i < nrows; i += NT) {
      if (sv.row(i, PO).w != 2.0f) continue;
      const uint32_t pair = fbits(sv.row(i, PO + 3).w);  // anchor-table pair id; the bodies from the row's unit
      const uint64_t ukey = w.sch64(S.UKEY)[sched[S.RUNIT + i]];
      const int a = key_a(ukey), b = key_b(ukey);
      f3 fra = xyz(sv.row(i, PO)), frb = xyz(sv.row(i, PO + 1));
      uint32_t k = tu(sv.row(i, PO + 1).w);
      float4 ex = sv.row(i, PO + 2);
      float lim = ex.w;
      if (lim <= 0.0f) continue;
      f3 n = xyz(sv.row(i, 0)), t1 = xyz(sv.row(i, 1)), t2 = xyz(sv.row(i, 2));
      f3 ta[2], tb[2];
      if (w.h(W_VELITERS)) { sv.snap_tw(a, ta); sv.snap_tw(b, tb); } else { sv.load_tw(a, ta); sv.load_tw(b, tb); }
      f3 vrel = (tb[0] + cross(tb[1], frb)) - (ta[0] + cross(ta[1], fra));
      float4 lam = sv.row(i, 6);
      f3 lt = t1 * lam.y + t2 * lam.z;
      float r = mx(mn(w.body_radius(a), w.body_radius(b)), 1e-4f);
      if (len(lt) >= (1.0f - 1e-5f) * lim && dot(vrel, lt) < -lim * 1e-5f * r / dt) continue;
      q4 qa = mk4(0, 0, 0, 1);
      f3 ca = mk3(0, 0, 0);
      if (a >= 0) { qa = w.pose_q(a); ca = w.pose_c(a); }
      f3 ll = xyz(sv.row(i, PO + 3));
      f3 x = ca + qrot(qa, ll);
      f3 e = xyz(ex) + sv.body_motion(b, x) - sv.body_motion(a, x);
      e -= n * dot(n, e);
      f3 el = qinvrot(qa, e);
      uint32_t hk = max(hash_u32(pair * 2u + k), 1u);
      uint32_t h = hk % AH;
      for (uint32_t probe = 0; probe < 32; probe++) {
        if (atomicCAS(&wd.akey[ab + h], 0u, hk) == 0u) {
          uint32_t* av = wd.aval + (size_t)(ab + h) * AVN;
          f3 il = qinvrot(qa, lt), nl_ = qinvrot(qa, n);
          av[0] = pair; av[1] = k;
          av[2] = fbits(ll.x); av[3] = fbits(ll.y); av[4] = fbits(ll.z);
          av[5] = fbits(el.x); av[6] = fbits(el.y); av[7] = fbits(el.z);
          av[8] = fbits(il.x); av[9] = fbits(il.y); av[10] = fbits(il.z);
          av[11] = fbits(nl_.x); av[12] = fbits(nl_.y); av[13] = fbits(nl_.z);
          break;
        }
        h = (h + 1) % AH;
      }
    }
  }
  tm.sync();
  if (tid == 0) {
    uint32_t* st = wd.stats;
    st[1] = max(st[1], wd.ctr[1]);
    st[2] = max(st[2], ncol);
    st[3] = max(st[3], nu);
    st[4] = max(st[4], wd.ctr[0]);
    st[5] = wd.ctr[1];
    st[6] = ncol;
    st[8] = sched[S.META + 5];
    st[7] = max(st[7], sched[S.META + 3]);
    st[9] = wd.ctr[0];   // last substep: candidate pairs, rows, units (per-step traffic accounting)
    st[10] = nrows;
    st[11] = nu;
    st[12] += wd.ctr[0];  // running sums since reset: pairs, contacts, rows, substeps
    st[13] += min(wd.ctr[1], MAXC);
    st[14] += nrows;
    st[15] += 1;
    wd.ctr[0] = 0; wd.ctr[1] = 0;
    wd.ctr[2] += 1;
    if (wd.ctr[2] % w.h(W_SUBSTEPS) == 0) {  // end of a frame: the frame clock advances by 1 / fps, in double
      double ph = __hiloint2double((int)wd.ctr[5], (int)wd.ctr[4]) + 1.0 / (double)w.hf(W_FPS);
      wd.ctr[4] = (uint32_t)__double2loint(ph);
      wd.ctr[5] = (uint32_t)__double2hiint(ph);
    }
  }
}

template <int BS, bool TWL, bool MGL = false, bool PRED = true>
__global__ void __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(STEEL_WPE_SOLVE))) k_solve(const WorldDev* __restrict__ worlds, uint32_t nworlds) {
  __shared__ uint32_t Hs[STEEL_WORLD_HEADER_WORDS];
  __shared__ TeamLds<BS> tl;
  extern __shared__ float4 s_dyn4[];
  const uint32_t wid = blockIdx.x;
  if (wid >= nworlds) return;
  const BlockTeam<BS> tm{threadIdx.x, &tl, nullptr};
  solve_world<BS, TWL, MGL, BlockTeam<BS>, PRED>(tm, worlds[wid], Hs, s_dyn4);
}

// The body predictor of world 0 over the device (a large world alone in its batch, twists and impulses in its
// buffers), before its k_solve<1024, false, true, false>.
__global__ void __launch_bounds__(256) k_solve_pred(const WorldDev* __restrict__ world) {
  __shared__ uint32_t Hs[STEEL_WORLD_HEADER_WORDS];
  const WorldDev wd = world[0];
  load_header(Hs, wd);
  Solver sv;
  sv.w = make_world(Hs, wd);
  sv.tw = wd.tw;
  sv.imp = wd.imp;
  const float dt = sv.w.hf(W_DT);
  const f3 g = mk3(sv.w.hf(W_GX), sv.w.hf(W_GY), sv.w.hf(W_GZ));
  const uint32_t NB = sv.w.NB();
  for (uint32_t b = blockIdx.x * 256 + threadIdx.x; b < NB; b += gridDim.x * 256) predict_body(sv, wd, b, g, dt);
}

// One large world over a cooperative launch: twists, impulses and the link mirror in global memory (g: grid team
// scratch).
template <int BS>
__global__ void __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(STEEL_WPE_SOLVE))) k_solve_grid(const WorldDev* __restrict__ world, uint32_t* g) {
  __shared__ uint32_t Hs[STEEL_WORLD_HEADER_WORDS];
  __shared__ TeamLds<BS> tl;
  const GridTeam<BS> tm{gridDim.x * BS, blockIdx.x * BS + threadIdx.x, &tl, nullptr, g};
  solve_world<BS, false, true>(tm, world[0], Hs, nullptr);
}

// ------------------------------------------------------------------------------------------------
// Launchers
// ------------------------------------------------------------------------------------------------
hipError_t steel_launch_substep(const WorldDev* worlds, uint32_t nworlds, uint32_t block, const LaunchLds& sl, uint32_t* pair_offs,
                                uint32_t narrow_groups, hipStream_t stream, uint32_t stages, hipEvent_t* marks) {
  const bool twl = sl.tw_lds(block);
  const size_t solve_lds = sl.bytes(block);
  if (nworlds == 0) return hipSuccess;
  dim3 grid(nworlds);
  if (marks) (void)hipEventRecord(marks[0], stream);
  if (stages & STAGE_COLLIDE) {
    // Workgroups of 1024 (large worlds) keep the broadphase's sorted bounds in global memory (cap 0): 1024 threads'
    // sort LDS and a 4096-shape bounds cache would not fit one CU's LDS together.
    if (block == 1024 && nworlds == 1 && sl.cell) {
      // A large world alone in its batch: the cell broadphase's phases over the device, a launch each.
      const dim3 g(1024), bs(CELLW_BS);
      hipLaunchKernelGGL(k_cellw_init, dim3(1), dim3(64), 0, stream, worlds);
      hipLaunchKernelGGL(k_cellw_pose, g, bs, 0, stream, worlds);
      hipLaunchKernelGGL(k_cellw_origin, g, bs, 0, stream, worlds);
      hipLaunchKernelGGL(k_cellw_keys, g, bs, 0, stream, worlds);
      for (uint32_t sh = 0; sh < 64; sh += 8) {
        hipLaunchKernelGGL(k_cellw_count, dim3(CELLW_MAXG), bs, 0, stream, worlds, sh);
        hipLaunchKernelGGL(k_cellw_scan, dim3(1), dim3(1024), 0, stream, worlds, sh);
        hipLaunchKernelGGL(k_cellw_scatter, dim3(CELLW_MAXG), bs, 0, stream, worlds, sh);
      }
      hipLaunchKernelGGL(k_cellw_sortfix, g, bs, 0, stream, worlds);
      hipLaunchKernelGGL(k_cellw_gather, g, bs, 0, stream, worlds);
      hipLaunchKernelGGL(k_cellw_insert, g, bs, 0, stream, worlds);
      hipLaunchKernelGGL(k_cellw_pairs, g, bs, 0, stream, worlds);
      hipLaunchKernelGGL(k_cellw_big, g, bs, 0, stream, worlds);
      hipLaunchKernelGGL(k_cellw_finish, dim3(1), dim3(64), 0, stream, worlds);
    } else if (block == 1024) hipLaunchKernelGGL((k_broad<1024>), grid, dim3(1024), 16, stream, worlds, nworlds, 0u);
    else hipLaunchKernelGGL((k_broad<256>), grid, dim3(256), sl.broad_bytes(), stream, worlds, nworlds, sl.broad_cap());
    if (stages & STAGE_NARROW) {
      hipLaunchKernelGGL((k_pair_offsets<1024>), dim3(1), dim3(1024), 0, stream, worlds, nworlds, pair_offs);
      hipLaunchKernelGGL(k_narrow, dim3(narrow_groups), dim3(256), 0, stream, worlds, nworlds, (const uint32_t*)pair_offs);
      if (block == 1024) hipLaunchKernelGGL((k_csort<1024>), grid, dim3(1024), 0, stream, worlds, nworlds);
      else hipLaunchKernelGGL((k_csort<256>), grid, dim3(256), 0, stream, worlds, nworlds);
    }
  }
  if (marks) (void)hipEventRecord(marks[1], stream);
  if (stages & STAGE_PREPARE) {
    const uint32_t pw = sl.prep_words();
    const size_t pb = (size_t)pw * 4;
    if (block == 64) hipLaunchKernelGGL((k_prepare<64>), grid, dim3(64), pb, stream, worlds, nworlds, pw);
    else if (block == 128) hipLaunchKernelGGL((k_prepare<128>), grid, dim3(128), pb, stream, worlds, nworlds, pw);
    else if (block == 1024 && nworlds == 1) {
      // One large world: its many articulations' kinematics and factor over the device, between prepare's halves.
      hipLaunchKernelGGL((k_prepare<1024, 1>), grid, dim3(1024), pb, stream, worlds, nworlds, pw);
      hipLaunchKernelGGL(k_art_factor, dim3(1024), dim3(64), 0, stream, worlds);
      hipLaunchKernelGGL((k_prepare<1024, 2>), grid, dim3(1024), pb, stream, worlds, nworlds, pw);
    } else if (block == 1024) hipLaunchKernelGGL((k_prepare<1024>), grid, dim3(1024), pb, stream, worlds, nworlds, pw);
    else hipLaunchKernelGGL((k_prepare<256>), grid, dim3(256), pb, stream, worlds, nworlds, pw);
  }
  if (marks) (void)hipEventRecord(marks[2], stream);
  if ((stages & STAGE_SOLVE) && block == 1024) {
    // sixteen waves per world, for large worlds: the global-mirror instantiation serves every world
    if (nworlds == 1) {  // one large world: its body predictor over the device first
      hipLaunchKernelGGL(k_solve_pred, dim3(512), dim3(256), 0, stream, worlds);
      hipLaunchKernelGGL((k_solve<1024, false, true, false>), grid, dim3(1024), sl.notw, stream, worlds, nworlds);
    } else {
      hipLaunchKernelGGL((k_solve<1024, false, true>), grid, dim3(1024), sl.notw, stream, worlds, nworlds);
    }
  } else if ((stages & STAGE_SOLVE) && sl.mgl) {
    if (block == 64) hipLaunchKernelGGL((k_solve<64, false, true>), grid, dim3(64), solve_lds, stream, worlds, nworlds);
    else if (block == 128) hipLaunchKernelGGL((k_solve<128, false, true>), grid, dim3(128), solve_lds, stream, worlds, nworlds);
    else hipLaunchKernelGGL((k_solve<256, false, true>), grid, dim3(256), solve_lds, stream, worlds, nworlds);
  } else if (stages & STAGE_SOLVE) {
    if (block == 64) {
      if (twl) hipLaunchKernelGGL((k_solve<64, true>), grid, dim3(64), solve_lds, stream, worlds, nworlds);
      else hipLaunchKernelGGL((k_solve<64, false>), grid, dim3(64), solve_lds, stream, worlds, nworlds);
    } else if (block == 128) {
      if (twl) hipLaunchKernelGGL((k_solve<128, true>), grid, dim3(128), solve_lds, stream, worlds, nworlds);
      else hipLaunchKernelGGL((k_solve<128, false>), grid, dim3(128), solve_lds, stream, worlds, nworlds);
    } else {
      if (twl) hipLaunchKernelGGL((k_solve<256, true>), grid, dim3(256), solve_lds, stream, worlds, nworlds);
      else hipLaunchKernelGGL((k_solve<256, false>), grid, dim3(256), solve_lds, stream, worlds, nworlds);
    }
  }
  if (marks) (void)hipEventRecord(marks[3], stream);
  return hipGetLastError();
}

// ------------------------------------------------------------------------------------------------
// One large world over the whole device: cooperative launches of the stage bodies with a grid team.
// ------------------------------------------------------------------------------------------------
static constexpr int GRID_BS = 512;

hipError_t steel_grid_groups(uint32_t cus, uint32_t* groups) {
  const void* ks[] = {(const void*)k_broad_grid<GRID_BS>, (const void*)k_csort_grid<GRID_BS>, (const void*)k_prepare_grid<GRID_BS>,
                      (const void*)k_solve_grid<GRID_BS>};
  int m = 1 << 20;
  for (const void* k : ks) {
    int x = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&x, k, GRID_BS, 0);
    if (e != hipSuccess) return e;
    m = x < m ? x : m;
  }
  if (m < 1) return hipErrorInvalidConfiguration;
  uint32_t gr = (uint32_t)m * cus;
  *groups = gr < GRID_MAX_GROUPS ? gr : GRID_MAX_GROUPS;
  return hipSuccess;
}

hipError_t steel_launch_substep_grid(const WorldDev* world, uint32_t groups, uint32_t* g, uint32_t* pair_offs, uint32_t narrow_groups,
                                     hipStream_t stream, uint32_t stages, hipEvent_t* marks) {
  void* args[] = {(void*)&world, (void*)&g};
  const dim3 grid(groups), block(GRID_BS);
  hipError_t e = hipSuccess;
  if (marks) (void)hipEventRecord(marks[0], stream);
  if (stages & STAGE_COLLIDE) {
    e = hipLaunchCooperativeKernel((const void*)k_broad_grid<GRID_BS>, grid, block, args, 0, stream);
    if (e != hipSuccess) return e;
    if (stages & STAGE_NARROW) {
      hipLaunchKernelGGL((k_pair_offsets<1024>), dim3(1), dim3(1024), 0, stream, world, 1u, pair_offs);
      hipLaunchKernelGGL(k_narrow, dim3(narrow_groups), dim3(256), 0, stream, world, 1u, (const uint32_t*)pair_offs);
      e = hipLaunchCooperativeKernel((const void*)k_csort_grid<GRID_BS>, grid, block, args, 0, stream);
      if (e != hipSuccess) return e;
    }
  }
  if (marks) (void)hipEventRecord(marks[1], stream);
  if (stages & STAGE_PREPARE) {
    e = hipLaunchCooperativeKernel((const void*)k_prepare_grid<GRID_BS>, grid, block, args, 0, stream);
    if (e != hipSuccess) return e;
  }
  if (marks) (void)hipEventRecord(marks[2], stream);
  if (stages & STAGE_SOLVE) {
    e = hipLaunchCooperativeKernel((const void*)k_solve_grid<GRID_BS>, grid, block, args, 0, stream);
    if (e != hipSuccess) return e;
  }
  if (marks) (void)hipEventRecord(marks[3], stream);
  return hipGetLastError();
}
