// Collision stages of one substep for one world, run by one workgroup:
//   shape poses + AABBs (wgsl.js:117-150),
//   sort-and-sweep broadphase with the reference's all-pairs predicate (wgsl.js:154-177),
//   narrowphase, one wave64 per candidate pair (wgsl.js:182-1215),
//   canonical contact order by an LDS radix sort (wgsl.js:3895-3980).
#pragma once
#include "steel_world.cuh"

struct FaceQd { float sep; uint32_t face; };
struct EdgeQd { float sep; f3 axis, a0, a1, b0, b1; uint32_t id; };
struct SegQd { float d2; f3 ps, ph; uint32_t face, kind; };
struct Geod { uint32_t t; f3 s; uint32_t h; };
struct Trkd { f3 r, p, n; float a2; bool np; };
struct Projd { float d; f3 n; };
struct Mand { f3 p[5]; float d[5]; uint32_t n; };

DEV uint32_t lane64() { return __lane_id(); }
// Tie order of a lane-strided search in the reference: candidate index f is visited by lane f % 32 at
// step f / 32; the first maximum within a lane and the lowest lane across lanes win.
DEV uint32_t tie_key(uint32_t f) { return ((f & 31u) << 16) | (f >> 5); }
// Cross-lane argmax of (sep desc, key asc); every lane returns the same winning lane.
DEV int wave_argmax(float sep, uint32_t key) {
  int win = (int)lane64();
  for (int off = 32; off >= 1; off >>= 1) {
    float s2 = __shfl_xor(sep, off, 64);
    uint32_t k2 = __shfl_xor(key, off, 64);
    int w2 = __shfl_xor(win, off, 64);
    if (s2 > sep || (s2 == sep && k2 < key)) { sep = s2; key = k2; win = w2; }
  }
  return win;
}
DEV f3 shfl3(f3 v, int src) { return mk3(__shfl(v.x, src, 64), __shfl(v.y, src, 64), __shfl(v.z, src, 64)); }

// Per-wave LDS work area of the serial (lane 0) manifold code: the two clip polygons, caliper indices and
// the manifold; private arrays indexed at run time would otherwise live in scratch memory.
struct ManifoldLds { f2 A[10], B[10]; int idx[4]; Mand m; };

struct Collider {
  WC w;
  uint32_t* s_nc;      // the world's contact counter
  ManifoldLds* mlds;   // this wave's LDS work area

  DEV float4 shape_rec(uint32_t s, uint32_t k) const { return w.sf(w.so(s) + k); }
  DEV Geod geo(uint32_t s) const {
    float4 gg = shape_rec(s, 6);
    Geod r; r.t = tu(gg.x); r.s = mk3(gg.y, gg.z, gg.w); r.h = tu(shape_rec(s, 0).y);
    if (r.t == GT_SPHERE || r.t == GT_CAPSULE) r.s.x = 1e-4f;
    return r;
  }
  // ---- lane-strided searches (stride 64, tie order of the 32-lane reference)
  DEV void face_query_lane(uint32_t ha, uint32_t hb, q4 rq, f3 t, uint32_t lane, float& bsep, uint32_t& bface) const {
    bsep = -1e30f; bface = 0;
    uint32_t fo = w.hfo(ha), nf = w.hnf(ha), nv = w.hnv(hb);
    for (uint32_t f = lane; f < nf; f += 64) {
      q4 pl = w.fplane(fo + f);
      f3 nb = qinvrot(rq, xyz(pl));
      float m = 1e30f;
      for (uint32_t v = 0; v < nv; v++) m = mn(m, dot(nb, w.hv(hb, v)));
      float s = m + dot(xyz(pl), t) - pl.w;
      if (s > bsep) { bsep = s; bface = f; }
    }
  }
  DEV FaceQd face_query(uint32_t ha, uint32_t hb, q4 rq, f3 t, uint32_t lane) const {
    float s; uint32_t f;
    face_query_lane(ha, hb, rq, t, lane, s, f);
    uint32_t key = s > -1e30f ? tie_key(f) : (0x80000000u | lane);
    int win = wave_argmax(s, key);
    FaceQd r; r.sep = __shfl(s, win, 64); r.face = __shfl(f, win, 64);
    return r;
  }
  DEV float support_min(uint32_t h, f3 n) const { float m = 1e30f; for (uint32_t v = 0; v < w.hnv(h); v++) m = mn(m, dot(n, w.hv(h, v))); return m; }
  DEV float support_max(uint32_t h, f3 n) const { float m = -1e30f; for (uint32_t v = 0; v < w.hnv(h); v++) m = mx(m, dot(n, w.hv(h, v))); return m; }
  DEV EdgeQd edge_query(uint32_t ha, uint32_t hb, q4 rq, f3 t, uint32_t lane) const {
    EdgeQd best;
    best.sep = -1e30f; best.axis = best.a0 = best.a1 = best.b0 = best.b1 = mk3(0, 0, 0); best.id = 0;
    uint32_t bk = 0;
    bool have = false;
    uint32_t foa = w.hfo(ha), fob = w.hfo(hb);
    f3 cena = xyz(w.sf(w.h(W_SF_HAABB) + ha * 2));
    f3 cenb = qrot(rq, xyz(w.sf(w.h(W_SF_HAABB) + hb * 2))) + t;
    uint32_t nea = w.hne(ha), neb = w.hne(hb);
    for (uint32_t k = lane; k < nea * neb; k += 64) {
      uint4 ea_ = w.hedge(ha, k / neb), eb_ = w.hedge(hb, k % neb);
      f3 a = xyz(w.fplane(foa + ea_.z)), b = xyz(w.fplane(foa + ea_.w));
      f3 c = -qrot(rq, xyz(w.fplane(fob + eb_.z))), dd = -qrot(rq, xyz(w.fplane(fob + eb_.w)));
      f3 bxa = cross(b, a), dxc = cross(dd, c);
      float cba = dot(c, bxa), dba = dot(dd, bxa), adc = dot(a, dxc), bdc = dot(b, dxc);
      if (cba * dba > 1e-3f || adc * bdc > 1e-3f || cba * bdc < -1e-3f) continue;
      f3 a0 = w.hv(ha, ea_.x), a1 = w.hv(ha, ea_.y);
      f3 ea = a1 - a0;
      f3 b0 = qrot(rq, w.hv(hb, eb_.x)) + t, b1 = qrot(rq, w.hv(hb, eb_.y)) + t;
      f3 ax = cross(ea, b1 - b0);
      float l = len(ax);
      if (l < 1e-6f * len(ea) * len(b1 - b0)) continue;
      ax = ax / l;
      if (dot(ax, cenb - cena) < 0.0f) ax = -ax;
      float amax = support_max(ha, ax);
      float bmin = dot(ax, t) + support_min(hb, qinvrot(rq, ax));
      float s = bmin - amax;
      float tol = 1e-4f + 1e-3f * (len(ea) + len(b1 - b0));
      if (dot(ax, a0) < amax - tol || dot(ax, a1) < amax - tol || dot(ax, b0) > bmin + tol || dot(ax, b1) > bmin + tol) continue;
      if (s > best.sep) {
        best.sep = s; best.axis = ax; best.a0 = a0; best.a1 = a1; best.b0 = b0; best.b1 = b1;
        best.id = (ea_.x << 24) ^ (ea_.y << 16) ^ (eb_.x << 8) ^ eb_.y;
        bk = k; have = true;
      }
    }
    uint32_t key = have ? tie_key(bk) : (0x80000000u | lane);
    int win = wave_argmax(best.sep, key);
    EdgeQd r;
    r.sep = __shfl(best.sep, win, 64);
    r.axis = shfl3(best.axis, win); r.a0 = shfl3(best.a0, win); r.a1 = shfl3(best.a1, win);
    r.b0 = shfl3(best.b0, win); r.b1 = shfl3(best.b1, win);
    r.id = __shfl(best.id, win, 64);
    return r;
  }
  static DEV float combine(float a, float b, uint32_t mode) {
    if (mode == 1) return mn(a, b);
    if (mode == 2) return a * b;
    if (mode == 3) return mx(a, b);
    return 0.5f * (a + b);
  }
  DEV void emit(uint32_t sa, uint32_t sb, f3 p, f3 n, float phi, uint32_t key) const {  // wgsl.js:269-307
    float4 A0 = shape_rec(sa, 0), B0 = shape_rec(sb, 0);
    int ba = ti(A0.x), bb = ti(B0.x);
    float4 ma = shape_rec(sa, 4), mb = shape_rec(sb, 4);
    float mu = combine(A0.z, B0.z, w.h(W_FRICTION_COMBINE));
    bool art_c = (ba >= 0 && w.body_kind(ba) >= 2) || (bb >= 0 && w.body_kind(bb) >= 2);
    float mt = 0.0f, mr = 0.0f;
    if (art_c) {
      if (w.h(W_ANG_ART)) { mt = combine(ma.y, mb.y, w.h(W_FRICTION_COMBINE)); mr = combine(ma.z, mb.z, w.h(W_FRICTION_COMBINE)); }
    } else {
      mt = mx(ma.w >= 4.0f ? ma.y : 0.0f, mb.w >= 4.0f ? mb.y : 0.0f);
      mr = mx(ma.w >= 6.0f ? ma.z : 0.0f, mb.w >= 6.0f ? mb.z : 0.0f);
    }
    float trd = art_c ? mx(shape_rec(sa, 3).w, shape_rec(sb, 3).w) : 0.0f;
    f3 mat = mk3(combine(clampf_(ma.x, 0.0f, 1.0f), clampf_(mb.x, 0.0f, 1.0f), w.h(W_RESTITUTION_COMBINE)), mt, mr);
    int a = ba, b = bb;
    f3 nn = n;
    float smu0 = A0.z, smu1 = B0.z;
    if (a > b) { a = bb; b = ba; nn = -n; float t = smu0; smu0 = smu1; smu1 = t; }
    uint32_t k = atomicAdd(s_nc, 1u);
    if (k >= w.MAXC()) return;
    float4* c = w.d.craw + (size_t)k * 6;
    float ka = shape_rec(sa, 5).x, kb = shape_rec(sb, 5).x;
    float kc = (ka > 0.0f && kb > 0.0f) ? ka * kb / (ka + kb) : 0.0f;
    float ga = shape_rec(sa, 5).z, gb = shape_rec(sb, 5).z;
    if (ba > bb) { float t = ga; ga = gb; gb = t; }
    c[0] = F4(p, phi);
    c[1] = F4(nn, mu);
    c[2] = make_float4((float)a, (float)b, (float)(key & 0xffffu), (float)(key >> 16));
    c[3] = F4(mat, trd);
    c[4] = make_float4(kc, shape_rec(sa, 5).y + shape_rec(sb, 5).y, ga, gb);
    c[5] = make_float4(smu0, smu1, 0.0f, 0.0f);
  }
  static DEV void closest_seg(f3 a0, f3 a1, f3 b0, f3 b1, f3 out[2]) {
    f3 d1 = a1 - a0, d2 = b1 - b0, r = a0 - b0;
    float a = dot(d1, d1), e = dot(d2, d2), f = dot(d2, r);
    float s = 0.0f, t = 0.0f;
    if (a <= 1e-12f && e <= 1e-12f) { out[0] = a0; out[1] = b0; return; }
    if (a <= 1e-12f) t = clampf_(f / e, 0.0f, 1.0f);
    else {
      float c = dot(d1, r);
      if (e <= 1e-12f) s = clampf_(-c / a, 0.0f, 1.0f);
      else {
        float b = dot(d1, d2), den = a * e - b * b;
        if (den > 1e-12f) s = clampf_((b * f - c * e) / den, 0.0f, 1.0f);
        t = (b * s + f) / e;
        if (t < 0.0f) { t = 0.0f; s = clampf_(-c / a, 0.0f, 1.0f); }
        else if (t > 1.0f) { t = 1.0f; s = clampf_((b - c) / a, 0.0f, 1.0f); }
      }
    }
    out[0] = a0 + d1 * s; out[1] = b0 + d2 * t;
  }
  static DEV void sphere_sphere(f3 p1, float r1, f3 p2, float r2, q4 out[2]) {
    f3 dv = p2 - p1;
    float l = len(dv);
    f3 n = (l == 0.0f) ? mk3(1, 0, 0) : dv / l;
    float d = l - (r1 + r2);
    out[0] = mk4(p1 + n * (r1 + 0.5f * d), d);
    out[1] = mk4(n, 0.0f);
  }
  DEV void prim_prim(uint32_t sa, uint32_t sb, f3 pa, q4 qa, f3 pb, q4 qb, float thresh) const {
    float4 A = shape_rec(sa, 6), B = shape_rec(sb, 6);
    f3 ua = qrot(qa, mk3(0, 0, A.z)), ub = qrot(qb, mk3(0, 0, B.z));
    uint32_t base = hash_u32(sa * 7919u + sb * 104729u);
    q4 c[2][2];
    uint32_t nc = 1;
    if (tu(A.x) == GT_SPHERE) {
      f3 a = pb - ub, ab = (pb + ub) - a;
      sphere_sphere(pa, A.y, a + clampf_(dot(pa - a, ab) / (dot(ab, ab) + 1e-6f), 0.0f, 1.0f) * ab, B.y, c[0]);
    } else {
      f3 dif = pa - pb;
      float ma = dot(ua, ua), mb = -dot(ua, ub), mc = dot(ub, ub);
      float u = -dot(ua, dif), v = dot(ub, dif);
      float det = ma * mc - mb * mb;
      if (fabsf(det) >= 1e-15f) {
        float inv_det = 1.0f / det;
        float x1 = (mc * u - mb * v) * inv_det, x2 = (ma * v - mb * u) * inv_det;
        if (x1 > 1.0f) { x1 = 1.0f; x2 = (v - mb) / mc; } else if (x1 < -1.0f) { x1 = -1.0f; x2 = (v + mb) / mc; }
        if (x2 > 1.0f) { x2 = 1.0f; x1 = clampf_((u - mb) / ma, -1.0f, 1.0f); } else if (x2 < -1.0f) { x2 = -1.0f; x1 = clampf_((u + mb) / ma, -1.0f, 1.0f); }
        sphere_sphere(pa + ua * x1, A.y, pb + ub * x2, B.y, c[0]);
      } else {
        sphere_sphere(pa + ua, A.y, pb + ub * clampf_((v - mb) / mc, -1.0f, 1.0f), B.y, c[0]);
        sphere_sphere(pa - ua, A.y, pb + ub * clampf_((v + mb) / mc, -1.0f, 1.0f), B.y, c[1]);
        c[1][1] = c[0][1];
        nc = 2;
      }
    }
    for (uint32_t i = 0; i < nc; i++)
      if (c[i][0].w <= thresh) emit(sa, sb, xyz(c[i][0]), xyz(c[0][1]), c[i][0].w, hash_u32(base ^ (i * 977u + 13u)));
  }
  DEV void plane_contacts(uint32_t pl, uint32_t os, f3 pp, q4 pq, f3 op, q4 oq, float thresh) const {
    f3 n = qrot(pq, mk3(0, 0, 1));
    uint32_t sa = pl, sb = os;
    uint32_t base = hash_u32(sa * 7919u + sb * 104729u);
    float4 g = shape_rec(os, 6);
    uint32_t gt = tu(g.x);
    float d[4] = {1e30f, 1e30f, 1e30f, 1e30f};
    f3 x[4];
    uint32_t nc = 0;
    if (gt == GT_SPHERE || gt == GT_CAPSULE) {
      f3 u = qrot(oq, mk3(0, 0, g.z));
      uint32_t cnt = gt == GT_SPHERE ? 1 : 2;
      for (uint32_t i = 0; i < cnt; i++) {
        f3 e = op + (i == 1 ? -u : u);
        d[i] = dot(n, e - pp) - g.y;
        x[i] = e - n * (g.y + 0.5f * d[i]);
      }
      nc = cnt;
    } else if (gt == GT_BOX) {
      float cd0 = dot(op - pp, n);
      uint32_t worst = 0;
      m3 rot = qmat(oq);
      for (uint32_t i = 0; i < 8; i++) {
        f3 c = mul(rot, mk3((i & 1) ? g.y : -g.y, (i & 2) ? g.z : -g.z, (i & 4) ? g.w : -g.w));
        float cd = cd0 + dot(n, c);
        if (cd > thresh) continue;
        f3 cp = c + op - (0.5f * n) * cd;
        if (nc < 4) {
          d[nc] = cd; x[nc] = cp;
          if (nc == 0 || cd > d[worst]) worst = nc;
          nc++;
        } else if (cd < d[worst]) {
          d[worst] = cd; x[worst] = cp;
          worst = 0;
          for (uint32_t k = 1; k < 4; k++) if (d[k] > d[worst]) worst = k;
        }
      }
    } else if (gt == GT_CYL) {
      // Deepest rim point, then three cap points (within 22.5 degrees of upright) or the opposite rim and the
      // cap's side point (rolling).
      f3 ax = qrot(oq, mk3(0, 0, 1));
      float dna = dot(n, ax);
      if (dna > 0.0f) { ax = -ax; dna = -dna; }
      f3 cap = op + ax * g.z;
      f3 pa = -n + ax * dna;
      float l2 = dot(pa, pa);
      bool has = l2 > 1e-10f;
      if (has) pa = pa * (1.0f / sqrtf(l2));
      bool flat = -dna >= 0.9238795325112867f;
      f3 pf = mk3(0, 0, 0);
      if (flat || !has) {
        f3 rf = mk3(1, 0, 0);
        if (fabsf(dot(ax, rf)) > 0.9f) rf = mk3(0, 1, 0);
        pf = nrm(rf - ax * dot(ax, rf));
      }
      f3 dp = has ? pa : pf;
      f3 dpt = cap + dp * g.y;
      d[0] = dot(dpt - pp, n);
      x[0] = dpt - n * (d[0] * 0.5f);
      nc = 1;
      float mt = 0.01f * mx(g.y, g.z);
      f3 cand[3];
      uint32_t ncand = 0;
      if (flat) {
        f3 u = pf * g.y, v = cross(ax, pf) * g.y;
        cand[0] = cap + u;
        cand[1] = cap - 0.5f * u + 0.8660254f * v;
        cand[2] = cap - 0.5f * u - 0.8660254f * v;
        ncand = 3;
      } else {
        f3 u = dp * g.y, v = cross(ax, dp) * g.y;
        cand[0] = op - ax * g.z + u;
        float dpv = dot(cap + v - pp, n), dnv = dot(cap - v - pp, n);
        cand[1] = dpv <= dnv ? cap + v : cap - v;
        ncand = 2;
      }
      for (uint32_t k = 0; k < ncand; k++) {
        float dk = dot(cand[k] - pp, n);
        f3 pk = cand[k] - n * (dk * 0.5f);
        if (nc < 4 && dot(pk - x[0], pk - x[0]) > mt * mt) { d[nc] = dk; x[nc] = pk; nc++; }
      }
    }
    for (uint32_t i = 0; i < nc; i++)
      if (d[i] <= thresh) emit(sa, sb, x[i], n, d[i], hash_u32(base ^ (i + 7u)));
  }
  DEV FaceQd face_query_seg(uint32_t h, f3 e0, f3 e1, float r, uint32_t lane) const {
    float bsep = -1e30f; uint32_t bface = 0;
    uint32_t fo = w.hfo(h);
    for (uint32_t f = lane; f < w.hnf(h); f += 64) {
      q4 pl = w.fplane(fo + f);
      float s = mn(dot(xyz(pl), e0), dot(xyz(pl), e1)) - pl.w - r;
      if (s > bsep) { bsep = s; bface = f; }
    }
    uint32_t key = bsep > -1e30f ? tie_key(bface) : (0x80000000u | lane);
    int win = wave_argmax(bsep, key);
    FaceQd q; q.sep = __shfl(bsep, win, 64); q.face = __shfl(bface, win, 64);
    return q;
  }
  DEV bool seg_hits_hull(uint32_t h, f3 e0, f3 e1) const {
    float t0 = 0.0f, t1 = 1.0f;
    for (uint32_t f = 0; f < w.hnf(h); f++) {
      q4 pl = w.fplane(w.hfo(h) + f);
      float d0 = dot(xyz(pl), e0) - pl.w, d1 = dot(xyz(pl), e1) - pl.w;
      if (d0 > 0.0f && d1 > 0.0f) return false;
      if (d0 > 0.0f) t0 = mx(t0, d0 / (d0 - d1));
      else if (d1 > 0.0f) t1 = mn(t1, d0 / (d0 - d1));
    }
    return t0 <= t1;
  }
  // Closest features of a segment and a hull (wgsl.js:568-590): the per-lane rule keeps face ties, so this
  // search runs on the reference's 32 lanes exactly (lanes 32..63 idle).
  DEV SegQd seg_hull_query(uint32_t h, f3 e0, f3 e1, uint32_t lane) const {
    SegQd best;
    best.d2 = 1e30f; best.ps = best.ph = mk3(0, 0, 0); best.face = 0; best.kind = 1;
    uint32_t fo = w.hfo(h);
    if (lane < NLANE) {
      for (uint32_t f = lane; f < w.hnf(h); f += NLANE) {
        q4 pl = w.fplane(fo + f);
        uint32_t sr = w.fstart(fo + f), cr = w.fcount(fo + f);
        bool in0 = dot(xyz(pl), e0) > pl.w, in1 = dot(xyz(pl), e1) > pl.w;
        for (uint32_t k = 0; k < cr; k++) {
          f3 r0 = w.hv(h, w.fvert(sr + k)), r1 = w.hv(h, w.fvert(sr + (k + 1) % cr));
          f3 out = cross(r1 - r0, xyz(pl));
          if (dot(e0 - r0, out) > 0.0f) in0 = false;
          if (dot(e1 - r0, out) > 0.0f) in1 = false;
          f3 c[2];
          closest_seg(e0, e1, r0, r1, c);
          float d2 = dot(c[0] - c[1], c[0] - c[1]);
          if (d2 < best.d2) { best.d2 = d2; best.ps = c[0]; best.ph = c[1]; best.face = f; best.kind = 1; }
        }
        if (in0) { float sd = dot(xyz(pl), e0) - pl.w; if (sd * sd <= best.d2 * 1.0001f) { best.d2 = sd * sd; best.ps = e0; best.ph = e0 - xyz(pl) * sd; best.face = f; best.kind = 0; } }
        if (in1) { float sd = dot(xyz(pl), e1) - pl.w; if (sd * sd <= best.d2 * 1.0001f) { best.d2 = sd * sd; best.ps = e1; best.ph = e1 - xyz(pl) * sd; best.face = f; best.kind = 0; } }
      }
    }
    // lowest lane among the minima (lanes 32..63 never win: their key is larger)
    float neg = lane < NLANE ? -best.d2 : -3e38f;
    int win = wave_argmax(neg, lane);
    SegQd r;
    r.d2 = __shfl(best.d2, win, 64); r.ps = shfl3(best.ps, win); r.ph = shfl3(best.ph, win);
    r.face = __shfl(best.face, win, 64); r.kind = __shfl(best.kind, win, 64);
    return r;
  }
  DEV void seg_on_face(uint32_t hs, uint32_t ps, uint32_t h, uint32_t f, f3 e0, f3 e1, float r, f3 wp, q4 wq, float thresh, bool flip) const {
    q4 pl = w.fplane(w.hfo(h) + f);
    float t0 = 0.0f, t1 = 1.0f;
    uint32_t sr = w.fstart(w.hfo(h) + f), cr = w.fcount(w.hfo(h) + f);
    for (uint32_t k = 0; k < cr; k++) {
      f3 r0 = w.hv(h, w.fvert(sr + k)), r1 = w.hv(h, w.fvert(sr + (k + 1) % cr));
      f3 out = cross(r1 - r0, xyz(pl));
      float d0 = dot(e0 - r0, out), d1 = dot(e1 - r0, out);
      if (d0 > 0.0f && d1 > 0.0f) { t0 = 2.0f; break; }
      if (d0 > 0.0f) t0 = mx(t0, d0 / (d0 - d1));
      if (d1 > 0.0f) t1 = mn(t1, d0 / (d0 - d1));
    }
    float ts[2] = {t0, t1};
    uint32_t n = 2;
    if (t0 > t1) { ts[0] = dot(xyz(pl), e0) <= dot(xyz(pl), e1) ? 0.0f : 1.0f; n = 1; }
    else if (t1 - t0 < 1e-4f || len(e1 - e0) < 1e-7f) n = 1;
    f3 nw = qrot(wq, xyz(pl));
    uint32_t sa = flip ? ps : hs, sb = flip ? hs : ps;
    uint32_t base = hash_u32(sa * 7919u + sb * 104729u);
    for (uint32_t i = 0; i < n; i++) {
      f3 x = e0 + (e1 - e0) * ts[i];
      float sd = dot(xyz(pl), x) - pl.w;
      float phi = sd - r;
      if (phi > thresh) continue;
      f3 mid = x - xyz(pl) * (0.5f * (sd + r));
      emit(sa, sb, wp + qrot(wq, mid), flip ? -nw : nw, phi, hash_u32(base ^ (f * 131u + i * 7u + 3u)));
    }
  }
  // ---- generic manifold (wgsl.js:625-977), lane 0
  DEV f3 gsup(const Geod& g, f3 dv) const {
    if (g.t == GT_BOX) {
      float th = 1e-10f * mx(fabsf(dv.x), mx(fabsf(dv.y), fabsf(dv.z)));
      return mk3(dv.x >= -th ? g.s.x : -g.s.x, dv.y >= -th ? g.s.y : -g.s.y, dv.z >= -th ? g.s.z : -g.s.z);
    }
    if (g.t == GT_SPHERE || g.t == GT_CAPSULE) {
      float l2 = dot(dv, dv);
      f3 p = (l2 > 1e-12f ? dv * rsq(l2) : mk3(1, 0, 0)) * g.s.x;
      if (g.t == GT_CAPSULE) p.z += dv.z >= 0.0f ? g.s.y : -g.s.y;
      return p;
    }
    if (g.t == GT_CYL) {  // rim point toward dv's xy, on the cap dv.z points to
      f3 xy = mk3(dv.x, dv.y, 0.0f);
      float l2 = dot(xy, xy);
      f3 p = l2 > 1e-12f ? xy * rsq(l2) * g.s.x : mk3(g.s.x, 0, 0);
      if (dv.z > 0.0f) p.z = g.s.y; else if (dv.z < 0.0f) p.z = -g.s.y;
      return p;
    }
    float best = -1e10f;
    f3 bp = mk3(0, 0, 0);
    for (uint32_t v = 0; v < w.hnv(g.h); v++) { f3 x = w.hv(g.h, v); float s = dot(x, dv); if (s > best) { best = s; bp = x; } }
    return bp;
  }
  DEV f3 fsup(const Geod& g, f3 dv, f3 n) const {
    f3 p = gsup(g, dv);
    if ((g.t == GT_CAPSULE || (g.t == GT_CYL && g.s.z == 0.0f)) && fabsf(n.z) <= 1e-4f) {
      f3 rad = nrm(mk3(n.x, n.y, 0)) * g.s.x;
      p = mk3(rad.x, rad.y, dv.z >= 0.0f ? g.s.y : -g.s.y);
    }
    return p;
  }
  static DEV bool planar_geo(const Geod& g) { return g.t == GT_BOX || g.t == GT_CONVEX; }
  static DEV Trkd trk_add(Trkd t, f3 x, uint32_t id, bool chk) {
    const float MEPS = 1e-5f;
    Trkd o = t;
    if (id == 0) { o.r = x; o.a2 = 0.0f; o.n = mk3(0, 0, 0); o.np = false; }
    else if (id == 1) o.p = x;
    else {
      if (chk && o.a2 > 0.0f) { float off = dot(o.n, x - o.r); if (off * off > MEPS * MEPS * o.a2) o.np = true; }
      f3 c = cross(o.p - o.r, x - o.r);
      float a2 = dot(c, c);
      if (a2 > o.a2) { o.a2 = a2; o.n = c; }
      o.p = x;
    }
    return o;
  }
  static DEV bool trk_anchor(const Trkd& t, f3 x) {
    const float MEPS = 1e-5f;
    float off = dot(t.n, x - t.r);
    return !t.np && off * off <= MEPS * MEPS * t.a2;
  }
  static DEV f3 line_proj_n(f3 dir, f3 rn) {
    f3 nn = cross(cross(dir, rn), dir);
    float l = len(nn);
    return l > 1e-12f ? nn * (1.0f / l) : rn;
  }
  static DEV void make_proj(const Trkd& ta, f3 pa, const Trkd& tb, f3 pb, f3 n, Projd out[2]) {
    Projd A, B;
    A.d = 0.0f; A.n = mk3(0, 0, 0); B = A;
    if (ta.a2 == 0.0f && tb.a2 == 0.0f) {
      A.n = line_proj_n(ta.p - ta.r, n); A.d = -dot(0.5f * (ta.r + ta.p), A.n);
      B.n = line_proj_n(tb.p - tb.r, n); B.d = -dot(0.5f * (tb.r + tb.p), B.n);
      out[0] = A; out[1] = B; return;
    }
    if (ta.a2 > 0.0f) { A.n = ta.n * (1.0f / sqrtf(mx(1e-12f, ta.a2))); A.d = -dot(pa, A.n); }
    if (tb.a2 > 0.0f) { B.n = tb.n * (1.0f / sqrtf(mx(1e-12f, tb.a2))); B.d = -dot(pb, B.n); }
    if (ta.a2 == 0.0f) { A.n = line_proj_n(ta.p - ta.r, B.n); A.d = -dot(0.5f * (ta.r + ta.p), A.n); }
    if (tb.a2 == 0.0f) { B.n = line_proj_n(tb.p - tb.r, A.n); B.d = -dot(0.5f * (tb.r + tb.p), B.n); }
    out[0] = A; out[1] = B;
  }
  static DEV f3 proj_on(const Projd& pr, f3 x, f3 n) {
    float den = dot(n, pr.n);
    if (fabsf(den) < 1e-12f) return x;
    return x + n * (-(dot(x, pr.n) + pr.d) / den);
  }
  static DEV float sarea(f2 a, f2 b, f2 q) { return (b.x - a.x) * (q.y - a.y) - (b.y - a.y) * (q.x - a.x); }
  static DEV bool add_nodup(f2* arr, int cnt, f2 v) {
    const float MEPS = 1e-5f;
    if (cnt > 0) { f2 d = arr[0] - v; if (dot(d, d) < MEPS * MEPS) return false; }
    if (cnt > 1) { f2 d = arr[cnt - 1] - v; if (dot(d, d) < MEPS * MEPS) return false; }
    arr[cnt] = v;
    return true;
  }
  static DEV int trim_in_place(f2 s0, f2 s1, f2* lp, int cnt0) {
    int cnt = cnt0;
    if (cnt < 3) return cnt;
    f2 ia = mk2(0, 0), ib = mk2(0, 0);
    int ca = -1, cb = -1;
    bool keep = false;
    bool prev_out = sarea(s0, s1, lp[0]) <= 0.0f;
    for (int i = 0; i < cnt; i++) {
      int nx = (i + 1) % cnt;
      bool out = sarea(s0, s1, lp[nx]) <= 0.0f;
      if (out != prev_out) {
        f2 a = lp[i], b = lp[nx];
        float sa = sarea(s0, s1, a), sb = sarea(s0, s1, b);
        float t = fabsf(sa) / fabsf(sa - sb);
        f2 x = (1.0f - t) * a + t * b;
        if (ca < 0) { ca = i; keep = !prev_out; ia = x; } else { cb = i; ib = x; }
      }
      prev_out = out;
    }
    if (ca >= 0 && cb >= 0) {
      int li = -1, ncnt = cnt, i = 0;
      for (int guard = 0; guard < 24 && i < cnt; guard++) {
        if (keep) { li++; lp[li] = lp[i]; }
        if (i == ca || i == cb) {
          f2 pt = (i == ca) ? ia : ib;
          if (li == i && !keep) {
            li++;
            for (int k = min(ncnt, 9); k > li; k--) lp[k] = lp[k - 1];
            lp[li] = pt;
            ncnt++; i++; cb++; cnt++;
          } else { li++; lp[li] = pt; }
          keep = !keep;
        }
        i++;
      }
      return li + 1;
    }
    return prev_out ? 0 : cnt;
  }
  static DEV int trim_all(f2* tp, int tn, f2* lp, int ln) {
    const float MEPS = 1e-5f;
    if (tn <= 1) return min(1, ln);
    if (tn == 2 || ln == 2) {
      f2 st = lp[0], en = lp[1];
      if (tn == 2) { st = tp[0]; en = tp[1]; }
      f2 dir = en - st;
      float l2 = dot(dir, dir);
      if (l2 <= MEPS * MEPS) return 0;
      float lo_ = 0.0f, hi = 1.0f;
      if (tn == 2 && ln == 2) {
        f2 os = lp[0], od = lp[1] - os;
        float cr = dir.x * od.y - dir.y * od.x;
        if (fabsf(cr) > 1e-6f * sqrtf(l2 * dot(od, od))) return 0;
        float t0 = dot(os - st, dir) / l2, t1 = dot(lp[1] - st, dir) / l2;
        lo_ = mx(lo_, mn(t0, t1)); hi = mn(hi, mx(t0, t1));
      } else {
        bool useLoop = tn == 2;
        int pc = useLoop ? ln : tn;
        for (int i = 0; i < pc; i++) {
          f2 e0 = useLoop ? lp[i] : tp[i];
          f2 e1 = useLoop ? lp[(i + 1) % pc] : tp[(i + 1) % pc];
          float a = sarea(e0, e1, st), b = sarea(e0, e1, en);
          if (a < 0.0f && b < 0.0f) return 0;
          if (a < 0.0f) lo_ = mx(lo_, a / (a - b)); else if (b < 0.0f) hi = mn(hi, a / (a - b));
        }
      }
      if (hi <= lo_) return 0;
      lp[0] = st + lo_ * dir; lp[1] = st + hi * dir;
      return 2;
    }
    int c = ln;
    f2 t0 = tp[0];
    for (int i = 0; i < tn; i++) {
      f2 s1 = (i == tn - 1) ? t0 : tp[i + 1];
      c = trim_in_place(tp[i], s1, lp, c);
    }
    return c;
  }
  static DEV int remove_zero_edges(f2* lp, int cnt) {
    const float MEPS = 1e-5f;
    if (cnt < 2) return 0;
    int wi = 0;
    for (int r = 1; r < cnt; r++) { f2 d = lp[r] - lp[wi]; if (dot(d, d) > MEPS * MEPS) { wi++; lp[wi] = lp[r]; } }
    int n = wi + 1;
    if (wi > 0) { f2 d = lp[wi] - lp[0]; if (dot(d, d) < MEPS * MEPS) n = wi; }
    return n < 2 ? 0 : n;
  }
  static DEV void calipers(f2* h, int n, int out[4]) {
    int p1 = 0, p3 = 1;
    f2 d0 = h[0] - h[1];
    float md = dot(d0, d0);
    const float te = 1e-3f;
    int j = 1;
    for (int i = 0; i < n; i++) {
      f2 hi = h[i], hi1 = h[(i + 1) % n];
      for (int g = 0; g < 16; g++) {
        if (sarea(hi, hi1, h[(j + 1) % n]) > sarea(hi, hi1, h[j])) j = (j + 1) % n; else break;
      }
      f2 hj = h[j];
      f2 a = hi - hj;
      if (dot(a, a) > md * (1.0f + te)) { md = dot(a, a); p1 = i; p3 = j; }
      f2 b = hi1 - hj;
      if (dot(b, b) > md * (1.0f + te)) { md = dot(b, b); p1 = (i + 1) % n; p3 = j; }
    }
    int p2 = 0, p4 = 0;
    float m1 = 0.0f, m2 = 0.0f;
    f2 q1 = h[p1], q3 = h[p3];
    for (int i = 0; i < n; i++) {
      float ar = sarea(q1, q3, h[i]);
      if (ar > m1 * (1.0f + te)) { m1 = ar; p2 = i; } else if (-ar > m2 * (1.0f + te)) { m2 = -ar; p4 = i; }
    }
    out[0] = p1; out[1] = p2; out[2] = p3; out[3] = p4;
  }
  static DEV void onb(f3 n, f3 out[2]) {
    if (n.z < 0.0f) {
      float a = 1.0f / (1.0f - n.z), b = n.x * n.y * a;
      out[0] = mk3(1.0f - n.x * n.x * a, -b, n.x); out[1] = mk3(b, n.y * n.y * a - 1.0f, -n.y);
      return;
    }
    float a = 1.0f / (1.0f + n.z), b = -n.x * n.y * a;
    out[0] = mk3(1.0f - n.x * n.x * a, b, -n.x); out[1] = mk3(b, 1.0f - n.y * n.y * a, -n.y);
  }
  DEV bool plane_supporting(const Geod& g, f3 dir, float dd) const {
    const float MEPS = 1e-5f;
    float off = dot(dir, gsup(g, dir)) + dd;
    return off * off <= MEPS * MEPS * dot(dir, dir);
  }
  DEV void build_manifold(const Geod& ga, const Geod& gb, q4 rq, f3 rt, f3 wa, f3 wb, f3 n) const {
    const float COS_TILT = 0.9993908270190958f, SIN_TILT = 0.03489949670250097f, COS_DEEP = 0.9999984769132877f;
    const float PENTX[5] = {1.0f, 0.30901699437494745f, -0.8090169943749473f, -0.8090169943749476f, 0.30901699437494723f};
    const float PENTY[5] = {0.0f, 0.9510565162951535f, 0.5877852522924732f, -0.587785252292473f, -0.9510565162951536f};
    Mand& out = mlds->m;
    out.n = 0;
    f3 fna = n;
    f3 tb[2];
    onb(n, tb);
    bool pa = planar_geo(ga), pb = planar_geo(gb);
    f3 center = 0.5f * (wa + wb);
    f2* A = mlds->A;
    f2* B = mlds->B;
    int ac = 0, bc = 0;
    Trkd ta;
    ta.r = ta.p = ta.n = mk3(0, 0, 0); ta.a2 = 0.0f; ta.np = false;
    Trkd tbk = ta;
    f3 lnb = qinvrot(rq, -n), lfb = qinvrot(rq, -fna);
    f3 lta = qinvrot(rq, -tb[0]), ltb = qinvrot(rq, -tb[1]);
    for (uint32_t e = 0; e < 5; e++) {
      f2 cs = mk2(PENTX[e], PENTY[e]) * SIN_TILT;
      f3 x = fsup(ga, n * COS_TILT + cs.x * tb[0] + cs.y * tb[1], fna);
      f3 xa = x - center;
      if (add_nodup(A, ac, mk2(dot(tb[0], xa), dot(tb[1], xa)))) { ac++; ta = trk_add(ta, x, (uint32_t)(ac - 1), pa); }
      f3 y = qrot(rq, fsup(gb, lnb * COS_TILT + cs.x * lta + cs.y * ltb, lfb)) + rt;
      f3 yb = y - center;
      if (add_nodup(B, bc, mk2(dot(tb[0], yb), dot(tb[1], yb)))) { bc++; tbk = trk_add(tbk, y, (uint32_t)(bc - 1), pb); }
    }
    uint32_t cnt = 0;
    float ndot = 0.0f;
    if (ac >= 2 && bc >= 2 && !(pa && !trk_anchor(ta, wa)) && !(pb && !trk_anchor(tbk, wb))) {
      Projd pr[2];
      make_proj(ta, wa, tbk, wb, n, pr);
      bool ok = fabsf(dot(n, pr[0].n)) >= COS_TILT && fabsf(dot(n, pr[1].n)) >= COS_TILT;
      if (ok && pa && (ta.a2 == 0.0f || dot(cross(n, pr[0].n), cross(n, pr[0].n)) > 0.0f)) {
        bool flip = dot(pr[0].n, n) < 0.0f;
        ok = plane_supporting(ga, flip ? -pr[0].n : pr[0].n, flip ? -pr[0].d : pr[0].d);
      }
      if (ok && pb && (tbk.a2 == 0.0f || dot(cross(n, pr[1].n), cross(n, pr[1].n)) > 0.0f)) {
        bool flip = dot(pr[1].n, n) > 0.0f;
        f3 db = flip ? -pr[1].n : pr[1].n;
        float dd = flip ? -pr[1].d : pr[1].d;
        ok = plane_supporting(gb, qinvrot(rq, db), dd + dot(db, rt));
      }
      if (ok) {
        ndot = fabsf(dot(pr[0].n, pr[1].n));
        int lc = remove_zero_edges(B, trim_all(A, ac, B, bc));
        if (lc > 1) {
          int* idx = mlds->idx;
          idx[0] = 0; idx[1] = 1; idx[2] = 2; idx[3] = 3;
          if (lc > 4) { calipers(B, lc, idx); lc = 4; }
          for (int i = 0; i < lc; i++) {
            f2 q = B[idx[i]];
            f3 pl = q.x * tb[0] + q.y * tb[1] + center;
            f3 a = proj_on(pr[0], pl, n), b = proj_on(pr[1], pl, n);
            out.p[cnt] = 0.5f * (a + b); out.d[cnt] = dot(b - a, n); cnt++;
          }
        } else ndot = 0.0f;
      }
    }
    if (ndot < COS_DEEP || cnt == 0) { out.p[cnt] = 0.5f * (wa + wb); out.d[cnt] = dot(wb - wa, n); cnt++; }
    out.n = cnt;
  }
  DEV void gen_run(uint32_t sa, uint32_t sb, const Geod& ga, const Geod& gb, f3 pa, q4 qa, q4 rq, f3 rt, f3 wa, f3 wb, f3 n, float thresh) const {
    build_manifold(ga, gb, rq, rt, wa, wb, n);
    const Mand& m = mlds->m;
    f3 nw = qrot(qa, n);
    float ra = (ga.t == GT_SPHERE || ga.t == GT_CAPSULE) ? shape_rec(sa, 6).y : 0.0f;
    float rb = (gb.t == GT_SPHERE || gb.t == GT_CAPSULE) ? shape_rec(sb, 6).y : 0.0f;
    // A cylinder against a discrete shape rolls when its axis is within 2 degrees of the tangent plane: its
    // points are moved onto the line through the cylinder's center across the axis (wgsl.js:955-967).
    auto discrete = [](uint32_t t) { return t == GT_BOX || t == GT_CONVEX || t == GT_PLANE; };
    const bool axA = ga.t == GT_CYL || ga.t == GT_CONE, axB = gb.t == GT_CYL || gb.t == GT_CONE;
    bool roll = false;
    f3 rpos = mk3(0, 0, 0), rn = mk3(0, 0, 0);
    if ((discrete(ga.t) && axB) || (discrete(gb.t) && axA)) {
      const bool onB = discrete(ga.t) && axB;
      f3 ax = onB ? qrot(qmul(qa, rq), mk3(0, 0, 1)) : qrot(qa, mk3(0, 0, 1));
      f3 an = onB ? nw : -nw;
      if (fabsf(dot(ax, an)) <= 0.03489949670250097f) {
        roll = true;
        rpos = onB ? pa + qrot(qa, rt) : pa;
        rn = nrm(cross(ax, an));
      }
    }
    uint32_t base = hash_u32(sa * 7919u + sb * 104729u);
    for (uint32_t i = 0; i < m.n; i++) {
      f3 p = pa + qrot(qa, m.p[i]) + nw * (0.5f * ra) - nw * (0.5f * rb);
      float dd = m.d[i] - ra - rb;
      if (roll) p = p - rn * dot(p - rpos, rn);
      if (dd > thresh) continue;
      emit(sa, sb, p, nw, dd, hash_u32(base ^ (i * 0x9e3779b1u + 0x51u)));
    }
  }
  DEV float hull_out(uint32_t h, f3 x) const {
    float m = -1e30f;
    for (uint32_t f = 0; f < w.hnf(h); f++) { q4 pl = w.fplane(w.hfo(h) + f); m = mx(m, dot(xyz(pl), x) - pl.w); }
    return m;
  }
  DEV f3 ray_face(uint32_t h, f3 o, f3 n, f3 s) const {
    float best = -2.0f;
    q4 pl = mk4(0, 0, 0, 0);
    for (uint32_t f = 0; f < w.hnf(h); f++) { q4 q = w.fplane(w.hfo(h) + f); if (dot(xyz(q), n) > best) { best = dot(xyz(q), n); pl = q; } }
    f3 x = o + n * ((pl.w - dot(xyz(pl), o)) / dot(xyz(pl), n));
    if (fabsf(dot(x - s, n)) > 1e-3f) return o + n * dot(s - o, n);
    return x;
  }
  DEV void sup_witness(const Geod& ga, const Geod& gb, q4 rq, f3 rt, f3 n, bool refA, f3 wo[2]) const {
    f3 sA = gsup(ga, n);
    f3 lb = gsup(gb, qinvrot(rq, -n));
    f3 sB = qrot(rq, lb) + rt;
    f3 wb0 = ray_face(ga.h, sB, n, sA), wb1 = sB;
    f3 wa0 = sA, wa1 = qrot(rq, ray_face(gb.h, qinvrot(rq, sA - rt), qinvrot(rq, -n), lb)) + rt;
    float ob = mx(hull_out(ga.h, wb0), hull_out(gb.h, qinvrot(rq, wb1 - rt)));
    float oa = mx(hull_out(ga.h, wa0), hull_out(gb.h, qinvrot(rq, wa1 - rt)));
    if ((refA && !(oa < ob - 1e-6f)) || (!refA && ob < oa - 1e-6f)) { wo[0] = wb0; wo[1] = wb1; return; }
    wo[0] = wa0; wo[1] = wa1;
  }
  // ---- MPR for hulls of more than 16 vertices (wgsl.js:310-389, 1206-1214); whole-wave, uniform control flow
  // Support point of hull h along dv in the reference's 32-lane subgroup form: lane l < 32 scans vertices l,
  // l + 32, ... keeping its first maximum, and the lowest lane holding the overall maximum wins (lanes 32-63
  // hold no candidate).  Every lane returns the winner.
  DEV f3 sup_lane(uint32_t h, f3 dv, uint32_t lane) const {
    float best = -1e30f;
    f3 bp = mk3(0, 0, 0);
    if (lane < 32)
      for (uint32_t v = lane; v < w.hnv(h); v += 32) {
        f3 x = w.hv(h, v);
        float s = dot(dv, x);
        if (s > best) { best = s; bp = x; }
      }
    float m = best;
    for (int off = 32; off >= 1; off >>= 1) m = mx(m, __shfl_xor(m, off, 64));
    int win = (lane < 32 && best == m) ? (int)lane : 64;
    for (int off = 32; off >= 1; off >>= 1) win = min(win, __shfl_xor(win, off, 64));
    return mk3(__shfl(bp.x, win, 64), __shfl(bp.y, win, 64), __shfl(bp.z, win, 64));
  }
  // Support of (B + sphere r) - (A + sphere r) along dv, in A's frame.
  DEV f3 sup_md(uint32_t ha, uint32_t hb, q4 rq, f3 t, float r, f3 dv, uint32_t lane) const {
    f3 sb = qrot(rq, sup_lane(hb, qinvrot(rq, dv), lane)) + t;
    f3 sa = sup_lane(ha, -dv, lane);
    return sb - sa + (2.0f * r) * nrm(dv);
  }
  // Minkowski portal refinement (XenoCollide) on the shapes inflated by r: (outward Minkowski normal,
  // inflated depth), w < 0 when the inflated shapes miss.
  DEV q4 mpr(uint32_t ha, uint32_t hb, q4 rq, f3 t, float r, uint32_t lane) const {
    const q4 none = mk4(0, 0, 1, -1);
    f3 v0 = qrot(rq, xyz(w.sf(w.h(W_SF_HAABB) + hb * 2))) + t - xyz(w.sf(w.h(W_SF_HAABB) + ha * 2));
    if (dot(v0, v0) < 1e-14f) v0 = mk3(1e-5f, 0, 0);
    f3 n = -v0;
    f3 v1 = sup_md(ha, hb, rq, t, r, n, lane);
    if (dot(v1, n) <= 0.0f) return none;
    n = cross(v1, v0);
    if (dot(n, n) < 1e-14f) {
      f3 nn = nrm(v1 - v0);
      return mk4(nn, dot(v1, nn));
    }
    f3 v2 = sup_md(ha, hb, rq, t, r, n, lane);
    if (dot(v2, n) <= 0.0f) return none;
    n = cross(v1 - v0, v2 - v0);
    if (dot(n, v0) > 0.0f) { f3 tmp = v1; v1 = v2; v2 = tmp; n = -n; }
    f3 v3 = mk3(0, 0, 0);
    bool found = false;
    for (uint32_t i = 0; i < 32; i++) {
      v3 = sup_md(ha, hb, rq, t, r, n, lane);
      if (dot(v3, n) <= 0.0f) return none;
      if (dot(cross(v1, v3), v0) < 0.0f) { v2 = v3; n = cross(v1 - v0, v3 - v0); continue; }
      if (dot(cross(v3, v2), v0) < 0.0f) { v1 = v3; n = cross(v3 - v0, v2 - v0); continue; }
      found = true;
      break;
    }
    if (!found) return none;
    for (uint32_t i = 0; i < 32; i++) {
      n = cross(v2 - v1, v3 - v1);
      float ln = len(n);
      if (ln < 1e-12f) break;
      n = n / ln;
      if (dot(n, v1 - v0) < 0.0f) n = -n;
      f3 v4 = sup_md(ha, hb, rq, t, r, n, lane);
      float d1 = dot(n, v1);
      if (dot(v4, n) - d1 < 1e-6f || i == 31) {
        if (d1 < 0.0f) return none;
        return mk4(n, d1);
      }
      if (dot(v4, n) < 0.0f) return none;
      f3 c = cross(v4, v0);
      if (dot(v1, c) > 0.0f) { if (dot(v2, c) > 0.0f) v1 = v4; else v3 = v4; }
      else { if (dot(v3, c) > 0.0f) v2 = v4; else v1 = v4; }
    }
    float d1 = dot(n, v1);
    if (d1 < 0.0f) return none;
    return mk4(n, d1);
  }
  // Large hulls: MPR's normal, the reference side from face alignment.
  DEV void mpr_generic(uint32_t sa, uint32_t sb, uint32_t ha, uint32_t hb, f3 pa, q4 qa, q4 rBA, f3 tBA, float thresh, f3 n) const {
    float alA = -2.0f;
    for (uint32_t f = 0; f < w.hnf(ha); f++) alA = mx(alA, dot(xyz(w.fplane(w.hfo(ha) + f)), n));
    float alB = -2.0f;
    for (uint32_t f = 0; f < w.hnf(hb); f++) alB = mx(alB, -dot(qrot(rBA, xyz(w.fplane(w.hfo(hb) + f))), n));
    Geod ga = geo(sa), gb = geo(sb);
    f3 wv[2];
    sup_witness(ga, gb, rBA, tBA, n, alA >= alB, wv);
    gen_run(sa, sb, ga, gb, pa, qa, rBA, tBA, wv[0], wv[1], n, thresh);
  }
  DEV void sat_generic(uint32_t sa, uint32_t sb, uint32_t ha, uint32_t hb, f3 pa, q4 qa, q4 rBA, f3 tBA, float thresh, FaceQd fa, FaceQd fb,
                        const EdgeQd& eq) const {
    const float SAT_TIE = 1e-5f;
    Geod ga = geo(sa), gb = geo(sb);
    if (eq.sep > mx(fa.sep, fb.sep) + SAT_TIE) {
      f3 c[2];
      closest_seg(eq.a0, eq.a1, eq.b0, eq.b1, c);
      gen_run(sa, sb, ga, gb, pa, qa, rBA, tBA, c[0], c[1], eq.axis, thresh);
      return;
    }
    bool refA = !(fb.sep > fa.sep + SAT_TIE);
    f3 n = refA ? xyz(w.fplane(w.hfo(ha) + fa.face)) : -qrot(rBA, xyz(w.fplane(w.hfo(hb) + fb.face)));
    f3 wv[2];
    sup_witness(ga, gb, rBA, tBA, n, refA, wv);
    gen_run(sa, sb, ga, gb, pa, qa, rBA, tBA, wv[0], wv[1], n, thresh);
  }
  DEV void plane_convex(uint32_t sa, uint32_t sb, f3 pa, q4 qa, f3 pb, q4 qb, float thresh) const {
    Geod gb = geo(sb);
    f3 nw = qrot(qa, mk3(0, 0, 1));
    float L = 10.0f * (len(xyz(w.sf(w.h(W_SF_HAABB) + gb.h * 2))) + len(xyz(w.sf(w.h(W_SF_HAABB) + gb.h * 2 + 1))));
    f3 c = pb - nw * (dot(pb - pa, nw) + L);
    q4 rq = qmul(qconj(qa), qb);
    f3 rt = qinvrot(qa, pb - c);
    f3 n = mk3(0, 0, 1);
    f3 wb = qrot(rq, gsup(gb, qinvrot(rq, -n))) + rt;
    float sep = wb.z - L;
    if (sep > thresh) return;
    Geod box; box.t = GT_BOX; box.s = mk3(L, L, L); box.h = 0;
    gen_run(sa, sb, box, gb, c, qa, rq, rt, wb - n * sep, wb, n, thresh);
  }
  DEV void capsule_generic(uint32_t c, uint32_t h, f3 pc, q4 qc, f3 wp, q4 wq, f3 wc, f3 wh, f3 nh, float thresh) const {
    if (dot(wh - wc, nh) - shape_rec(c, 6).y > thresh) return;
    q4 rq = qmul(qconj(qc), wq);
    f3 rt = qinvrot(qc, wp - pc);
    gen_run(c, h, geo(c), geo(h), pc, qc, rq, rt, qrot(rq, wc) + rt, qrot(rq, wh) + rt, qrot(rq, nh), thresh);
  }
  // One candidate pair on one wave (wgsl.js:1072-1195): all lanes share the uniform control flow.
  DEV void narrow_pair(uint32_t pi, uint32_t lane) const {
    uint32_t x0 = w.d.pairs[pi * 2], y0 = w.d.pairs[pi * 2 + 1];
    bool swp = w.geo_type(x0) > w.geo_type(y0);
    uint32_t sa = swp ? y0 : x0, sb = swp ? x0 : y0;
    uint32_t PS = w.h(W_P_SHAPE);
    uint32_t ha = tu(shape_rec(sa, 0).y), hb = tu(shape_rec(sb, 0).y);
    f3 pa = xyz(w.d.poses[PS + sa * 2]);
    q4 qa = Q(w.d.poses[PS + sa * 2 + 1]);
    f3 pb = xyz(w.d.poses[PS + sb * 2]);
    q4 qb = Q(w.d.poses[PS + sb * 2 + 1]);
    float thresh = mn(shape_rec(sa, 0).w + shape_rec(sb, 0).w, w.hf(W_F_GAP_GATE));
    q4 rBA = qmul(qconj(qa), qb);
    f3 tBA = qinvrot(qa, pb - pa);
    q4 rAB = qconj(rBA);
    f3 tAB = qinvrot(qb, pa - pb);
    q4 ka = w.shape_prim(sa), kb = w.shape_prim(sb);
    if (ka.x == 4.0f || kb.x == 4.0f) {
      if (lane == 0) {
        if (w.geo_type(sb) == GT_CONVEX) plane_convex(sa, sb, pa, qa, pb, qb, thresh);
        else plane_contacts(sa, sb, pa, qa, pb, qb, thresh);
      }
      return;
    }
    if (ka.x != 0.0f && kb.x != 0.0f) { if (lane == 0) prim_prim(sa, sb, pa, qa, pb, qb, thresh); return; }
    if (ka.x != 0.0f || kb.x != 0.0f) {
      bool flip = ka.x != 0.0f;
      uint32_t hs = flip ? sb : sa, ps = flip ? sa : sb, h = flip ? hb : ha;
      q4 pr = flip ? ka : kb;
      q4 rq = flip ? rAB : rBA;
      f3 rt = flip ? tAB : tBA;
      f3 u = qrot(rq, mk3(0, 0, pr.z));
      f3 e0 = rt - u, e1 = rt + u;
      FaceQd fq = face_query_seg(h, e0, e1, pr.y, lane);
      if (fq.sep > thresh) return;
      f3 wp = flip ? pb : pa;
      q4 wq = flip ? qb : qa;
      if (fq.sep + pr.y > 1e-6f || (pr.x == 2.0f && !seg_hits_hull(h, e0, e1))) {
        SegQd q = seg_hull_query(h, e0, e1, lane);
        if (lane != 0) return;
        if (pr.x == 2.0f) {
          float dd = sqrtf(q.d2);
          if (dd >= 1e-9f) {
            f3 nh = (q.ph - q.ps) / dd;
            capsule_generic(ps, hs, flip ? pa : pb, flip ? qa : qb, wp, wq, q.ps + nh * 1e-4f, q.ph, nh, thresh);
          }
        } else {
          float dd = sqrtf(q.d2);
          float phi = dd - pr.y;
          f3 n = (q.ps - q.ph) / mx(dd, 1e-9f);
          if (phi > thresh || dd < 1e-9f) {
          } else if (dot(n, xyz(w.fplane(w.hfo(h) + fq.face))) > 0.999f) {
            seg_on_face(hs, ps, h, fq.face, e0, e1, pr.y, wp, wq, thresh, flip);
          } else {
            uint32_t sa2 = flip ? ps : hs, sb2 = flip ? hs : ps;
            f3 nw = qrot(wq, n);
            uint32_t key = hash_u32(hash_u32(sa2 * 7919u + sb2 * 104729u) ^ (q.face * 131u + q.kind * 17u + 0x9e37u));
            emit(sa2, sb2, wp + qrot(wq, q.ph + n * (0.5f * phi)), flip ? -nw : nw, phi, key);
          }
        }
        return;
      }
      if (lane != 0) return;
      if (pr.x == 2.0f) {
        q4 pl = w.fplane(w.hfo(h) + fq.face);
        f3 x = dot(xyz(pl), e0) <= dot(xyz(pl), e1) ? e0 : e1;
        capsule_generic(ps, hs, flip ? pa : pb, flip ? qa : qb, wp, wq, x - xyz(pl) * 1e-4f, x - xyz(pl) * (dot(xyz(pl), x) - pl.w), -xyz(pl), thresh);
      } else {
        seg_on_face(hs, ps, h, fq.face, e0, e1, pr.y, wp, wq, thresh, flip);
      }
      return;
    }
    FaceQd fa = face_query(ha, hb, rBA, tBA, lane);
    if (fa.sep > thresh) return;
    FaceQd fb = face_query(hb, ha, rAB, tAB, lane);
    if (fb.sep > thresh) return;
    if (!(w.hnv(ha) <= 16 && w.hnv(hb) <= 16)) {
      // Face-only SAT misjudges edge contacts on large hulls: MPR's normal instead, unless a face axis
      // separates the hulls better (wgsl.js:1164-1176).
      q4 m = mpr(ha, hb, rBA, tBA, 0.5f * thresh, lane);
      if (lane == 0 && m.w >= 0.0f) {
        f3 n = -xyz(m);
        float sep = dot(n, tBA) + support_min(hb, qinvrot(rBA, n)) - support_max(ha, n);
        if (mx(fa.sep, fb.sep) > sep + 1e-4f) {
          EdgeQd none;
          none.sep = -1e30f; none.axis = none.a0 = none.a1 = none.b0 = none.b1 = mk3(0, 0, 0); none.id = 0;
          sat_generic(sa, sb, ha, hb, pa, qa, rBA, tBA, thresh, fa, fb, none);
        } else {
          mpr_generic(sa, sb, ha, hb, pa, qa, rBA, tBA, thresh, n);
        }
      }
      return;
    }
    EdgeQd eq = edge_query(ha, hb, rBA, tBA, lane);
    if (eq.sep > thresh) return;
    if (lane == 0) sat_generic(sa, sb, ha, hb, pa, qa, rBA, tBA, thresh, fa, fb, eq);
  }
};
