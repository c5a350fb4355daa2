// Per-world accessors shared by the step kernels: record readers for the packed world arrays
// (packages/steel-webgpu/src/wgsl.js:70-105, 185-195; artwgsl.js:29-95) and the articulated-body
// algorithm (artwgsl.js:98-596), with the reference's operation order.
#pragma once
#include "vmath.h"
#include "steel_dev.h"
#include "../../include/steel_world.h"

#define DEV __device__ __forceinline__
#define DEVN __device__ __noinline__

static constexpr uint32_t NLANE = 32;   // the reference's lane stride; tie order of every SAT search
static constexpr uint32_t GT_PLANE = 1, GT_SPHERE = 3, GT_CAPSULE = 4, GT_CYL = 6, GT_BOX = 7, GT_CONE = 9, GT_CONVEX = 10;
static constexpr uint32_t JT_PRISMATIC = 0, JT_REVOLUTE = 1, JT_BALL = 2, JT_FREE = 4, JT_D6 = 6;
static constexpr uint32_t LB = 268, L_S = 0, L_U = 36, L_DI = 72, L_IA = 108, L_C = 135, L_PA = 141, L_V = 147, L_W = 153, L_A = 189,
                          L_UU = 195, L_P = 201, L_DV = 207, L_POS = 213, L_ROT = 216, L_COM = 220, L_TAU = 223, L_JR = 229, L_W2 = 232;
static constexpr uint32_t AVN = 14;
static constexpr uint32_t MAXPATH_DEV = 64;  // deepest articulation root path the kernels take
static constexpr uint32_t MAXJM = 6;
static constexpr uint32_t MAX_PATHDOF = 192;  // MAXPATH * JM bound of the link-response scratch

struct SVd { f3 a, l; };
// Orders LDS and memory between the lanes of one wave (tree passes run one wave per articulation).
DEV void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}
DEV SVd sv0() { SVd s; s.a = mk3(0, 0, 0); s.l = mk3(0, 0, 0); return s; }
DEV SVd mksv(f3 a, f3 l) { SVd s; s.a = a; s.l = l; return s; }
DEV SVd sadd(SVd x, SVd y) { return mksv(x.a + y.a, x.l + y.l); }
DEV SVd sscale(SVd x, float s) { return mksv(x.a * s, x.l * s); }
DEV float sdot(SVd m, SVd f) { return dot(m.a, f.a) + dot(m.l, f.l); }
DEV SVd crm(SVd v, SVd m) { return mksv(cross(v.a, m.a), cross(v.a, m.l) + cross(v.l, m.a)); }
DEV SVd crf(SVd v, SVd f) { return mksv(cross(v.a, f.a) + cross(v.l, f.l), cross(v.a, f.l)); }
DEV SVd mshift(SVd m, f3 r) { return mksv(m.a, m.l + cross(m.a, r)); }
DEV SVd fshift(SVd f, f3 r) { return mksv(f.a + cross(r, f.l), f.l); }

// Body-pair key of a contact or unit: (b + 1, a + 1) as one 64-bit word, so keys order by b, then a, as the
// reference's 16-bit packing (a + 1) | (b + 1) << 16 does (wgsl.js:1517-1524, 3906-3912), for any body count.
DEV uint64_t pair_key(int a, int b) { return ((uint64_t)(uint32_t)(b + 1) << 32) | (uint32_t)(a + 1); }
DEV int key_a(uint64_t k) { return (int)(uint32_t)k - 1; }
DEV int key_b(uint64_t k) { return (int)(uint32_t)(k >> 32) - 1; }

// A world's view: its header (LDS) and buffers.  SG: header words, uniform across a wave, are moved into scalar
// registers as they are read (a gain in the prepare and solve kernels; the collision kernels run faster leaving
// that to the compiler).
#ifndef AS_GLOBAL
#define AS_GLOBAL __attribute__((address_space(1)))
#endif
#ifndef AS_LDS
#define AS_LDS __attribute__((address_space(3)))
#endif
typedef float gv4f __attribute__((ext_vector_type(4)));

// Prepare's LDS mirror of the factored link data an articulation's solves read (null: read the link blocks).
// Per link: (parent, ndof, first dof, mirror offset), and at the offset S (6n), U (6n), D^-1 (n x n), link_r (3).
typedef uint32_t pm_u4 __attribute__((ext_vector_type(4)));
struct PrepMirror {
  const __attribute__((address_space(3))) float* f;
  const __attribute__((address_space(3))) pm_u4* meta;
};
DEV uint32_t prep_mirror_floats(uint32_t n) { return 12 * n + n * n + 3; }

template <bool SG>
struct WV {
  const uint32_t* H;  // header copy in LDS
  WorldDev d;
  SchedLayout S;
  PrepMirror pm{nullptr, nullptr};

  DEV uint32_t hu(uint32_t k) const {
    if constexpr (SG) return __builtin_amdgcn_readfirstlane(H[k]);
    else return H[k];
  }
  DEV uint32_t h(uint32_t k) const { return hu(k); }
  DEV float hf(uint32_t k) const { return ubits(hu(k)); }
  DEV uint32_t NB() const { return hu(W_NB); }
  DEV uint32_t NS() const { return hu(W_NS); }
  DEV uint32_t NA() const { return hu(W_NA); }
  DEV uint32_t NL() const { return hu(W_NL); }
  DEV uint32_t ND() const { return hu(W_ND); }
  DEV uint32_t NDA() const { return hu(W_NDA); }
  DEV uint32_t JM() const { return hu(W_JM); }
  DEV uint32_t MAXC() const { return hu(W_MAXC); }
  DEV uint32_t MAXP() const { return hu(W_MAXP); }
  DEV uint32_t MAXU() const { return hu(W_MAXU); }
  DEV uint32_t UH() const { return hu(W_UH); }
  DEV uint32_t AH() const { return hu(W_AH); }
  DEV uint32_t NCOL() const { return hu(W_NCOLW) * 32; }
  DEV uint32_t ROWS() const { return hu(W_ROWS); }
  DEV uint32_t PO() const { return hu(W_PO); }
  // World buffers are global memory: typed loads and stores (not flat ones, which wait on LDS traffic too).
  DEV AS_GLOBAL float* ga() const { return (AS_GLOBAL float*)d.art; }
  // (The collision kernels keep generic loads for the static records: they measured faster that way.)
  DEV float4 sf(uint32_t i) const {
    if constexpr (SG) { const gv4f v = ((const AS_GLOBAL gv4f*)d.statF)[i]; return make_float4(v.x, v.y, v.z, v.w); }
    else return d.statF[i];
  }
  DEV uint32_t su(uint32_t i) const {
    if constexpr (SG) return ((const AS_GLOBAL uint32_t*)d.statU)[i];
    else return d.statU[i];
  }
  DEV uint32_t* sch(uint32_t off) const { return d.sched + off; }
  DEV uint64_t* sch64(uint32_t off) const { return (uint64_t*)(d.sched + off); }  // even offsets (HKEY, UKEY, UTPW)

  // ---- body records
  DEV uint32_t bo(int b) const { return hu(W_SF_BODY) + (uint32_t)b * 8; }
  DEV float body_inv_mass(int b) const { return sf(bo(b)).x; }
  DEV uint32_t body_kind(int b) const { return tu(sf(bo(b)).y); }
  DEV int body_art(int b) const { return ti(sf(bo(b)).z); }
  DEV int body_link(int b) const { return ti(sf(bo(b)).w); }
  DEV f3 body_com_local(int b) const { return xyz(sf(bo(b) + 1)); }
  DEV bool same_art(int a, int b) const { return a >= 0 && b >= 0 && body_kind(a) == 2 && body_kind(b) == 2 && body_art(a) == body_art(b); }
  // Sleep nodes: free body b is node b, the bodies of articulation a node NB + a (artwgsl.js:41-49).
  DEV int body_node(int b) const {
    if (b < 0) return -1;
    uint32_t k = body_kind(b);
    if (k == 1) return b;
    if (k == 2 || k == 3) return (int)hu(W_NB) + body_art(b);
    return -1;
  }
  DEV bool node_asleep(int n) const { return hu(W_SLEEP) && n >= 0 && ga()[hu(W_A_SLP) + (uint32_t)n * 4 + 1] > 0.5f; }
  DEV bool pair_asleep(int a, int b) const {
    if (!hu(W_SLEEP)) return false;
    int na = body_node(a), nb = body_node(b);
    return (na < 0 || node_asleep(na)) && (nb < 0 || node_asleep(nb));
  }
  DEV float body_mass(int b) const { return sf(bo(b) + 1).w; }
  DEV q4 body_limits(int b) const { uint32_t o = bo(b); return mk4(sf(o + 3).w, sf(o + 4).w, sf(o + 5).w, sf(o + 6).w); }
  DEV float depen_cap(int a, int b) const {
    float m = 1e20f;
    if (a >= 0) { float l = body_limits(a).w; if (l > 0.0f) m = mn(m, l); }
    if (b >= 0) { float l = body_limits(b).w; if (l > 0.0f) m = mn(m, l); }
    return m;
  }
  DEV uint32_t body_flags(int b) const { if (b < 0) return 0; return tu(sf(bo(b) + 7).w); }
  DEV float body_radius(int b) const { if (b < 0) return 1e30f; return sf(bo(b) + 2).w; }
  DEV m3 rows_T(uint32_t o) const { m3 m; m.c[0] = xyz(sf(o)); m.c[1] = xyz(sf(o + 1)); m.c[2] = xyz(sf(o + 2)); return tr(m); }
  DEV m3 body_inertia(int b) const { return rows_T(bo(b) + 2); }
  DEV m3 body_inv_inertia(int b) const { return rows_T(bo(b) + 5); }
  // ---- hull records
  DEV f3 hv(uint32_t h_, uint32_t i) const { return xyz(sf(hu(W_SF_VERTS) + su(hu(W_SU_HULL) + h_ * 4) + i)); }
  DEV uint32_t hnv(uint32_t h_) const { return su(hu(W_SU_HULL) + h_ * 4 + 1); }
  DEV uint32_t hfo(uint32_t h_) const { return su(hu(W_SU_HULL) + h_ * 4 + 2); }
  DEV uint32_t hnf(uint32_t h_) const { return su(hu(W_SU_HULL) + h_ * 4 + 3); }
  DEV q4 fplane(uint32_t f) const { return Q(sf(hu(W_SF_FPLANE) + f)); }
  DEV uint32_t fstart(uint32_t f) const { return su(hu(W_SU_FACE) + f * 2); }
  DEV uint32_t fcount(uint32_t f) const { return su(hu(W_SU_FACE) + f * 2 + 1); }
  DEV uint32_t fvert(uint32_t k) const { return su(hu(W_SU_FV) + k); }
  DEV uint32_t hne(uint32_t h_) const { return su(hu(W_SU_HEDGE) + h_ * 2 + 1); }
  DEV uint4 hedge(uint32_t h_, uint32_t k) const {
    uint32_t o = hu(W_SU_EDGE) + (su(hu(W_SU_HEDGE) + h_ * 2) + k) * 4;
    return make_uint4(su(o), su(o + 1), su(o + 2), su(o + 3));
  }
  DEV uint32_t so(uint32_t s) const { return hu(W_SF_SHAPE) + s * 7; }
  DEV q4 shape_prim(uint32_t s) const { return Q(sf(so(s) + 3)); }
  DEV uint32_t geo_type(uint32_t s) const { return tu(sf(so(s) + 6).x); }
  // ---- articulation records
  DEV uint32_t lk(uint32_t l) const { return hu(W_SF_LINK) + l * 6; }
  DEV uint32_t link_type(uint32_t l) const { return tu(sf(lk(l)).x); }
  DEV int link_parent(uint32_t l) const { return pm.meta ? (int)pm.meta[l].x : ti(sf(lk(l)).y); }
  DEV uint32_t link_dof(uint32_t l) const { return pm.meta ? pm.meta[l].z : tu(sf(lk(l)).z); }
  DEV int link_body(uint32_t l) const { return ti(sf(lk(l)).w); }
  DEV uint32_t link_ndof(uint32_t l) const { return pm.meta ? pm.meta[l].y : tu(sf(lk(l) + 1).w); }
  DEV uint32_t link_qs(uint32_t l) const { return tu(sf(lk(l) + 3).w); }
  DEV float4 dof_drive(uint32_t dd) const { return sf(hu(W_SF_DOF) + dd * 5); }
  DEV bool driven_dof(uint32_t dd) const { const float4 dr = dof_drive(dd); return dr.y > 0.0f || dr.z > 0.0f; }
  // drive_mode rows: a driven dof's velocity limit is left to the sweep clamp (artwgsl.js:54)
  DEV bool fused_dof(uint32_t dd) const { return hu(W_DRIVE_ROWS) && hu(W_FUSE_VLIM) && driven_dof(dd); }
  DEV f3 dof_axis(uint32_t dd) const { return xyz(sf(hu(W_SF_DOF) + dd * 5 + 3)); }
  DEV uint32_t link_nlin(uint32_t l) const { return tu(sf(lk(l) + 5).w); }  // D6: leading translation dofs
  DEV float4 dof_limit(uint32_t dd) const { return sf(hu(W_SF_DOF) + dd * 5 + 1); }
  DEV float4 dof_passive(uint32_t dd) const { return sf(hu(W_SF_DOF) + dd * 5 + 2); }
  DEV bool scalar_dof(uint32_t dd) const { return sf(hu(W_SF_DOF) + dd * 5 + 2).w > 0.5f; }
  DEV uint4 art_info(uint32_t a) const { uint32_t o = hu(W_SU_ART) + a * 4; return make_uint4(su(o), su(o + 1), su(o + 2), su(o + 3)); }
  DEV uint32_t lo(uint32_t l) const { return hu(W_A_LINK) + l * LB; }
  DEV float* art() const { return d.art; }
  DEV f3 ld3(uint32_t o) const { const AS_GLOBAL float* a = ga(); return mk3(a[o], a[o + 1], a[o + 2]); }
  DEV void st3(uint32_t o, f3 v) const { AS_GLOBAL float* a = ga(); a[o] = v.x; a[o + 1] = v.y; a[o + 2] = v.z; }
  DEV SVd ldsv(uint32_t o) const { return mksv(ld3(o), ld3(o + 3)); }
  DEV void stsv(uint32_t o, SVd s) const { st3(o, s.a); st3(o + 3, s.l); }
  DEV m3 ldm(uint32_t o) const { m3 m; m.c[0] = ld3(o); m.c[1] = ld3(o + 3); m.c[2] = ld3(o + 6); return m; }
  DEV void stm(uint32_t o, const m3& m) const { st3(o, m.c[0]); st3(o + 3, m.c[1]); st3(o + 6, m.c[2]); }
  DEV SVd imul(uint32_t o, SVd m) const {
    m3 A = ldm(o), B = ldm(o + 9), C = ldm(o + 18);
    return mksv(mul(A, m.a) + mul(B, m.l), mul(tr(B), m.a) + mul(C, m.l));
  }
  DEV SVd pmsv(uint32_t o) const { const auto* p = pm.f + o; return mksv(mk3(p[0], p[1], p[2]), mk3(p[3], p[4], p[5])); }
  DEV SVd ldS(uint32_t l, uint32_t k) const {
    if (pm.meta) return pmsv(pm.meta[l].w + k * 6);
    return ldsv(lo(l) + L_S + k * 6);
  }
  DEV SVd ldU(uint32_t l, uint32_t k) const {
    if (pm.meta) { const pm_u4 m = pm.meta[l]; return pmsv(m.w + 6 * m.y + k * 6); }
    return ldsv(lo(l) + L_U + k * 6);
  }
  DEV float ldDI(uint32_t l, uint32_t r, uint32_t c) const {
    if (pm.meta) { const pm_u4 m = pm.meta[l]; return pm.f[m.w + 12 * m.y + r * m.y + c]; }
    return ga()[lo(l) + L_DI + r * JM() + c];
  }
  // COM(l) - COM(parent), zero at a root.
  DEV f3 link_r_blocks(uint32_t l) const {
    int par = ti(sf(lk(l)).y);
    if (par < 0) return mk3(0, 0, 0);
    return ld3(lo(l) + L_COM) - ld3(lo((uint32_t)par) + L_COM);
  }
  DEV f3 link_r(uint32_t l) const {
    if (pm.meta) { const pm_u4 m = pm.meta[l]; const auto* p = pm.f + m.w + 12 * m.y + m.y * m.y; return mk3(p[0], p[1], p[2]); }
    return link_r_blocks(l);
  }
  DEV q4 link_rot(uint32_t l) const { f3 r = ld3(lo(l) + L_ROT); return mk4(r, ga()[lo(l) + L_ROT + 3]); }
  DEV float4 pose4(uint32_t i) const {
    if constexpr (SG) { const gv4f v = ((const AS_GLOBAL gv4f*)d.poses)[i]; return make_float4(v.x, v.y, v.z, v.w); }
    else return d.poses[i];
  }
  DEV q4 pose_q(int b) const { return Q(pose4(hu(W_P_BODY) + (uint32_t)b * 2 + 1)); }
  DEV f3 pose_c(int b) const { return xyz(pose4(hu(W_P_BODY) + (uint32_t)b * 2)); }

  // ---- Featherstone (artwgsl.js), one thread per articulation
  // Forward kinematics of one link from its parent's pose (artwgsl.js:98-187).
  DEV void fk_link(uint32_t l) const {
    uint32_t JMv = JM();
    {
      uint32_t o = lk(l);
      f3 pp = mk3(0, 0, 0);
      q4 pq = mk4(0, 0, 0, 1);
      int par = link_parent(l);
      if (par >= 0) { pp = ld3(lo((uint32_t)par) + L_POS); pq = link_rot((uint32_t)par); }
      f3 jp = pp + qrot(pq, xyz(sf(o + 1)));
      q4 jq = qmul(pq, Q(sf(o + 2)));
      f3 axis = xyz(sf(o + 5));
      q4 mq = mk4(0, 0, 0, 1);
      f3 mp = mk3(0, 0, 0);
      uint32_t t = link_type(l), qs = link_qs(l);
      uint32_t AQ = hu(W_A_Q);
      AS_GLOBAL float* art_ = ga();
      if (t == JT_REVOLUTE) mq = qaxis(axis, art_[AQ + qs]);
      if (t == JT_PRISMATIC) mp = axis * art_[AQ + qs];
      if (t == JT_BALL) mq = nrm4(mk4(art_[AQ + qs], art_[AQ + qs + 1], art_[AQ + qs + 2], art_[AQ + qs + 3]));
      // D6: translations along the leading axes, then rotations about axes carried through the rotations before
      // them (artwgsl.js:117-141)
      const uint32_t d0 = link_dof(l), nlin = link_nlin(l), nang = link_ndof(l) - nlin;
      f3 ang[3] = {mk3(0, 0, 0), mk3(0, 0, 0), mk3(0, 0, 0)};
      if (t == JT_D6) {
        for (uint32_t c = 0; c < nlin; c++) mp += dof_axis(d0 + c) * art_[AQ + qs + c];
        const uint32_t iq = qs + nlin;
        if (nang == 1) {
          ang[0] = dof_axis(d0 + nlin);
          mq = qaxis(ang[0], art_[AQ + iq]);
        }
        if (nang == 2) {
          const f3 x0 = nrm(dof_axis(d0 + nlin));
          const f3 x2 = nrm(cross(x0, dof_axis(d0 + nlin + 1)));
          ang[0] = x0;
          const q4 q0 = qaxis(x0, art_[AQ + iq]);
          ang[1] = qrot(q0, cross(x2, x0));
          mq = qmul(qaxis(ang[1], art_[AQ + iq + 1]), q0);
        }
        if (nang == 3) {
          ang[0] = dof_axis(d0 + nlin);
          const q4 q0 = qaxis(ang[0], art_[AQ + iq]);
          ang[1] = qrot(q0, dof_axis(d0 + nlin + 1));
          const q4 q1 = qaxis(ang[1], art_[AQ + iq + 1]);
          ang[2] = qrot(qmul(q1, q0), dof_axis(d0 + nlin + 2));
          mq = qmul(qaxis(ang[2], art_[AQ + iq + 2]), qmul(q1, q0));
        }
      }
      if (t == JT_FREE) {
        mp = mk3(art_[AQ + qs], art_[AQ + qs + 1], art_[AQ + qs + 2]);
        mq = nrm4(mk4(art_[AQ + qs + 3], art_[AQ + qs + 4], art_[AQ + qs + 5], art_[AQ + qs + 6]));
      }
      q4 cq = qmul(jq, mq);
      f3 cp = jp + qrot(jq, mp);
      q4 iq = qconj(Q(sf(o + 4)));
      q4 lq = nrm4(qmul(cq, iq));
      f3 lp = cp + qrot(cq, -qrot(iq, xyz(sf(o + 3))));
      uint32_t L = lo(l);
      st3(L + L_POS, lp); st3(L + L_ROT, xyz(lq)); art_[L + L_ROT + 3] = lq.w;
      f3 com = lp + qrot(lq, body_com_local(link_body(l)));
      st3(L + L_COM, com);
      f3 aw = qrot(jq, axis);
      f3 rj = com - jp;
      for (uint32_t c = 0; c < JMv; c++) stsv(L + L_S + c * 6, sv0());
      if (t == JT_REVOLUTE) stsv(L + L_S, mksv(aw, cross(aw, rj)));
      if (t == JT_PRISMATIC) stsv(L + L_S, mksv(mk3(0, 0, 0), aw));
      if (t == JT_BALL)
        for (uint32_t c = 0; c < 3; c++) {
          f3 e = mk3(0, 0, 0); setc(e, c, 1.0f);
          f3 ew = qrot(jq, e);
          stsv(L + L_S + c * 6, mksv(ew, cross(ew, rj)));
        }
      f3 rf = com - cp;
      st3(L + L_JR, rf);
      if (t == JT_D6) {
        for (uint32_t c = 0; c < nlin; c++) stsv(L + L_S + c * 6, mksv(mk3(0, 0, 0), qrot(jq, dof_axis(d0 + c))));
        for (uint32_t c = 0; c < nang; c++) {
          const f3 ew = qrot(jq, ang[c]);
          stsv(L + L_S + (nlin + c) * 6, mksv(ew, cross(ew, rf)));
        }
      }
      if (t == JT_FREE)
        for (uint32_t c = 0; c < 3; c++) {
          f3 e = mk3(0, 0, 0); setc(e, c, 1.0f);
          f3 ew = qrot(jq, e);
          stsv(L + L_S + c * 6, mksv(mk3(0, 0, 0), ew));
          stsv(L + L_S + (c + 3) * 6, mksv(ew, cross(ew, rf)));
        }
    }
  }
  // FK of articulation a, one wave (lane k: link ai.x + k at depth dl), root level first.
  DEVN void art_fk_lv(uint32_t a, uint32_t dl) const {
    const uint32_t MP = hu(W_MAXPATH), l = art_info(a).x + (threadIdx.x & 63);
    wave_sync();
    for (uint32_t d = 0; d < MP; d++) {
      if (dl == d) fk_link(l);
      wave_sync();
    }
  }
  DEV void inv_spd(float* D, uint32_t n, float* Di) const {
    uint32_t JMv = JM();
    for (uint32_t i = 0; i < JMv * JMv; i++) Di[i] = 0.0f;
    for (uint32_t i = 0; i < n; i++) Di[i * JMv + i] = 1.0f;
    for (uint32_t c = 0; c < n; c++) {
      float pv = 1.0f / mx(D[c * JMv + c], 1e-12f);
      for (uint32_t k = 0; k < n; k++) { D[c * JMv + k] *= pv; Di[c * JMv + k] *= pv; }
      for (uint32_t r = 0; r < n; r++) {
        if (r == c) continue;
        float f = D[r * JMv + c];
        if (f == 0.0f) continue;
        for (uint32_t k = 0; k < n; k++) { D[r * JMv + k] -= f * D[c * JMv + k]; Di[r * JMv + k] -= f * Di[c * JMv + k]; }
      }
    }
  }
  DEV void add_shifted_inertia(uint32_t o, const m3& A, const m3& B, const m3& C, f3 r) const {
    m3 R = skew3(r);
    stm(o, msub(madd(msub(madd(ldm(o), A), mul(B, R)), mul(R, tr(B))), mul(mul(R, C), R)));
    stm(o + 9, madd(madd(ldm(o + 9), B), mul(R, C)));
    stm(o + 18, madd(ldm(o + 18), C));
  }
  // Articulated-body factor (artwgsl.js:210-369), split into per-link steps run level by level by
  // art_factor_lv.  Pass 1 (root first): velocity, bias, rigid inertia and bias force of the link.
  DEV void factor_fwd_link(uint32_t l) const {
    AS_GLOBAL float* art_ = ga();
    uint32_t AQD = hu(W_A_QD), ABF = hu(W_A_BF);
    f3 g = mk3(hf(W_GX), hf(W_GY), hf(W_GZ));
    uint32_t L = lo(l);
    int par = link_parent(l);
    SVd vp = sv0();
    if (par >= 0) vp = mshift(ldsv(lo((uint32_t)par) + L_V), link_r(l));
    SVd sq = sv0();
    for (uint32_t c = 0; c < link_ndof(l); c++) sq = sadd(sq, sscale(ldS(l, c), art_[AQD + link_dof(l) + c]));
    SVd v = sadd(vp, sq);
    stsv(L + L_V, v);
    SVd cb = crm(vp, sq);
    const uint32_t jt = link_type(l);
    if (jt == JT_D6) {
      // rotation axes transported by earlier D6 angles: d(a_k)/dt = sum_{j<k} a_j x a_k qd_j (artwgsl.js:223-233)
      const uint32_t nlin = link_nlin(l), n = link_ndof(l), dd = link_dof(l);
      f3 ca = mk3(0, 0, 0);
      for (uint32_t j = nlin; j < n; j++)
        for (uint32_t k = j + 1; k < n; k++) ca += cross(ldS(l, j).a, ldS(l, k).a) * art_[AQD + dd + j] * art_[AQD + dd + k];
      cb = sadd(cb, mksv(ca, cross(ca, ld3(L + L_JR))));
    }
    if (jt == JT_FREE || jt == JT_D6) {
      const uint32_t nlin = jt == JT_FREE ? 3u : link_nlin(l);
      f3 vo = mk3(0, 0, 0), wj = mk3(0, 0, 0);
      for (uint32_t c = 0; c < link_ndof(l); c++) {
        SVd sc = ldS(l, c);
        float q = art_[AQD + link_dof(l) + c];
        if (c < nlin) vo += sc.l * q; else wj += sc.a * q;
      }
      cb = sadd(cb, mksv(mk3(0, 0, 0), -cross(wj, vo)));
    }
    stsv(L + L_C, cb);
    int b = link_body(l);
    float m = body_mass(b);
    m3 R = qmat(link_rot(l));
    m3 Ic = mul(mul(R, body_inertia(b)), tr(R));
    stm(L + L_IA, Ic);
    stm(L + L_IA + 9, m3zero());
    m3 Mm; Mm.c[0] = mk3(m, 0, 0); Mm.c[1] = mk3(0, m, 0); Mm.c[2] = mk3(0, 0, m);
    stm(L + L_IA + 18, Mm);
    SVd pa0 = crf(v, imul(L + L_IA, v));
    if ((tu(sf(bo(b) + 7).w) & 2u) != 0) pa0 = sadd(pa0, mksv(mk3(0, 0, 0), m * g));
    pa0 = sadd(pa0, mksv(-ld3(ABF + (uint32_t)b * 8 + 4), -ld3(ABF + (uint32_t)b * 8)));
    stsv(L + L_PA, pa0);
  }
  // Pass 2 (leaves first), after the link's children were folded in: U, D^-1, u and the reduced inertia;
  // returns the bias force pa the parent folds in (meaningless for a root).
  DEV SVd factor_bwd_link(uint32_t l) const {
    AS_GLOBAL float* art_ = ga();
    uint32_t JMv = JM(), AQD = hu(W_A_QD), AJF = hu(W_A_JF);
    float dt = hf(W_DT);
    uint32_t L = lo(l);
    uint32_t n = link_ndof(l);
    SVd pA = ldsv(L + L_PA);
    float D[36], u[6] = {0, 0, 0, 0, 0, 0};
    for (uint32_t i = 0; i < JMv * JMv; i++) D[i] = 0.0f;
    for (uint32_t c = 0; c < n; c++) stsv(L + L_U + c * 6, imul(L + L_IA, ldS(l, c)));
    for (uint32_t r = 0; r < n; r++) {
      uint32_t dd = link_dof(l) + r;
      for (uint32_t c = 0; c < n; c++) D[r * JMv + c] = sdot(ldS(l, r), ldU(l, c));
      float4 dr = dof_drive(dd);
      D[r * JMv + r] += dr.x;
      float damp = dof_passive(dd).x;
      D[r * JMv + r] += dt * damp;
      float tau = -damp * art_[AQD + dd] + art_[AJF + dd];
      // Passive spring, explicit (artwgsl.js:298-300).
      float4 sp = sf(hu(W_SF_DOF) + dd * 5 + 4);
      if (sp.x != 0.0f) tau += sp.x * (sp.y - art_[hu(W_A_Q) + link_qs(l) + r]);
      // Implicit PD drive toward (A_TGT, A_TGTV), saturated on the implicit torque (artwgsl.js:301-313).
      if (!hu(W_DRIVE_ROWS) && (dr.y > 0.0f || dr.z > 0.0f)) {
        float q = art_[hu(W_A_Q) + link_qs(l) + r], qd = art_[AQD + dd];
        float err = art_[hu(W_A_TGT) + dd] - q;
        float tv = art_[hu(W_A_TGTV) + dd];
        float m = mx(D[r * JMv + r], 1e-9f);
        float qdn = (m * qd + dt * (dr.y * err + dr.z * tv)) / (m + dt * dr.z + dt * dt * dr.y);
        float ti = dr.y * (err - dt * qdn) + dr.z * (tv - qdn);
        if (fabsf(ti) <= dr.w || dr.w <= 0.0f) {
          D[r * JMv + r] += dt * dt * dr.y + dt * dr.z;
          tau += dr.y * err - dt * dr.y * qd + dr.z * (tv - qd);
        } else {
          tau += clampf_(ti, -dr.w, dr.w);
        }
      }
      u[r] = tau - sdot(ldS(l, r), pA);
    }
    float Di[36];
    inv_spd(D, n, Di);
    for (uint32_t i = 0; i < JMv * JMv; i++) art_[L + L_DI + i] = Di[i];
    for (uint32_t r = 0; r < JMv; r++) art_[L + L_UU + r] = u[r];
    if (link_parent(l) < 0) return sv0();
    m3 A = ldm(L + L_IA), B = ldm(L + L_IA + 9), C = ldm(L + L_IA + 18);
    SVd Duu = sv0();
    for (uint32_t r = 0; r < n; r++) {
      SVd Ur = ldU(l, r);
      float du = 0.0f;
      for (uint32_t c = 0; c < n; c++) {
        SVd Uc = ldU(l, c);
        float w = ldDI(l, r, c);
        A = msub(A, mscale(w, outer3(Ur.a, Uc.a)));
        B = msub(B, mscale(w, outer3(Ur.a, Uc.l)));
        C = msub(C, mscale(w, outer3(Ur.l, Uc.l)));
        du += w * u[c];
      }
      Duu = sadd(Duu, sscale(Ur, du));
    }
    stm(L + L_IA, A); stm(L + L_IA + 9, B); stm(L + L_IA + 18, C);
    SVd cbias = ldsv(L + L_C);
    return sadd(sadd(pA, imul(L + L_IA, cbias)), Duu);
  }
  // Fold child c (reduced inertia in its link block, bias force pa) into parent p.
  DEV void factor_fold_child(uint32_t p, uint32_t c, SVd pa) const {
    uint32_t P0 = lo(p), C0 = lo(c);
    f3 r = link_r(c);
    add_shifted_inertia(P0 + L_IA, ldm(C0 + L_IA), ldm(C0 + L_IA + 9), ldm(C0 + L_IA + 18), r);
    stsv(P0 + L_PA, sadd(ldsv(P0 + L_PA), fshift(pa, r)));
  }
  // Pass 3 (root first): accelerations and the free joint velocities qd_hat.
  DEV void factor_acc_link(uint32_t l) const {
    AS_GLOBAL float* art_ = ga();
    uint32_t AQD = hu(W_A_QD), AQDH = hu(W_A_QDH);
    float dt = hf(W_DT);
    f3 g = mk3(hf(W_GX), hf(W_GY), hf(W_GZ));
    uint32_t L = lo(l);
    {
      int par = link_parent(l);
      SVd ap = mksv(mk3(0, 0, 0), -g);
      if (par >= 0) ap = mshift(ldsv(lo((uint32_t)par) + L_A), link_r(l));
      SVd a1 = sadd(ap, ldsv(L + L_C));
      uint32_t n = link_ndof(l);
      SVd acc = a1;
      for (uint32_t r = 0; r < n; r++) {
        float qdd = 0.0f;
        for (uint32_t c = 0; c < n; c++) qdd += ldDI(l, r, c) * (art_[L + L_UU + c] - sdot(ldU(l, c), a1));
        uint32_t dd = link_dof(l) + r;
        art_[AQDH + dd] = art_[AQD + dd] + dt * qdd;
        acc = sadd(acc, sscale(ldS(l, r), qdd));
      }
      stsv(L + L_A, acc);
      if (hu(W_JFL)) {
        // Load-dependent joint friction: the joint's transmitted force and torque under the free dynamics
        // (articulated force at the child) raise its friction bound (artwgsl.js:361-367).
        SVd f = sadd(imul(L + L_IA, acc), ldsv(L + L_PA));
        float fm = len(f.l) + len(f.a);
        for (uint32_t r = 0; r < n; r++) {
          uint32_t dd = link_dof(l) + r;
          art_[hu(W_A_JFL) + dd] = sf(hu(W_SF_DOF) + dd * 5 + 3).w * fm;
        }
      }
    }
  }
  // Factor of articulation a, one wave (lane k: link ai.x + k at depth dl).  tree[k]: bits 0-7 depth, 8-15 the
  // highest-index child (0xff: none), 16-23 the next lower sibling; pabuf: 8 floats per link.  Children are
  // folded into a parent in decreasing link order, the order of the serial pass, so results match it bit
  // for bit.
  DEVN void art_factor_lv(uint32_t a, const AS_LDS uint32_t* tree, AS_LDS float* pabuf) const {
    const uint32_t lane = threadIdx.x & 63, MP = hu(W_MAXPATH);
    const uint4 ai = art_info(a);
    const bool has = lane < ai.y;
    const uint32_t l = ai.x + lane, dl = has ? (tree[lane] & 0xffu) : 0xffu;
    wave_sync();
    for (uint32_t dd = 0; dd < MP; dd++) {
      if (dl == dd) factor_fwd_link(l);
      wave_sync();
    }
    for (int dd = (int)MP - 1; dd >= 0; dd--) {
      if (dl == (uint32_t)dd) {
        SVd pa = factor_bwd_link(l);
        AS_LDS float* pb = pabuf + lane * 8;
        pb[0] = pa.a.x; pb[1] = pa.a.y; pb[2] = pa.a.z; pb[3] = pa.l.x; pb[4] = pa.l.y; pb[5] = pa.l.z;
      }
      wave_sync();
      if (dd > 0 && dl == (uint32_t)(dd - 1)) {
        for (uint32_t c = (tree[lane] >> 8) & 0xffu; c != 0xffu; c = (tree[c] >> 16) & 0xffu) {
          const AS_LDS float* pb = pabuf + c * 8;
          factor_fold_child(l, ai.x + c, mksv(mk3(pb[0], pb[1], pb[2]), mk3(pb[3], pb[4], pb[5])));
        }
      }
      wave_sync();
    }
    for (uint32_t dd = 0; dd < MP; dd++) {
      if (dl == dd) factor_acc_link(l);
      wave_sync();
    }
  }
  // tree of articulation a for art_factor_lv (one wave; lane k fills entry k).
  DEVN void art_tree(uint32_t a, AS_LDS uint32_t* tree) const {
    const uint32_t lane = threadIdx.x & 63;
    const uint4 ai = art_info(a);
    if (lane < ai.y) {
      uint32_t dep = 0;
      for (int p = link_parent(ai.x + lane); p >= 0; p = link_parent((uint32_t)p)) dep++;
      tree[lane] = dep | 0xffff00u;
    }
    wave_sync();
    if (lane == 0)
      for (uint32_t k = 0; k < ai.y; k++) {
        int p = link_parent(ai.x + k);
        if (p < 0) continue;
        uint32_t pk = (uint32_t)p - ai.x;
        tree[k] = (tree[k] & ~0xff0000u) | (((tree[pk] >> 8) & 0xffu) << 16);
        tree[pk] = (tree[pk] & ~0xff00u) | (k << 8);
      }
    wave_sync();
  }
  DEV uint32_t link_path(uint32_t l0, uint32_t* path) const {
    uint32_t n = 0, MP = hu(W_MAXPATH);
    int l = (int)l0;
    while (!(l < 0 || n >= MP)) { path[n++] = (uint32_t)l; l = link_parent((uint32_t)l); }
    return n;
  }
  DEVN float dof_hinv(uint32_t tl, uint32_t td) const {  // artwgsl.js:513-518 via path_test 375-414
    uint32_t path[MAXPATH_DEV];
    uint32_t np = link_path(tl, path);
    uint32_t JMv = JM();
    float ub[MAXPATH_DEV * MAXJM];
    SVd p = sscale(sv0(), -1.0f);
    for (uint32_t i = 0; i < np; i++) {
      uint32_t l = path[i], n = link_ndof(l);
      float u[6] = {0, 0, 0, 0, 0, 0};
      for (uint32_t r = 0; r < n; r++) {
        float t = 0.0f;
        if (l == tl && r == td) t = 1.0f;
        u[r] = t - sdot(ldS(l, r), p);
        ub[i * JMv + r] = u[r];
      }
      for (uint32_t r = 0; r < n; r++) {
        float du = 0.0f;
        for (uint32_t c = 0; c < n; c++) du += ldDI(l, r, c) * u[c];
        p = sadd(p, sscale(ldU(l, r), du));
      }
      p = fshift(p, link_r(l));
    }
    SVd dv = sv0();
    float out = 0.0f;
    for (uint32_t ii = 0; ii < np; ii++) {
      uint32_t i = np - 1 - ii, l = path[i], n = link_ndof(l);
      dv = mshift(dv, link_r(l));
      float dq[6];
      for (uint32_t r = 0; r < n; r++) {
        dq[r] = 0.0f;
        for (uint32_t c = 0; c < n; c++) dq[r] += ldDI(l, r, c) * (ub[i * JMv + c] - sdot(ldU(l, c), dv));
      }
      for (uint32_t r = 0; r < n; r++) {
        dv = sadd(dv, sscale(ldS(l, r), dq[r]));
        if (l == tl && r == td) out = dq[r];
      }
    }
    return out;
  }
  DEVN void link_response(uint32_t l, const uint32_t* jsplit) const {  // artwgsl.js:431-510
    uint32_t path[MAXPATH_DEV];
    uint32_t np = link_path(l, path);
    uint32_t JMv = JM();
    const uint32_t PD = np * JMv;  // compact stride (indexes only path dofs)
    float ub[6 * MAX_PATHDOF];
    SVd p[6];
    for (uint32_t e = 0; e < 6; e++) {
      f3 f = mk3(0, 0, 0), t = mk3(0, 0, 0);
      if (e < 3) setc(f, e, 1.0f); else setc(t, e - 3, 1.0f);
      p[e] = mksv(-t, -f);
    }
    for (uint32_t i = 0; i < np; i++) {
      uint32_t pl = path[i], n = link_ndof(pl);
      f3 r = link_r(pl);
      for (uint32_t e = 0; e < 6; e++) {
        float u[6];
        for (uint32_t c = 0; c < n; c++) { u[c] = 0.0f - sdot(ldS(pl, c), p[e]); ub[e * PD + i * JMv + c] = u[c]; }
        SVd pe = p[e];
        for (uint32_t rr = 0; rr < n; rr++) {
          float du = 0.0f;
          for (uint32_t c = 0; c < n; c++) du += ldDI(pl, rr, c) * u[c];
          pe = sadd(pe, sscale(ldU(pl, rr), du));
        }
        p[e] = fshift(pe, r);
      }
    }
    SVd dv[6];
    for (uint32_t e = 0; e < 6; e++) dv[e] = sv0();
    for (uint32_t ii = 0; ii < np; ii++) {
      uint32_t i = np - 1 - ii, pl = path[i], n = link_ndof(pl);
      f3 r = link_r(pl);
      for (uint32_t e = 0; e < 6; e++) {
        SVd dd = mshift(dv[e], r);
        float dq[6];
        for (uint32_t rr = 0; rr < n; rr++) {
          dq[rr] = 0.0f;
          for (uint32_t c = 0; c < n; c++) dq[rr] += ldDI(pl, rr, c) * (ub[e * PD + i * JMv + c] - sdot(ldU(pl, c), dd));
        }
        for (uint32_t rr = 0; rr < n; rr++) dd = sadd(dd, sscale(ldS(pl, rr), dq[rr]));
        dv[e] = dd;
      }
    }
    float Wm[36];
    for (uint32_t e = 0; e < 6; e++) {
      Wm[0 * 6 + e] = dv[e].l.x; Wm[1 * 6 + e] = dv[e].l.y; Wm[2 * 6 + e] = dv[e].l.z;
      Wm[3 * 6 + e] = dv[e].a.x; Wm[4 * 6 + e] = dv[e].a.y; Wm[5 * 6 + e] = dv[e].a.z;
    }
    AS_GLOBAL float* art_ = ga();
    for (uint32_t k = 0; k < 36; k++) art_[lo(l) + L_W2 + k] = Wm[k];
    for (uint32_t i = 0; i < np; i++) {
      float extra = (float)jsplit[path[i]] - 1.0f;
      if (extra <= 0.0f) continue;
      uint32_t pl = path[i], n = link_ndof(pl);
      float t[36];
      for (uint32_t cc = 0; cc < 6; cc++)
        for (uint32_t x = 0; x < n; x++) {
          float v = 0.0f;
          for (uint32_t y = 0; y < n; y++) v += ldDI(pl, x, y) * ub[cc * PD + i * JMv + y];
          t[cc * 6 + x] = v;
        }
      for (uint32_t rr = 0; rr < 6; rr++)
        for (uint32_t cc = 0; cc < 6; cc++) {
          float sv = 0.0f;
          for (uint32_t x = 0; x < n; x++) sv += ub[rr * PD + i * JMv + x] * t[cc * 6 + x];
          Wm[rr * 6 + cc] += extra * sv;
        }
    }
    for (uint32_t k = 0; k < 36; k++) art_[lo(l) + L_W + k] = Wm[k];
  }
  // link_response with the six unit impulses on six lanes: lane base + e computes direction e (column e of
  // W2 and W).  The split terms need every direction's u along the path; a lane gets them from its link's
  // other five lanes by shuffles.  Each element keeps the serial pass's summation order.  All six lanes of
  // a link must call this together (same l, base).
  DEVN void link_response_e(uint32_t l, uint32_t e, uint32_t base, const uint32_t* jsplit) const {  // artwgsl.js:431-510
    uint32_t path[MAXPATH_DEV];
    uint32_t np = link_path(l, path);
    uint32_t JMv = JM();
    float ub[MAX_PATHDOF];
    f3 f = mk3(0, 0, 0), t = mk3(0, 0, 0);
    if (e < 3) setc(f, e, 1.0f); else setc(t, e - 3, 1.0f);
    SVd p = mksv(-t, -f);
    for (uint32_t i = 0; i < np; i++) {
      uint32_t pl = path[i], n = link_ndof(pl);
      f3 r = link_r(pl);
      float u[6];
      for (uint32_t c = 0; c < n; c++) { u[c] = 0.0f - sdot(ldS(pl, c), p); ub[i * JMv + c] = u[c]; }
      SVd pe = p;
      for (uint32_t rr = 0; rr < n; rr++) {
        float du = 0.0f;
        for (uint32_t c = 0; c < n; c++) du += ldDI(pl, rr, c) * u[c];
        pe = sadd(pe, sscale(ldU(pl, rr), du));
      }
      p = fshift(pe, r);
    }
    SVd dv = sv0();
    for (uint32_t ii = 0; ii < np; ii++) {
      uint32_t i = np - 1 - ii, pl = path[i], n = link_ndof(pl);
      f3 r = link_r(pl);
      SVd dd = mshift(dv, r);
      float dq[6];
      for (uint32_t rr = 0; rr < n; rr++) {
        dq[rr] = 0.0f;
        for (uint32_t c = 0; c < n; c++) dq[rr] += ldDI(pl, rr, c) * (ub[i * JMv + c] - sdot(ldU(pl, c), dd));
      }
      for (uint32_t rr = 0; rr < n; rr++) dd = sadd(dd, sscale(ldS(pl, rr), dq[rr]));
      dv = dd;
    }
    float col[6] = {dv.l.x, dv.l.y, dv.l.z, dv.a.x, dv.a.y, dv.a.z};  // column e of W2
    AS_GLOBAL float* art_ = ga();
    for (uint32_t k = 0; k < 6; k++) art_[lo(l) + L_W2 + k * 6 + e] = col[k];
    for (uint32_t i = 0; i < np; i++) {
      float extra = (float)jsplit[path[i]] - 1.0f;
      if (extra <= 0.0f) continue;
      uint32_t pl = path[i], n = link_ndof(pl);
      float tc[6];
      for (uint32_t x = 0; x < n; x++) {
        float v = 0.0f;
        for (uint32_t y = 0; y < n; y++) v += ldDI(pl, x, y) * ub[i * JMv + y];
        tc[x] = v;
      }
      for (uint32_t rr = 0; rr < 6; rr++) {
        float sv = 0.0f;
        for (uint32_t x = 0; x < n; x++) sv += __shfl(ub[i * JMv + x], (int)(base + rr), 64) * tc[x];
        col[rr] += extra * sv;
      }
    }
    for (uint32_t k = 0; k < 6; k++) art_[lo(l) + L_W + k * 6 + e] = col[k];
  }
  // Angular Jacobian rows d0, d1 of link x's world angular velocity over its root path (artwgsl.js:542-554).
  DEV void add_gen_ang(uint32_t x, float sgn, f3 d0, f3 d1, uint32_t ds, float* J) const {
    int l = (int)x;
    const uint32_t NDAv = NDA();
    while (l >= 0) {
      const uint32_t k = (uint32_t)l;
      for (uint32_t r = 0; r < link_ndof(k); r++) {
        const f3 wv = ldS(k, r).a;
        const uint32_t i = link_dof(k) + r - ds;
        J[i] += sgn * dot(d0, wv); J[NDAv + i] += sgn * dot(d1, wv);
      }
      l = link_parent(k);
    }
  }
  // Dense rows of the loop closures: one slot per closure, two for a revolute one (wgsl.js:2367-2368).
  DEV uint32_t loop_rows() const {
    uint32_t n = 0;
    for (uint32_t k = 0; k < hu(W_NLOOP); k++) n += sf(hu(W_SF_LOOP) + k * 5).w > 0.5f ? 2u : 1u;
    return n;
  }
  DEV void add_gen_jac(uint32_t x, f3 p, float sgn, f3 d0, f3 d1, f3 d2, uint32_t ds, float* J) const {
    int l = (int)x;
    uint32_t NDAv = NDA();
    while (l >= 0) {
      uint32_t k = (uint32_t)l;
      f3 c = ld3(lo(k) + L_COM);
      for (uint32_t r = 0; r < link_ndof(k); r++) {
        SVd Sv = ldS(k, r);
        f3 v = Sv.l + cross(Sv.a, p - c);
        uint32_t i = link_dof(k) + r - ds;
        J[i] += sgn * dot(d0, v); J[NDAv + i] += sgn * dot(d1, v); J[2 * NDAv + i] += sgn * dot(d2, v);
      }
      l = link_parent(k);
    }
  }
  // qd change of articulation a for generalized impulse tau (artwgsl.js:557-596); scratch: p, dvs in SVd[MAXLP], uu float[6*MAXLP]
  DEVN void art_solve_gen(uint32_t a, const float* tau, float* out, SVd* p, SVd* dvs, float* uu) const {
    uint4 ai = art_info(a);
    for (uint32_t k = 0; k < ai.y; k++) p[k] = sv0();
    for (uint32_t kk = 0; kk < ai.y; kk++) {
      uint32_t k = ai.y - 1 - kk, l = ai.x + k, n = link_ndof(l);
      float u[6] = {0, 0, 0, 0, 0, 0};
      for (uint32_t r = 0; r < n; r++) u[r] = tau[link_dof(l) + r - ai.z] - sdot(ldS(l, r), p[k]);
      for (uint32_t r = 0; r < n; r++) uu[k * 6 + r] = u[r];
      int par = link_parent(l);
      if (par >= 0) {
        SVd pp = p[k];
        for (uint32_t r = 0; r < n; r++) {
          float du = 0.0f;
          for (uint32_t c = 0; c < n; c++) du += ldDI(l, r, c) * u[c];
          pp = sadd(pp, sscale(ldU(l, r), du));
        }
        uint32_t pk = (uint32_t)par - ai.x;
        p[pk] = sadd(p[pk], fshift(pp, link_r(l)));
      }
    }
    for (uint32_t k = 0; k < ai.y; k++) {
      uint32_t l = ai.x + k;
      int par = link_parent(l);
      SVd dv = sv0();
      if (par >= 0) dv = mshift(dvs[(uint32_t)par - ai.x], link_r(l));
      uint32_t n = link_ndof(l);
      float dq[6];
      for (uint32_t r = 0; r < n; r++) {
        dq[r] = 0.0f;
        for (uint32_t c = 0; c < n; c++) dq[r] += ldDI(l, r, c) * (uu[k * 6 + c] - sdot(ldU(l, c), dv));
      }
      for (uint32_t r = 0; r < n; r++) { dv = sadd(dv, sscale(ldS(l, r), dq[r])); out[link_dof(l) + r - ai.z] = dq[r]; }
      dvs[k] = dv;
    }
  }
  DEV void body_response(int b, f3 f, f3 t, f3 out[2]) const {  // wgsl.js:1312-1328
    if (b < 0) { out[0] = mk3(0, 0, 0); out[1] = mk3(0, 0, 0); return; }
    if (body_kind(b) == 1) {
      m3 R = qmat(pose_q(b));
      out[0] = f * body_inv_mass(b);
      out[1] = mul(R, mul(body_inv_inertia(b), mul(tr(R), t)));
      return;
    }
    uint32_t wo = lo((uint32_t)body_link(b)) + L_W;
    float x[6] = {f.x, f.y, f.z, t.x, t.y, t.z}, o[6];
    for (uint32_t r = 0; r < 6; r++) { float s = 0.0f; for (uint32_t c = 0; c < 6; c++) s += ga()[wo + r * 6 + c] * x[c]; o[r] = s; }
    out[0] = mk3(o[0], o[1], o[2]); out[1] = mk3(o[3], o[4], o[5]);
  }
  DEV void body_response_unsplit(int b, f3 f, f3 t, f3 out[2]) const {
    if (b < 0 || body_kind(b) != 2) { body_response(b, f, t, out); return; }
    uint32_t wo = lo((uint32_t)body_link(b)) + L_W2;
    float x[6] = {f.x, f.y, f.z, t.x, t.y, t.z}, o[6];
    for (uint32_t r = 0; r < 6; r++) { float s = 0.0f; for (uint32_t c = 0; c < 6; c++) s += ga()[wo + r * 6 + c] * x[c]; o[r] = s; }
    out[0] = mk3(o[0], o[1], o[2]); out[1] = mk3(o[3], o[4], o[5]);
  }
};

typedef WV<true> W;    // prepare and solve
typedef WV<false> WC;  // broadphase, narrowphase, contact sort

template <bool SG = true>
DEV WV<SG> make_world(const uint32_t* Hs, const WorldDev& wd) {
  WV<SG> w;
  w.H = Hs;
  w.d = wd;
  w.S = sched_layout(w.hu(W_MAXC), w.hu(W_MAXU), w.hu(W_UH), w.hu(W_NCOLW) * 32, w.hu(W_NB), w.hu(W_NCOLW), w.hu(W_NL), w.hu(W_WH), w.hu(W_MAXDENSE));
  return w;
}
