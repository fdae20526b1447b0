// jr_common.h -- per-triangle setup and per-pixel visibility arithmetic shared
// by the forward and backward kernels, so both evaluate the reference's
// formulas (renderer/pipeline.py:76-113, 170-279) with the very same code.
#pragma once

#include "../../include/jrhip.h"
#include "jr_math.h"
#include "jr_shade.h"

namespace jr {

// ---------------------------------------------------------------------------
// triangle records
//
// rec[4t+k] for k < 3 is column k of M^-1 with the clip-space z of vertex k:
// (Minv[0][k], Minv[1][k], Minv[2][k], z_k), so the edge value is
// e_k = (x * r.x + y * r.y) + r.z.  rec[4t+3] = (det, flags, 0, 0).
// ---------------------------------------------------------------------------
constexpr uint32_t kKeep = 1u;     // |det| > 1e-6
constexpr uint32_t kFront = 2u;    // det >= 0
constexpr uint32_t kSane = 4u;     // all coefficients finite and < 1e18

struct Rect {
  int32_t tx0, tx1, ty0, ty1;  // inclusive range, empty when tx0 > tx1
};

// NDC sample position of pixel (i, j) (renderer/pipeline.py:177).
JR_DEV float pix_x(int32_t i, const float* vp) {
  return ((float)i - vp[3]) / vp[0];
}
JR_DEV float pix_y(int32_t j, const float* vp) {
  return ((float)j - vp[7]) / vp[5];
}

// Edge value e_k of the reference: dot((x, y, 1), Minv[:, k]).
JR_DEV float edge(const float4& r, float x, float y) {
  float e = x * r.x;
  e = e + y * r.y;
  e = e + r.z;
  return e;
}

// Window-space depth: interpolate(z_clip, e) * vp[2][2] + vp[2][3]
// (renderer/pipeline.py:212-219).
JR_DEV float win_depth(const float4& r0, const float4& r1, const float4& r2,
                       float e0, float e1, float e2, float vp22, float vp23) {
  float z = e0 * r0.w;
  z = z + e1 * r1.w;
  z = z + e2 * r2.w;
  return z * vp22 + vp23;
}

struct Frame {
  int32_t W, H, TW, TH, tiles_x, tiles_y, T;
  int32_t tw_sh, th_sh;  // log2 TW, log2 TH
  // n / T and n / tiles_y for 0 <= n < 2^31 as (n * m) >> sh (udiv31)
  uint32_t T_m, ty_m;
  int32_t T_sh, ty_sh;
  int32_t n_scenes;
  int64_t capacity;
  // depth-only pass whose output is the depth alone (no primitive ids or
  // barycentrics) plus a nonzero offset: the small-triangle keys are 32-bit
  // depth keys kept in the zbuffer itself
  int32_t narrow;
};

// floor(n / d) for 0 <= n < 2^31 with m = ceil(2^(31 + l) / d), sh = 31 + l,
// l = ceil(log2 d): exact (Granlund and Montgomery, "Division by invariant
// integers using multiplication", 1994, theorem 4.2 with N = 31).
JR_DEV int32_t udiv31(int32_t n, uint32_t m, int32_t sh) {
  return (int32_t)(((uint64_t)(uint32_t)n * m) >> sh);
}

// m, sh of udiv31 for the divisor d (1 <= d < 2^31), on the host.
inline void magic31(uint32_t d, uint32_t* m, int32_t* sh) {
  int l = 0;
  while ((1ull << l) < d) ++l;
  *m = (uint32_t)(((1ull << (31 + l)) + d - 1) / d);
  *sh = 31 + l;
}

// q / h for 0 <= q < 2^22 and 1 <= h < 2^22, without an integer division:
// q and h are exact floats, and the float quotient by the hardware
// reciprocal (within 1 ulp, the product within half an ulp more) is within
// 2e-7 q / h < 1 of the exact one, so its truncation is the quotient or one
// off, and one correction step either way makes it exact.
JR_DEV int32_t div_small(int32_t q, int32_t h) {
  int32_t d = (int32_t)((float)q * __builtin_amdgcn_rcpf((float)h));
  if (d * h > q) --d;
  if ((d + 1) * h <= q) ++d;
  return d;
}

// The scene of triangle t (the last scene whose tri_begin <= t; scene ranges
// ascend) for the wavefront of triangles [t0, t0 + 64), t0 = t & ~63, from a
// table holding the scene of every wavefront's first triangle
// (k_clear fills it): one scalar load, then a wavefront-uniform walk over the
// scenes that begin inside the wavefront's range.  Call from converged code
// of a kernel whose wavefronts hold 64 consecutive triangle ids.
JR_DEV int32_t scene_of_wave(const jr_scene* __restrict__ scenes, int32_t n,
                             const int32_t* __restrict__ wave_scene, int32_t t,
                             int32_t n_tris) {
  const int32_t t0 = __builtin_amdgcn_readfirstlane(t & ~63);
  int32_t su = __builtin_amdgcn_readfirstlane(wave_scene[t0 >> 6]);
  su = min(max(su, 0), n - 1);
  const int32_t tl = min(t0 + 63, n_tris - 1);
  int32_t s = su;
  while (su + 1 < n) {
    const int32_t nb = scenes[su + 1].tri_begin;
    if (nb > tl) break;
    ++su;
    if (t >= nb) s = su;
  }
  return s;
}

// Open-addressing set of non-negative ids in LDS (-1: free), linear probing
// from w mod cap over at most `probes` entries (all of them by default): the
// entry holding w, or -1 when those entries are all taken by other ids.  A
// table's insertions all use one window length, so an id present is always
// found; a short window bounds what a miss costs once the table is full.
JR_DEV int32_t table_insert(int32_t* tab, int32_t cap, int32_t w, int32_t probes = 0) {
  int32_t e = (int32_t)((uint32_t)w % (uint32_t)cap);
  const int32_t n_probe = (probes > 0 && probes < cap) ? probes : cap;
  for (int32_t n = 0; n < n_probe; ++n) {
    const int32_t cur = atomicCAS(tab + e, -1, w);
    if (cur == -1 || cur == w) return e;
    e = (e + 1 == cap) ? 0 : e + 1;
  }
  return -1;
}

// Slot of corner k of face row t: faces[t][k], or 3 t + k when faces is NULL
// (the identity table of a per-corner flattening, not read from memory).
JR_DEV int32_t face_slot(const jr_inputs& in, int32_t t, int k) {
  return in.faces ? in.faces[3 * (int64_t)t + k] : 3 * t + k;
}

// The slots of face row w's corners (the face table's test made once, so its
// three loads issue together).
JR_DEV Corner corner_slots(const jr_inputs& in, int32_t w) {
  Corner c;
  if (in.faces) {
#pragma unroll
    for (int k = 0; k < 3; ++k) c.slot[k] = in.faces[3 * (int64_t)w + k];
  } else {
#pragma unroll
    for (int k = 0; k < 3; ++k) c.slot[k] = 3 * w + k;
  }
  return c;
}

// PerPrimitive.create (renderer/pipeline.py:76-113) for face row t: clip
// coordinates of its three slots, the [x, y, w] matrix, its determinant,
// keep / front-facing flags and the inverse.
struct TriSetup {
  float clip[3][4];
  float M[9];
  float Mi[9];
  float det;
  bool keep, front, sane;
};

// Clip coordinates of face row t's three slots (the vertex stage's positions).
// The uniform choices (face table, clip input, position index) are made once,
// outside the per-corner code, so the three corners' loads of each level are
// issued together: two round trips to memory instead of six.
JR_DEV void tri_corners(const jr_inputs& in, const jr_scene& sc, int32_t t, float clip[3][4]) {
  int32_t slot[3];
  if (in.faces) {
#pragma unroll
    for (int k = 0; k < 3; ++k) slot[k] = in.faces[3 * (int64_t)t + k];
  } else {
#pragma unroll
    for (int k = 0; k < 3; ++k) slot[k] = 3 * t + k;
  }
  if (in.clip) {
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int q = 0; q < 4; ++q) clip[k][q] = in.clip[4 * (int64_t)slot[k] + q];
    return;
  }
  int32_t row[3];
  if (in.pos_index) {
#pragma unroll
    for (int k = 0; k < 3; ++k) row[k] = in.pos_index[slot[k]];
  } else {
#pragma unroll
    for (int k = 0; k < 3; ++k) row[k] = slot[k];
  }
  float p[3][3];
#pragma unroll
  for (int k = 0; k < 3; ++k) load3(in.positions, row[k], p[k]);
  float W[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) W[i] = sc.world_to_clip[i];
#pragma unroll
  for (int k = 0; k < 3; ++k) apply4(W, p[k][0], p[k][1], p[k][2], 1.0f, clip[k]);
}

// The [x, y, w] matrix of the corners and its determinant.
JR_DEV float tri_det(const float clip[3][4], float M[9]) {
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    M[3 * k + 0] = clip[k][0];
    M[3 * k + 1] = clip[k][1];
    M[3 * k + 2] = clip[k][3];
  }
  return det3(M);
}

// Keep / front-facing flags and the inverse, given the corners and det.
JR_DEV void tri_finish(TriSetup& o) {
  o.keep = fabsf(o.det) > 1e-6f;
  o.front = o.det >= 0.0f;
  inv3(o.M, o.Mi);
  bool sane = true;
#pragma unroll
  for (int i = 0; i < 9; ++i) sane = sane && (fabsf(o.Mi[i]) < 1e18f);
  o.sane = sane;
}

JR_DEV void tri_setup(const jr_inputs& in, const jr_scene& sc, int32_t t, TriSetup& o) {
  tri_corners(in, sc, t, o.clip);
  o.det = tri_det(o.clip, o.M);
  tri_finish(o);
}

JR_DEV float4 rec_col(const TriSetup& o, int k) {
  return make_float4(o.Mi[k], o.Mi[3 + k], o.Mi[6 + k], o.clip[k][2]);
}

}  // namespace jr
