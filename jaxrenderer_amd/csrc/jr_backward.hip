// jr_backward.hip -- reverse-mode pass of the forward rasteriser (gfx950).
//
// What jax.grad computes through renderer/pipeline.py with visibility held
// piecewise constant: the chosen primitive, the inside test, keep flags,
// floor/round texture and shadow lookups and all comparisons carry no
// gradient; gradients flow through the barycentrics b = e / sum(e) with
// e = (x, y, 1) M^-1 (so through jnp.linalg.inv: dM = -M^-T G M^-T), the
// window depth, the interpolated varyings, the shading formulas, texture and
// specular-map values (a gather transposes to a scatter-add at the same
// clamped index), the light and the camera matrices.  JAX's conventions are
// kept: max(a, b) splits the gradient 1/2-1/2 on ties, pow(x, y) has
// d/dy = log(x) x^y with log(0) read as log(1), and d/dx = 0 when y == 0.
//
// Nothing is stored from the forward but prim_id: for every written pixel the
// kernels recompute the chosen primitive's setup, edge values, barycentrics
// and fragment attributes with the forward's own code (jr_common.h,
// jr_shade.h), so they are bit-identical to what the forward used.
//
//   k_bwd_prepare / k_bwd_fill  check the workspace header; clear the
//                 per-triangle flags and accumulators only for a workspace that
//                 no earlier call left zero (each call clears what it used)
//   k_bwd_pixels  one wavefront per 16x16 tile (4 pixels per lane): each
//                 distinct primitive's setup and corner stage once, in LDS;
//                 per pixel, shading backward (texture / specular gradients
//                 summed per texel in LDS), the visibility backward (dL/dM^-1,
//                 dL/dz_clip, per-corner varying cotangents) summed per
//                 triangle in LDS, then one set of global atomics per
//                 (tile, triangle) and per (tile, texel)
//   k_bwd_tris    one thread per triangle, the touched ones packed together:
//                 M^-1 -> M -> clip -> positions and the vertex stage of each
//                 corner (normals, colours, light, camera matrices)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>

#include "../../include/jrhip.h"
#include "jr_common.h"
#include "jr_math.h"
#include "jr_shade.h"

namespace jr {

// Per-triangle accumulators: dMinv[3][3] (row r, column k), dz_clip[3],
// dvar[3][3]; phong_darboux adds the cotangents of its face's NDC corners
// (3 x 3) and of their uvs (3 x 2).
constexpr int kAcc = 21;
constexpr int A_MINV = 0;
constexpr int A_Z = 9;
constexpr int A_VAR = 12;
constexpr int A_TRI = 21;
constexpr int A_TUV = 30;
template <int SHADER>
struct AccOf {
  static constexpr int N = kAcc;
};
template <>
struct AccOf<JR_SHADER_PHONG_DARBOUX> {
  static constexpr int N = kAcc + 15;
};

// Accumulator floats per triangle of a shader (host side).
inline int acc_floats(int shader) {
  return shader == JR_SHADER_PHONG_DARBOUX ? AccOf<JR_SHADER_PHONG_DARBOUX>::N : kAcc;
}

// per-scene gradients reduced in k_bwd_pixels
enum {
  S_LIGHT_COLOUR = 0,
  S_LIGHT_DIR_EYE = 3,
  S_AMBIENT = 6,
  S_DIFFUSE = 9,
  S_SPECULAR = 12,
  S_STRENGTH = 15,
  S_LIGHT_DIR = 18,
  S_VP00 = 21,
  S_VP03 = 22,
  S_VP11 = 23,
  S_VP13 = 24,
  S_VP22 = 25,
  S_VP23 = 26,
  S_COUNT = 27
};

__device__ __constant__ int kSceneSlot[S_COUNT] = {
    JR_DSCENE_LIGHT_COLOUR, JR_DSCENE_LIGHT_COLOUR + 1, JR_DSCENE_LIGHT_COLOUR + 2,
    JR_DSCENE_LIGHT_DIR_EYE, JR_DSCENE_LIGHT_DIR_EYE + 1, JR_DSCENE_LIGHT_DIR_EYE + 2,
    JR_DSCENE_AMBIENT, JR_DSCENE_AMBIENT + 1, JR_DSCENE_AMBIENT + 2,
    JR_DSCENE_DIFFUSE, JR_DSCENE_DIFFUSE + 1, JR_DSCENE_DIFFUSE + 2,
    JR_DSCENE_SPECULAR, JR_DSCENE_SPECULAR + 1, JR_DSCENE_SPECULAR + 2,
    JR_DSCENE_SHADOW_STRENGTH, JR_DSCENE_SHADOW_STRENGTH + 1, JR_DSCENE_SHADOW_STRENGTH + 2,
    JR_DSCENE_LIGHT_DIRECTION, JR_DSCENE_LIGHT_DIRECTION + 1, JR_DSCENE_LIGHT_DIRECTION + 2,
    JR_DSCENE_VIEWPORT + 0, JR_DSCENE_VIEWPORT + 3, JR_DSCENE_VIEWPORT + 5,
    JR_DSCENE_VIEWPORT + 7, JR_DSCENE_VIEWPORT + 10, JR_DSCENE_VIEWPORT + 11};

// y = v / |v|: dv = (dy - y (y . dy)) / |v|
JR_DEV void normalise_bwd(const float v[3], const float dy[3], float dv[3]) {
  float ss = v[0] * v[0];
  ss = ss + v[1] * v[1];
  ss = ss + v[2] * v[2];
  const float n = sqrtf(ss);
  const float y[3] = {v[0] / n, v[1] / n, v[2] / n};
  const float yd = dot3(y, dy);
  for (int q = 0; q < 3; ++q) dv[q] = (dy[q] - y[q] * yd) / n;
}

// d max(a, b) / da under JAX's balanced rule
JR_DEV float max_grad(float a, float b) {
  return a > b ? 1.0f : (a == b ? 0.5f : 0.0f);
}

struct PixGrad {
  float db[3];
  float dvar[3][3];
  float sg[S_COUNT];
  // the texel, specular-map and normal-map entry this pixel's gradient lands
  // on (-1: none); scattered after the per-pixel work, with the wavefront
  // converged
  int64_t tex_at, spec_at, nmap_at;
  int32_t tex_n;
  float tex_g[3], spec_g, nmap_g[3];
  // phong_darboux: cotangents of the face's NDC corners and their uvs
  float dtri[9], dtuv[6];
};

// A texel lookup transposes to an add at the same (already clamped) index.
JR_DEV void add_tex(PixGrad& pg, int64_t base, int32_t th, int32_t tc, int32_t u, int32_t v,
                    const float g[3]) {
  pg.tex_at = base + ((int64_t)u * th + v) * tc;
  if (tc > 1) {
    pg.tex_n = 3;
    pg.tex_g[0] = g[0];
    pg.tex_g[1] = g[1];
    pg.tex_g[2] = g[2];
  } else {
    pg.tex_n = 1;
    pg.tex_g[0] = (g[0] + g[1]) + g[2];
  }
}

template <int SHADER>
struct ShadeBwd;

template <>
struct ShadeBwd<JR_SHADER_DEPTH> {
  static JR_DEV void run(const jr_inputs&, const jr_scene&, const jr_grads&, const Corner&,
                         const float*, const float*, PixGrad&, int) {}
};

// The shaders whose corner stage a tile stages once per triangle
// (Varyings<SHADER>): run() computes it, run_v() takes it.
template <int SHADER>
struct CornerBwd {
  static JR_DEV void run(const jr_inputs& in, const jr_scene& sc, const jr_grads& gr,
                         const Corner& c, const float b[3], const float dcol[3], PixGrad& g,
                         int scene) {
    float v[Varyings<SHADER>::N];
    Varyings<SHADER>::corners(in, sc, c, v);
    ShadeBwd<SHADER>::run_v(in, sc, gr, v, b, dcol, g, scene);
  }
};

// shaders/gouraud.py:90-124: col = sum_k b_k col_k
template <>
struct ShadeBwd<JR_SHADER_GOURAUD> : CornerBwd<JR_SHADER_GOURAUD> {
  // v: the three corners' lit colours
  static JR_DEV void run_v(const jr_inputs&, const jr_scene&, const jr_grads&,
                           const float* v, const float b[3], const float dcol[3], PixGrad& g,
                           int) {
    for (int k = 0; k < 3; ++k) {
      g.db[k] = dot3(dcol, v + 3 * k);
      for (int q = 0; q < 3; ++q) g.dvar[k][q] = b[k] * dcol[q];
    }
  }
};

// shaders/gouraud_texture.py:103-144: col = texture[uv] * sum_k b_k lc_k
template <>
struct ShadeBwd<JR_SHADER_GOURAUD_TEXTURE> : CornerBwd<JR_SHADER_GOURAUD_TEXTURE> {
  // v: the corners' light colours lc_k (3 x 3), then their uvs (3 x 2)
  static JR_DEV void run_v(const jr_inputs& in, const jr_scene& sc, const jr_grads&,
                           const float* v, const float b[3], const float dcol[3], PixGrad& g,
                           int) {
    const float* uvv = v + 9;
    float light[3];
    for (int q = 0; q < 3; ++q) light[q] = interp(b, v[q], v[3 + q], v[6 + q]);
    const float uu = interp(b, uvv[0], uvv[2], uvv[4]);
    const float vv = interp(b, uvv[1], uvv[3], uvv[5]);
    const int32_t u = floor_mod(f2i_sat(floorf(uu)), sc.tex_w);
    const int32_t t = floor_mod(f2i_sat(floorf(vv)), sc.tex_h);
    float tex[3];
    tex_rgb(in.textures, sc.tex_base, sc.tex_h, sc.tex_c, u, t, tex);
    float dt[3], dl[3];
    for (int q = 0; q < 3; ++q) {
      dt[q] = dcol[q] * light[q];
      dl[q] = dcol[q] * tex[q];
    }
    add_tex(g, sc.tex_base, sc.tex_h, sc.tex_c, u, t, dt);
    for (int k = 0; k < 3; ++k) {
      g.db[k] = dot3(dl, v + 3 * k);
      for (int q = 0; q < 3; ++q) g.dvar[k][q] = b[k] * dl[q];
    }
  }
};

// shaders/phong.py:106-164
template <>
struct ShadeBwd<JR_SHADER_PHONG> : CornerBwd<JR_SHADER_PHONG> {
  // v: the corners' eye-space normals (3 x 3), then their uvs (3 x 2)
  static JR_DEV void run_v(const jr_inputs& in, const jr_scene& sc, const jr_grads&,
                           const float* v, const float b[3], const float dcol[3], PixGrad& g,
                           int) {
    const float* ne = v;
    const float* uvv = v + 9;
    float n[3];
    for (int q = 0; q < 3; ++q) n[q] = interp(b, ne[q], ne[3 + q], ne[6 + q]);
    const float uu = interp(b, uvv[0], uvv[2], uvv[4]);
    const float vv = interp(b, uvv[1], uvv[3], uvv[5]);
    const int32_t u = floor_mod(f2i_sat(floorf(uu)), sc.tex_w);
    const int32_t t = floor_mod(f2i_sat(floorf(vv)), sc.tex_h);
    float tex[3], nn[3], ln[3];
    tex_rgb(in.textures, sc.tex_base, sc.tex_h, sc.tex_c, u, t, tex);
    normalise3(n, nn);
    normalise3(sc.light_direction, ln);
    const float d = dot3(nn, ln);
    float lc[3];
    bool pos = true;
    for (int q = 0; q < 3; ++q) {
      lc[q] = sc.light_colour[q] * d;
      pos = pos && (lc[q] >= 0.0f);
    }
    if (!pos) return;  // the zero branch of the cond carries no gradient
    float dt[3], dlc[3];
    float dd = 0.0f;
    for (int q = 0; q < 3; ++q) {
      dt[q] = dcol[q] * lc[q];
      dlc[q] = dcol[q] * tex[q];
      g.sg[S_LIGHT_COLOUR + q] += dlc[q] * d;
      dd += dlc[q] * sc.light_colour[q];
    }
    add_tex(g, sc.tex_base, sc.tex_h, sc.tex_c, u, t, dt);
    float dnn[3], dln[3], dn[3];
    for (int q = 0; q < 3; ++q) {
      dnn[q] = dd * ln[q];
      dln[q] = dd * nn[q];
      g.sg[S_LIGHT_DIR + q] += dln[q];  // normalise's backward applied per tile
    }
    normalise_bwd(n, dnn, dn);
    for (int k = 0; k < 3; ++k) {
      g.db[k] = dot3(dn, ne + 3 * k);
      for (int q = 0; q < 3; ++q) g.dvar[k][q] = b[k] * dn[q];
    }
  }
};

// shaders/phong_darboux.py:116-293: the forward of Shade<PHONG_DARBOUX>
// with its intermediates, then its reverse.  Beyond phong's terms the
// gradient reaches the tangent frame: B = [i/|i|, j/|j|, n] with
// (i, j) = inv(A) @ [du, dv] and A = [t1 - t0; t2 - t0; n] built from the
// face's NDC corners t_m = to_cartesian(world_to_clip @ [p_m, 1]) and their
// uvs; d(inv(A)) transposes to dA = -inv(A)^T G inv(A)^T.  The corners'
// cotangents go to the per-triangle accumulators (k_bwd_tris takes them to
// positions, uvs and world_to_clip); the normal-map texel gets d(B @ nm).
template <>
struct ShadeBwd<JR_SHADER_PHONG_DARBOUX> {
  static JR_DEV void run(const jr_inputs& in, const jr_scene& sc, const jr_grads& gr,
                         const Corner& c, const float b[3], const float dcol[3], PixGrad& g,
                         int) {
    float ne[3][3], uvv[3][2];
    for (int k = 0; k < 3; ++k) {
      eye_normal(in, sc, c.slot[k], ne[k]);
      load2(in.uvs, attr_row(in.uv_index, c.slot[k]), uvv[k]);
    }
    const int32_t face = in.id_to_face[c.slot[0]];
    float tri[3][3], tuv[3][2];
    for (int m = 0; m < 3; ++m) {
      const int32_t q = in.face_slots[3 * (int64_t)face + m];
      float h[4];
      position_clip(in, sc.world_to_clip, q, h);
      const bool wz = (h[3] == 0.0f);
      tri[m][0] = wz ? h[0] : h[0] / h[3];
      tri[m][1] = wz ? h[1] : h[1] / h[3];
      tri[m][2] = wz ? h[2] : h[2] / h[3];
      load2(in.uvs, attr_row(in.uv_index, q), tuv[m]);
    }
    float n_i[3];
    for (int q = 0; q < 3; ++q) n_i[q] = interp(b, ne[0][q], ne[1][q], ne[2][q]);
    const float uu = interp(b, uvv[0][0], uvv[1][0], uvv[2][0]);
    const float vv = interp(b, uvv[0][1], uvv[1][1], uvv[2][1]);
    const int32_t u = floor_mod(f2i_sat(floorf(uu)), sc.tex_w);
    const int32_t v = floor_mod(f2i_sat(floorf(vv)), sc.tex_h);
    float n[3];
    normalise3(n_i, n);
    float A[9], AI[9];
    for (int q = 0; q < 3; ++q) {
      A[q] = tri[1][q] - tri[0][q];
      A[3 + q] = tri[2][q] - tri[0][q];
      A[6 + q] = n[q];
    }
    inv3(A, AI);
    const float du1 = tuv[1][0] - tuv[0][0], du2 = tuv[2][0] - tuv[0][0];
    const float dv1 = tuv[1][1] - tuv[0][1], dv2 = tuv[2][1] - tuv[0][1];
    float iv[3], jv[3];
    for (int r = 0; r < 3; ++r) {
      float s = AI[3 * r + 0] * du1;
      s = s + AI[3 * r + 1] * du2;
      s = s + AI[3 * r + 2] * 0.0f;
      iv[r] = s;
      float t = AI[3 * r + 0] * dv1;
      t = t + AI[3 * r + 1] * dv2;
      t = t + AI[3 * r + 2] * 0.0f;
      jv[r] = t;
    }
    float in_[3], jn[3];
    normalise3(iv, in_);
    normalise3(jv, jn);
    const int32_t un = index_clamp(u, sc.nmap_w);
    const int32_t vn = index_clamp(v, sc.nmap_h);
    float nm[3];
    load3(in.normal_maps + sc.nmap_base, un * sc.nmap_h + vn, nm);
    float bn[3];
    for (int r = 0; r < 3; ++r) {
      float s = in_[r] * nm[0];
      s = s + jn[r] * nm[1];
      s = s + n[r] * nm[2];
      bn[r] = s;
    }
    float n2[3], ln[3];
    normalise3(bn, n2);
    normalise3(sc.light_direction, ln);
    const float d = dot3(n2, ln);
    float tex[3];
    tex_rgb(in.textures, sc.tex_base, sc.tex_h, sc.tex_c, u, v, tex);
    float lc[3];
    bool pos = true;
    for (int ch = 0; ch < 3; ++ch) {
      lc[ch] = sc.light_colour[ch] * d;
      pos = pos && (lc[ch] >= 0.0f);
    }
    if (!pos) return;  // the zero branch of the cond carries no gradient

    // colour = tex * light_colour * dot(n2, ln)
    float dt[3], dd = 0.0f;
    for (int q = 0; q < 3; ++q) {
      dt[q] = dcol[q] * lc[q];
      const float dlc = dcol[q] * tex[q];
      g.sg[S_LIGHT_COLOUR + q] += dlc * d;
      dd += dlc * sc.light_colour[q];
    }
    add_tex(g, sc.tex_base, sc.tex_h, sc.tex_c, u, v, dt);
    float dn2[3], dln[3], dbn[3];
    for (int q = 0; q < 3; ++q) {
      dn2[q] = dd * ln[q];
      dln[q] = dd * n2[q];
      g.sg[S_LIGHT_DIR + q] += dln[q];  // normalise's backward applied per tile
    }
    normalise_bwd(bn, dn2, dbn);
    // bn = B @ nm with B's columns in_, jn, n
    float din[3], djn[3], dn[3];
    for (int r = 0; r < 3; ++r) {
      din[r] = dbn[r] * nm[0];
      djn[r] = dbn[r] * nm[1];
      dn[r] = dbn[r] * nm[2];
    }
    if (gr.d_normal_maps) {
      g.nmap_at = sc.nmap_base + 3 * ((int64_t)un * sc.nmap_h + vn);
      g.nmap_g[0] = dot3(dbn, in_);
      g.nmap_g[1] = dot3(dbn, jn);
      g.nmap_g[2] = dot3(dbn, n);
    }
    float div[3], djv[3];
    normalise_bwd(iv, din, div);
    normalise_bwd(jv, djn, djv);
    // (i, j) = AI @ (du, dv): dAI = div du^T + djv dv^T, d(du) = AI^T div
    float dAI[9];
    float ddu1 = 0.0f, ddu2 = 0.0f, ddv1 = 0.0f, ddv2 = 0.0f;
    for (int r = 0; r < 3; ++r) {
      dAI[3 * r + 0] = div[r] * du1 + djv[r] * dv1;
      dAI[3 * r + 1] = div[r] * du2 + djv[r] * dv2;
      dAI[3 * r + 2] = div[r] * 0.0f + djv[r] * 0.0f;
      ddu1 += div[r] * AI[3 * r + 0];
      ddu2 += div[r] * AI[3 * r + 1];
      ddv1 += djv[r] * AI[3 * r + 0];
      ddv2 += djv[r] * AI[3 * r + 1];
    }
    g.dtuv[0] = -(ddu1 + ddu2);  // corner 0: u, v
    g.dtuv[1] = -(ddv1 + ddv2);
    g.dtuv[2] = ddu1;  // corner 1
    g.dtuv[3] = ddv1;
    g.dtuv[4] = ddu2;  // corner 2
    g.dtuv[5] = ddv2;
    // dA = -AI^T dAI AI^T
    float T1[9], dA[9];
    for (int p = 0; p < 3; ++p)
      for (int k = 0; k < 3; ++k) {
        float s = 0.0f;
        for (int r = 0; r < 3; ++r) s += AI[3 * r + p] * dAI[3 * r + k];
        T1[3 * p + k] = s;
      }
    for (int p = 0; p < 3; ++p)
      for (int q = 0; q < 3; ++q) {
        float s = 0.0f;
        for (int k = 0; k < 3; ++k) s += T1[3 * p + k] * AI[3 * q + k];
        dA[3 * p + q] = -s;
      }
    for (int q = 0; q < 3; ++q) {
      g.dtri[0 + q] = -(dA[q] + dA[3 + q]);
      g.dtri[3 + q] = dA[q];
      g.dtri[6 + q] = dA[3 + q];
      dn[q] += dA[6 + q];
    }
    float dni[3];
    normalise_bwd(n_i, dn, dni);
    for (int k = 0; k < 3; ++k) {
      g.db[k] = dot3(dni, ne[k]);
      for (int q = 0; q < 3; ++q) g.dvar[k][q] = b[k] * dni[q];
    }
  }
};

// shaders/phong_reflection.py:154-235 and phong_reflection_shadow.py:173-271
template <bool SHADOW>
struct ReflectionBwd {
  using V = Varyings<SHADOW ? JR_SHADER_PHONG_REFLECTION_SHADOW : JR_SHADER_PHONG_REFLECTION>;
  static JR_DEV void run(const jr_inputs& in, const jr_scene& sc, const jr_grads& gr,
                         const Corner& c, const float b[3], const float dcol[3], PixGrad& g,
                         int scene) {
    float v[V::N];
    V::corners(in, sc, c, v);
    run_v(in, sc, gr, v, b, dcol, g, scene);
  }
  // v: the triangle's corner stage (Varyings<>::corners)
  static JR_DEV void run_v(const jr_inputs& in, const jr_scene& sc, const jr_grads& gr,
                           const float* vary, const float b[3], const float dcol[3],
                           PixGrad& g, int scene) {
    const float* nev = vary;      // eye-space normal of corner k: nev[3k..3k+2]
    const float* uvv = vary + 9;  // uv of corner k: uvv[2k..2k+1]
    const int32_t ti = __float_as_int(vary[V::kTi]);
    float n[3], uv[2];
    for (int q = 0; q < 3; ++q) n[q] = interp(b, nev[q], nev[3 + q], nev[6 + q]);
    uv[0] = interp(b, uvv[0], uvv[2], uvv[4]);
    uv[1] = interp(b, uvv[1], uvv[3], uvv[5]);
    bool lit = true;
    if (SHADOW) {
      const float* scd = vary + 15;  // homogeneous-normalised shadow clip of corner k
      float s4[4], sv[4];
      for (int q = 0; q < 4; ++q) s4[q] = interp(b, scd[q], scd[4 + q], scd[8 + q]);
      apply4(sc.shadow_viewport, s4[0], s4[1], s4[2], s4[3], sv);
      const float* map = in.shadow_maps + (int64_t)scene * in.shadow_w * in.shadow_h;
      lit = (sv[2] / sv[3]) <= shadow_get(map, in.shadow_w, in.shadow_h, sv[0] / sv[3],
                                          sv[1] / sv[3]);
    }
    // forward of reflection_terms with its intermediates
    const int32_t s0 = __float_as_int(vary[V::kShape]), s1 = __float_as_int(vary[V::kShape + 1]);
    float f0 = modf_frac(uv[0]), f1 = modf_frac(uv[1]);
    f0 = (f0 < 0.0f) ? f0 + 1.0f : f0;
    f1 = (f1 < 0.0f) ? f1 + 1.0f : f1;
    const int32_t shift = (int32_t)((uint32_t)ti * (uint32_t)sc.tex_offset);
    const int32_t u0 = f2i_sat(floorf(f0 * (float)s0 + (float)shift));
    const int32_t v0 = f2i_sat(floorf(f1 * (float)s1));
    const int32_t u = index_clamp(u0, sc.tex_w), v = index_clamp(v0, sc.tex_h);
    const int32_t us = index_clamp(u0, sc.spec_w), vs = index_clamp(v0, sc.spec_h);
    float tex[3];
    tex_rgb(in.textures, sc.tex_base, sc.tex_h, sc.tex_c, u, v, tex);
    const int64_t spec_at = sc.spec_base + (int64_t)us * sc.spec_h + vs;
    const float S = in.specular_maps[spec_at];
    float nh[3], lh[3];
    normalise3(n, nh);
    normalise3(sc.light_dir_eye, lh);
    const float d = dot3(nh, lh);
    const float diff = max_nan(d, 0.0f);
    const float two_d = 2.0f * d;
    float qv[3] = {two_d * nh[0] - lh[0], two_d * nh[1] - lh[1], two_d * nh[2] - lh[2]};
    float rh[3];
    normalise3(qv, rh);
    const float m = max_nan(rh[2], 0.0f);
    // the forward's power (pow_unit: within 3e-7 of powf for m in [0, 1])
    const float spec = pow_unit(m, S);

    float dt[3], ddiff = 0.0f, dspec = 0.0f;
    for (int q = 0; q < 3; ++q) {
      const float k = sc.diffuse[q] * diff + sc.specular[q] * spec;
      const float sh = lit ? 1.0f : 1.0f - sc.shadow_strength[q];
      g.sg[S_AMBIENT + q] += dcol[q] * tex[q];
      float dk;
      if (SHADOW) {
        dt[q] = dcol[q] * (sc.ambient[q] + (sh * k) * sc.light_colour[q]);
        dk = dcol[q] * sh * tex[q] * sc.light_colour[q];
        g.sg[S_LIGHT_COLOUR + q] += dcol[q] * (sh * k) * tex[q];
        if (!lit) g.sg[S_STRENGTH + q] += -dcol[q] * k * tex[q] * sc.light_colour[q];
      } else {
        dt[q] = dcol[q] * (sc.ambient[q] + k * sc.light_colour[q]);
        dk = dcol[q] * sc.light_colour[q] * tex[q];
        g.sg[S_LIGHT_COLOUR + q] += dcol[q] * k * tex[q];
      }
      g.sg[S_DIFFUSE + q] += dk * diff;
      g.sg[S_SPECULAR + q] += dk * spec;
      ddiff += dk * sc.diffuse[q];
      dspec += dk * sc.specular[q];
    }
    add_tex(g, sc.tex_base, sc.tex_h, sc.tex_c, u, v, dt);
    if (gr.d_specular_maps) {
      const float lg = logf(m == 0.0f ? 1.0f : m);
      g.spec_at = spec_at;
      g.spec_g = dspec * lg * spec;
    }
    const float dm = dspec * (S == 0.0f ? 0.0f : S * pow_unit(m, S - 1.0f));
    const float dr[3] = {0.0f, 0.0f, dm * max_grad(rh[2], 0.0f)};
    float dq[3];
    normalise_bwd(qv, dr, dq);
    float dnh[3], dlh[3];
    float dtwo_d = dot3(dq, nh);
    for (int q = 0; q < 3; ++q) {
      dnh[q] = two_d * dq[q];
      dlh[q] = -dq[q];
    }
    const float dd = 2.0f * dtwo_d + ddiff * max_grad(d, 0.0f);
    for (int q = 0; q < 3; ++q) {
      dnh[q] += dd * lh[q];
      dlh[q] += dd * nh[q];
    }
    float dn[3];
    normalise_bwd(n, dnh, dn);
    // normalise's backward is linear in its cotangent: applied once per tile
    for (int q = 0; q < 3; ++q) g.sg[S_LIGHT_DIR_EYE + q] += dlh[q];
    for (int k = 0; k < 3; ++k) {
      g.db[k] = dot3(dn, nev + 3 * k);
      for (int q = 0; q < 3; ++q) g.dvar[k][q] = b[k] * dn[q];
    }
  }
};

template <>
struct ShadeBwd<JR_SHADER_PHONG_REFLECTION> : ReflectionBwd<false> {};
template <>
struct ShadeBwd<JR_SHADER_PHONG_REFLECTION_SHADOW> : ReflectionBwd<true> {};

// v moved across lanes by the DPP pattern CTRL (lanes without a source get 0).
template <int CTRL>
JR_DEV float dpp_move(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xf, 0xf, false));
}

// Sum over the 64 lanes of the wavefront, returned to every lane: lane pairs
// and quads (quad_perm [1,0,3,2], [2,3,0,1]), rows of 16 (row_ror 4, 8), then
// rows 0-1 into 1 and 2-3 into 3 (row_bcast 15) and rows 0-1 into 2-3
// (row_bcast 31), so lane 63 holds the total; all register-to-register, no
// LDS.  Call with every lane active.
JR_DEV float wave_sum(float v) {
  v += dpp_move<0xb1>(v);
  v += dpp_move<0x4e>(v);
  v += dpp_move<0x124>(v);
  v += dpp_move<0x128>(v);
  v += dpp_move<0x142>(v);
  v += dpp_move<0x143>(v);
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

// d[at + c] += v[c] (c < n), summed first in an LDS table of (index, sum)
// entries shared by the tile's pixels, so a texel that several pixels read
// costs one global atomic per channel and tile; indices beyond the table's
// capacity (or beyond 2^31) are added directly.
// Sums a lane could not stage (its table full), kept in registers (at most
// KMAX, later ones add at once) and added to global memory after the pixel
// loop, so that no global atomic sits between the loop's loads (a
// wavefront's memory operations complete in order, so a later load would
// wait for the atomic's round trip).  A 16x16 tile whose 256 pixels read 256
// distinct texels fills a 128-entry table after two of a lane's four pixels,
// which leaves two per lane.
template <int NC, int KMAX>
struct Pending {
  int64_t at[KMAX];
  float v[KMAX][NC];
  JR_DEV Pending() {
#pragma unroll
    for (int k = 0; k < KMAX; ++k) at[k] = -1;
  }
  // false when every register entry is taken
  JR_DEV bool push(int64_t a, int32_t n, const float* g) {
#pragma unroll
    for (int k = 0; k < KMAX; ++k) {
      if (at[k] < 0) {
        at[k] = a;
#pragma unroll
        for (int c = 0; c < NC; ++c) v[k][c] = c < n ? g[c] : 0.0f;
        return true;
      }
    }
    return false;
  }
  JR_DEV void flush(float* __restrict__ d) {
#pragma unroll
    for (int k = 0; k < KMAX; ++k)
      if (at[k] >= 0)
#pragma unroll
        for (int c = 0; c < NC; ++c)
          if (v[k][c] != 0.0f) atomicAdd(d + at[k] + c, v[k][c]);
  }
};

constexpr int32_t kStageProbes = 16;

template <int CAP, int NC, int KMAX>
JR_DEV void stage_add(int32_t* tab, float (*val)[NC], float* __restrict__ d, int64_t at,
                      int32_t n, const float* v, Pending<NC, KMAX>& pend) {
  if (!d || at < 0) return;
  // a texel table fills up where nearly every pixel reads its own texel: a
  // miss then costs kStageProbes LDS operations, not CAP
  const int32_t e = at < 0x7FFFFFFFLL ? table_insert(tab, CAP, (int32_t)at, kStageProbes) : -1;
  if (e < 0) {
    if (!pend.push(at, n, v))
      for (int c = 0; c < n; ++c)
        if (v[c] != 0.0f) atomicAdd(d + at + c, v[c]);
    return;
  }
  for (int c = 0; c < n; ++c)
    if (v[c] != 0.0f) atomicAdd(&val[e][c], v[c]);
}

// Adds the staged sums to d (n channels per entry) and leaves the table free.
// Lanes take consecutive (entry, channel) pairs, so the channels of one entry
// are added by neighbouring lanes of one instruction (same cache line).
template <int CAP, int NC>
JR_DEV void flush_adds(int32_t* tab, float (*val)[NC], float* __restrict__ d, int32_t n) {
  if (!d) return;
  for (int idx = threadIdx.x; idx < CAP * n; idx += blockDim.x) {
    const int e = n == 3 ? idx / 3 : (n == 1 ? idx : idx / n);
    const int c = idx - e * n;
    const int32_t at = tab[e];
    if (at >= 0 && val[e][c] != 0.0f) atomicAdd(d + at + c, val[e][c]);
  }
}

struct BwdArgs {
  jr_inputs in;
  const jr_scene* scenes;
  jr_grads gr;
  const float* d_canvas;
  // fused L2 loss: with target set the canvas cotangent of a pixel is
  // l2_scale * (canvas - target), read in place of d_canvas
  const float* canvas;
  const float* target;
  float l2_scale;
  const float* d_zbuffer;
  const int32_t* prim_id;
  float* tacc;              // kAcc floats per triangle, zero between calls
  int32_t* touched;         // per triangle: its scene + 1 once a pixel adds to it, zero between calls
  uint32_t* header;         // workspace header (kHdr* words)
  int32_t W, H, n_scenes;
  // the 2x2 pixel blocks of a scene (k_bwd_pixels), nq of them numbered:
  // FX x FY tiles of 8 x 8 blocks, rw and rh blocks left over right of and
  // below them; blocks [0, full_end) are the tiles', [full_end, bottom_end)
  // those below them; T wavefronts per scene
  int32_t nq, T, FX, FY, rw, rh, full_end, bottom_end;
  uint32_t T_m, fy_m, rh_m, rw_m;  // udiv31 by T, FY, rh and rw (by 1 for a zero)
  int32_t T_sh, fy_sh, rh_sh, rw_sh;
};

// Workspace header words.  The accumulators and flags are left zero by every
// call (k_bwd_tris clears what it consumed), so a call only clears them in
// full when the header does not describe this workspace yet.
constexpr int kHdrMagicLo = 0, kHdrMagicHi = 1, kHdrTris = 2, kHdrFill = 3, kHdrAcc = 4;
constexpr uint32_t kMagicLo = 0x6a724277u, kMagicHi = 0x73706b31u;

// A scene's pixels in 2x2 blocks, one wavefront per 64 consecutive blocks of
// this enumeration: the full tiles of 8 x 8 blocks (16 x 16 pixels) first,
// tile by tile, so that each is one wavefront with lane l owning the block
// (16 tx + 2 (l / 8), 16 ty + 2 (l % 8)); then the blocks below them (rh per
// row) and those right of them (rw per column), which share wavefronts
// instead of leaving most lanes of a partial tile idle (an 84 x 84 frame
// takes 28 wavefronts, not 36).  Where that saves few wavefronts every tile,
// partial ones included, is a wavefront of its own.  The wavefront's
// distinct primitives (up to kBwdCap) get their setup (M^-1 columns and clip
// z) and, for the split shaders, their corner stage computed once, by one
// lane each, into LDS.  A lane sums the cotangents of consecutive pixels of
// one triangle (and the gradients of consecutive pixels on one texel) in
// registers and adds a run to the LDS sums only when it ends, so neighbouring
// pixels, which mostly share both, cost one LDS atomic instead of four; the
// LDS sums reach the global accumulators once per (wavefront, triangle) and
// per (wavefront, texel).  Pixels of primitives beyond the table's capacity compute
// their own and add directly.
constexpr int kBwdCap = 32;
constexpr int kTexCap = 128;  // distinct texels summed in LDS per wavefront
constexpr int kSpecCap = 64;  // distinct specular-map entries

template <int SHADER>
__global__ __launch_bounds__(64) void k_bwd_pixels(BwdArgs a) {
  constexpr int NV = Varyings<SHADER>::N;
  constexpr int NS = (12 + NV) | 1;  // record stride (odd: no LDS bank conflicts)
  __shared__ float trec[kBwdCap * NS];
  constexpr int ACC = AccOf<SHADER>::N;
  constexpr bool DARBOUX = SHADER == JR_SHADER_PHONG_DARBOUX;
  constexpr int kNmapCap = DARBOUX ? kTexCap : 1;  // distinct normal-map texels
  constexpr int AS = ACC | 1;  // odd row stride: lanes of different rows hit different banks
  __shared__ float acc[kBwdCap][AS];
  __shared__ int32_t tab[kBwdCap];
  __shared__ int32_t tex_tab[kTexCap];
  __shared__ float tex_val[kTexCap][3];
  __shared__ int32_t spec_tab[kSpecCap];
  __shared__ float spec_val[kSpecCap][1];
  __shared__ int32_t nmap_tab[kNmapCap];
  __shared__ float nmap_val[kNmapCap][3];
  const int lane = threadIdx.x;
  const int32_t gt = (int32_t)blockIdx.x;
  const int32_t s = udiv31(gt, a.T_m, a.T_sh);
  // the first pixel (bi, bj) of the lane's block, number Q of the scene's blocks
  const int32_t Q = 64 * (gt - s * a.T) + lane;
  int32_t bi, bj;
  if (Q < a.full_end) {  // a full tile per wavefront, its blocks row by row
    const int32_t tile = Q >> 6;
    const int32_t tx = udiv31(tile, a.fy_m, a.fy_sh), ty = tile - tx * a.FY;
    bi = 16 * tx + 2 * ((Q >> 3) & 7);
    bj = 16 * ty + 2 * (Q & 7);
  } else if (Q < a.bottom_end) {  // below the full tiles, rh blocks per row
    const int32_t q = Q - a.full_end;
    const int32_t row = udiv31(q, a.rh_m, a.rh_sh);
    bi = 2 * row;
    bj = 16 * a.FY + 2 * (q - row * a.rh);
  } else {  // right of them over the whole height, rw blocks per column
    const int32_t q = Q - a.bottom_end;
    const int32_t col = udiv31(q, a.rw_m, a.rw_sh);
    bi = 16 * a.FX + 2 * (q - col * a.rw);
    bj = 2 * col;
  }
  if (Q >= a.nq) bi = a.W;  // past the scene's blocks: no pixels
  const jr_scene& sc = a.scenes[s];
  const float* vp = sc.viewport;
  // reciprocals for the viewport gradients only (x and y themselves are the
  // forward's quotients)
  const float ivp0 = 1.0f / vp[0], ivp5 = 1.0f / vp[5];
  auto pixel_of = [&](int q, int32_t* i_out, int32_t* j_out) -> int32_t {
    const int32_t i = bi + (q >> 1);
    const int32_t j = bj + (q & 1);
    *i_out = i;
    *j_out = j;
    if (i >= a.W || j >= a.H) return -1;
    const int32_t t = a.prim_id[(int64_t)s * a.W * a.H + (int64_t)i * a.H + j];
    return t < a.in.n_tris ? t : -1;
  };

  // the tile's distinct primitives; the texel and specular-map sum tables
  if (lane < kBwdCap) tab[lane] = -1;
  for (int e = lane; e < kTexCap; e += 64) {
    tex_tab[e] = -1;
    tex_val[e][0] = tex_val[e][1] = tex_val[e][2] = 0.0f;
  }
  for (int e = lane; e < kSpecCap; e += 64) {
    spec_tab[e] = -1;
    spec_val[e][0] = 0.0f;
  }
  for (int e = lane; e < kNmapCap; e += 64) {
    nmap_tab[e] = -1;
    nmap_val[e][0] = nmap_val[e][1] = nmap_val[e][2] = 0.0f;
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    int32_t i, j;
    const int32_t t = pixel_of(q, &i, &j);
    if (t >= 0) (void)table_insert(tab, kBwdCap, t);
  }
  __syncthreads();
  if (lane < kBwdCap) {
    for (int k = 0; k < ACC; ++k) acc[lane][k] = 0.0f;
    const int32_t t = tab[lane];
    if (t >= 0) {
      TriSetup ts;
      tri_setup(a.in, sc, t, ts);
      float* r = trec + lane * NS;
      for (int k = 0; k < 3; ++k) {
        const float4 c = rec_col(ts, k);
        r[4 * k + 0] = c.x;
        r[4 * k + 1] = c.y;
        r[4 * k + 2] = c.z;
        r[4 * k + 3] = c.w;
      }
    }
  }
  // the corner stage: one lane per corner of each listed triangle
  if constexpr (NV > 0) {
    for (int idx = lane; idx < 3 * kBwdCap; idx += 64) {
      const int e = idx / 3, k = idx - 3 * e;
      const int32_t t = tab[e];
      if (t >= 0) Varyings<SHADER>::corner(a.in, sc, corner_slots(a.in, t), k, trec + e * NS + 12);
    }
  }
  __syncthreads();

  PixGrad g;
  for (int q = 0; q < S_COUNT; ++q) g.sg[q] = 0.0f;
  // the lane's open runs: triangle (its table entry, or -1 beyond the table),
  // texel and specular-map entry, with their sums
  int32_t run_t = -1, run_slot = -1;
  float run_v[ACC];
  for (int k = 0; k < ACC; ++k) run_v[k] = 0.0f;
  int64_t run_tex = -1, run_spec = -1;
  int32_t run_tex_n = 0;
  float run_tex_g[3] = {0.0f, 0.0f, 0.0f}, run_spec_g = 0.0f;
  int64_t run_nmap = -1;
  float run_nmap_g[3] = {0.0f, 0.0f, 0.0f};
  Pending<3, 2> pend_tex;
  Pending<1, 1> pend_spec;
  Pending<3, 2> pend_nmap;
  auto flush_tri = [&]() {
    if (run_t < 0) return;
    if (run_slot >= 0) {
      for (int k = 0; k < ACC; ++k)
        if (run_v[k] != 0.0f) atomicAdd(&acc[run_slot][k], run_v[k]);
    } else {
      a.touched[run_t] = s + 1;  // k_bwd_tris only visits touched triangles
      for (int k = 0; k < ACC; ++k)
        if (run_v[k] != 0.0f) atomicAdd(a.tacc + (int64_t)run_t * ACC + k, run_v[k]);
    }
    for (int k = 0; k < ACC; ++k) run_v[k] = 0.0f;
  };
#pragma unroll 1
  for (int q = 0; q < 4; ++q) {
    int32_t i, j;
    const int32_t t = pixel_of(q, &i, &j);
    if (t != run_t) {
      flush_tri();
      run_t = t;
      run_slot = t >= 0 ? table_insert(tab, kBwdCap, t) : -1;  // finds its entry
    }
    const int32_t e_slot = run_slot;
    const int64_t pix = (int64_t)s * a.W * a.H + (int64_t)i * a.H + j;
    g.tex_at = -1;
    g.spec_at = -1;
    g.nmap_at = -1;
    if constexpr (DARBOUX) {
      for (int k = 0; k < 9; ++k) g.dtri[k] = 0.0f;
      for (int k = 0; k < 6; ++k) g.dtuv[k] = 0.0f;
    }
    g.tex_n = 0;
    g.spec_g = 0.0f;
    for (int k = 0; k < 3; ++k) g.tex_g[k] = 0.0f;
    float vals[ACC];
    for (int k = 0; k < ACC; ++k) vals[k] = 0.0f;
    if (t >= 0) {
      float4 r0, r1, r2;
      float v[NV > 0 ? NV : 1];
      if (e_slot >= 0) {
        const float* r = trec + e_slot * NS;
        r0 = make_float4(r[0], r[1], r[2], r[3]);
        r1 = make_float4(r[4], r[5], r[6], r[7]);
        r2 = make_float4(r[8], r[9], r[10], r[11]);
        if constexpr (NV > 0)
          for (int k = 0; k < NV; ++k) v[k] = r[12 + k];
      } else {
        TriSetup ts;
        tri_setup(a.in, sc, t, ts);
        r0 = rec_col(ts, 0);
        r1 = rec_col(ts, 1);
        r2 = rec_col(ts, 2);
        if constexpr (NV > 0) Varyings<SHADER>::corners(a.in, sc, corner_slots(a.in, t), v);
      }
      const float x = pix_x(i, vp), y = pix_y(j, vp);
      const float e[3] = {edge(r0, x, y), edge(r1, x, y), edge(r2, x, y)};
      float wr = e[0] + e[1];
      wr = wr + e[2];
      const float b[3] = {e[0] / wr, e[1] / wr, e[2] / wr};
      for (int k = 0; k < 3; ++k) {
        g.db[k] = 0.0f;
        for (int c = 0; c < 3; ++c) g.dvar[k][c] = 0.0f;
      }
      if (a.d_canvas || a.target) {
        float dcol[3];
        if (a.target) {
          for (int c = 0; c < 3; ++c)
            dcol[c] = a.l2_scale * (a.canvas[3 * pix + c] - a.target[3 * pix + c]);
        } else {
          for (int c = 0; c < 3; ++c) dcol[c] = a.d_canvas[3 * pix + c];
        }
        if constexpr (NV > 0)
          ShadeBwd<SHADER>::run_v(a.in, sc, a.gr, v, b, dcol, g, s);
        else
          ShadeBwd<SHADER>::run(a.in, sc, a.gr, corner_slots(a.in, t), b, dcol, g, s);
      }
      // b = e / sum(e)
      const float tdb = (g.db[0] * e[0] + g.db[1] * e[1]) + g.db[2] * e[2];
      float de[3];
      for (int k = 0; k < 3; ++k) de[k] = g.db[k] / wr - tdb / (wr * wr);
      // zbuffer = (sum_k e_k z_k) * vp22 + vp23
      if (a.d_zbuffer) {
        const float dz = a.d_zbuffer[pix];
        float zc = e[0] * r0.w;
        zc = zc + e[1] * r1.w;
        zc = zc + e[2] * r2.w;
        const float zk[3] = {r0.w, r1.w, r2.w};
        for (int k = 0; k < 3; ++k) {
          de[k] += dz * vp[10] * zk[k];
          vals[A_Z + k] = dz * vp[10] * e[k];
        }
        g.sg[S_VP22] += dz * zc;
        g.sg[S_VP23] += dz;
      }
      // e_k = x Minv[0][k] + y Minv[1][k] + Minv[2][k]
      float dx = 0.0f, dy = 0.0f;
      const float m0[3] = {r0.x, r1.x, r2.x}, m1[3] = {r0.y, r1.y, r2.y};
      for (int k = 0; k < 3; ++k) {
        vals[A_MINV + 0 * 3 + k] = x * de[k];
        vals[A_MINV + 1 * 3 + k] = y * de[k];
        vals[A_MINV + 2 * 3 + k] = de[k];
        dx += m0[k] * de[k];
        dy += m1[k] * de[k];
      }
      // x = (i - vp03) / vp00, y = (j - vp13) / vp11
      g.sg[S_VP03] += -dx * ivp0;
      g.sg[S_VP00] += -dx * x * ivp0;
      g.sg[S_VP13] += -dy * ivp5;
      g.sg[S_VP11] += -dy * y * ivp5;
      for (int k = 0; k < 3; ++k)
        for (int c = 0; c < 3; ++c) vals[A_VAR + 3 * k + c] = g.dvar[k][c];
      if constexpr (DARBOUX) {
        for (int k = 0; k < 9; ++k) vals[A_TRI + k] = g.dtri[k];
        for (int k = 0; k < 6; ++k) vals[A_TUV + k] = g.dtuv[k];
      }
      for (int k = 0; k < ACC; ++k) run_v[k] += vals[k];
    }
    // texel and specular-map runs
    if (g.tex_at != run_tex) {
      stage_add<kTexCap, 3>(tex_tab, tex_val, a.gr.d_textures, run_tex, run_tex_n, run_tex_g,
                            pend_tex);
      run_tex = g.tex_at;
      run_tex_n = g.tex_n;
      run_tex_g[0] = run_tex_g[1] = run_tex_g[2] = 0.0f;
    }
    for (int c = 0; c < 3; ++c) run_tex_g[c] += g.tex_g[c];
    if (g.spec_at != run_spec) {
      stage_add<kSpecCap, 1>(spec_tab, spec_val, a.gr.d_specular_maps, run_spec, 1, &run_spec_g,
                             pend_spec);
      run_spec = g.spec_at;
      run_spec_g = 0.0f;
    }
    run_spec_g += g.spec_g;
    if constexpr (DARBOUX) {
      if (g.nmap_at != run_nmap) {
        stage_add<kNmapCap, 3>(nmap_tab, nmap_val, a.gr.d_normal_maps, run_nmap, 3, run_nmap_g,
                               pend_nmap);
        run_nmap = g.nmap_at;
        run_nmap_g[0] = run_nmap_g[1] = run_nmap_g[2] = 0.0f;
      }
      for (int c = 0; c < 3; ++c) run_nmap_g[c] += g.nmap_g[c];
    }
  }
  flush_tri();
  stage_add<kTexCap, 3>(tex_tab, tex_val, a.gr.d_textures, run_tex, run_tex_n, run_tex_g,
                        pend_tex);
  stage_add<kSpecCap, 1>(spec_tab, spec_val, a.gr.d_specular_maps, run_spec, 1, &run_spec_g,
                         pend_spec);
  if constexpr (DARBOUX)
    stage_add<kNmapCap, 3>(nmap_tab, nmap_val, a.gr.d_normal_maps, run_nmap, 3, run_nmap_g,
                           pend_nmap);
  if (a.gr.d_textures) pend_tex.flush(a.gr.d_textures);
  if (a.gr.d_specular_maps) pend_spec.flush(a.gr.d_specular_maps);
  if constexpr (DARBOUX)
    if (a.gr.d_normal_maps) pend_nmap.flush(a.gr.d_normal_maps);
  __syncthreads();
  // one texel's channels: tex_c is the scene's (one scene per tile)
  flush_adds<kTexCap, 3>(tex_tab, tex_val, a.gr.d_textures, sc.tex_c > 1 ? 3 : 1);
  flush_adds<kSpecCap, 1>(spec_tab, spec_val, a.gr.d_specular_maps, 1);
  if constexpr (DARBOUX) flush_adds<kNmapCap, 3>(nmap_tab, nmap_val, a.gr.d_normal_maps, 3);
  if (lane < kBwdCap && tab[lane] >= 0) a.touched[tab[lane]] = s + 1;
  // lanes take consecutive accumulators of a triangle's row: the addresses of
  // one atomic instruction fall in few cache lines
  for (int idx = lane; idx < kBwdCap * ACC; idx += 64) {
    const int e = idx / ACC, k = idx - e * ACC;
    const int32_t t = tab[e];
    if (t < 0) continue;
    const float v = acc[e][k];
    if (v != 0.0f) atomicAdd(a.tacc + (int64_t)t * ACC + k, v);
  }
  // per-scene parameters: one reduction per tile.  The pixels summed the
  // cotangents of normalise(light_direction) and normalise(light_dir_eye);
  // normalise's backward is linear in them, so it is applied to the sums.
  if (a.gr.d_scene) {
    float tot[S_COUNT];
    for (int q = 0; q < S_COUNT; ++q) tot[q] = wave_sum(g.sg[q]);
    auto through_normalise = [&](const float* v, int at) {
      if (tot[at] == 0.0f && tot[at + 1] == 0.0f && tot[at + 2] == 0.0f) return;
      float d[3];
      normalise_bwd(v, tot + at, d);
      for (int q = 0; q < 3; ++q) tot[at + q] = d[q];
    };
    through_normalise(sc.light_direction, S_LIGHT_DIR);
    through_normalise(sc.light_dir_eye, S_LIGHT_DIR_EYE);
    for (int q = 0; q < S_COUNT; ++q)
      if (lane == 0 && tot[q] != 0.0f)
        atomicAdd(a.gr.d_scene + (int64_t)s * JR_DSCENE_FLOATS + kSceneSlot[q], tot[q]);
  }
}

// vertex-stage backward of one slot for its varying cotangent dv: the
// gradient dn of the slot's normal row *nrow_out (and for gouraud dc of its
// colour row *crow_out; rows stay -1 when there is nothing to add)
template <int SHADER>
JR_DEV void vertex_bwd(const BwdArgs& a, const jr_scene& sc, int32_t slot, const float dv[3],
                       float* dW, float* dE, float* dLd, float* dLc, int64_t* nrow_out,
                       float dn[3], int64_t* crow_out, float dc[3]) {
  if (SHADER == JR_SHADER_DEPTH) return;
  if (dv[0] == 0.0f && dv[1] == 0.0f && dv[2] == 0.0f) return;
  const jr_inputs& in = a.in;
  const int32_t nrow = attr_row(in.nrm_index, slot);
  float n[3], nn[3];
  load3(in.normals, nrow, n);
  normalise3(n, nn);
  float dnn[3];
  if (SHADER == JR_SHADER_GOURAUD || SHADER == JR_SHADER_GOURAUD_TEXTURE) {
    float ln[3];
    normalise3(sc.light_direction, ln);
    const float it = dot3(nn, ln);
    float dI = 0.0f;
    if (SHADER == JR_SHADER_GOURAUD) {
      const int32_t crow = attr_row(in.col_index, slot);
      float col[3];
      load3(in.colours, crow, col);
      *crow_out = crow;
      for (int q = 0; q < 3; ++q) {
        dc[q] = dv[q] * sc.light_colour[q] * it;
        dLc[q] += dv[q] * col[q] * it;
        dI += dv[q] * col[q] * sc.light_colour[q];
      }
    } else {
      for (int q = 0; q < 3; ++q) {
        dLc[q] += dv[q] * it;
        dI += dv[q] * sc.light_colour[q];
      }
    }
    float dln[3], dL[3];
    for (int q = 0; q < 3; ++q) {
      dnn[q] = dI * ln[q];
      dln[q] = dI * nn[q];
    }
    normalise_bwd(sc.light_direction, dln, dL);
    for (int q = 0; q < 3; ++q) dLd[q] += dL[q];
  } else {
    // ne = normalise((E @ [normalise(nn), 0])[:3])  (Camera.apply_vec)
    float u2[3];
    normalise3(nn, u2);
    float h[4];
    apply4(sc.world_to_eye_norm, u2[0], u2[1], u2[2], 0.0f, h);
    const float t3[3] = {h[0], h[1], h[2]};
    float dt3[3], du2[3], du1[3];
    normalise_bwd(t3, dv, dt3);
    for (int jj = 0; jj < 3; ++jj) {
      float s = 0.0f;
      for (int ii = 0; ii < 3; ++ii) {
        s += sc.world_to_eye_norm[4 * ii + jj] * dt3[ii];
        dE[4 * ii + jj] += dt3[ii] * u2[jj];
      }
      du2[jj] = s;
    }
    normalise_bwd(nn, du2, du1);
    for (int q = 0; q < 3; ++q) dnn[q] = du1[q];
  }
  normalise_bwd(n, dnn, dn);
  *nrow_out = nrow;
}

// Adds each lane's 3-float row g at d + 3 row (row < 0: nothing) with
// neighbouring lanes on the three floats of one row, so the adds of an atomic
// instruction share cache lines.  W: the wavefront's LDS staging; call with
// every lane of the wavefront.
struct WaveRows {
  int64_t at[64];
  float v[64][3];
};
JR_DEV void wave_row_adds(WaveRows& W, float* __restrict__ d, int64_t row, const float g[3]) {
  const int lane = threadIdx.x & 63;
  W.at[lane] = row;
  W.v[lane][0] = g[0];
  W.v[lane][1] = g[1];
  W.v[lane][2] = g[2];
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
  for (int it = 0; it < 3; ++it) {
    const int idx = lane + 64 * it;
    const int e = idx / 3, c = idx - 3 * e;
    const int64_t r = W.at[e];
    const float x = W.v[e][c];
    if (r >= 0 && x != 0.0f) atomicAdd(d + 3 * r + c, x);
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Triangle t (-1: none) of the calling lane; the accumulators are read and
// cleared only when `touched`.  Every lane computes its scene, so that a
// wavefront of one scene reduces its camera and light gradients once.
template <int SHADER>
JR_DEV void bwd_tri(const BwdArgs& a, int32_t t, int32_t s, bool touched, WaveRows& W) {
  float dW[16], dE[16], dLd[3], dLc[3];
  for (int q = 0; q < 16; ++q) dW[q] = dE[q] = 0.0f;
  for (int q = 0; q < 3; ++q) dLd[q] = dLc[q] = 0.0f;
  if (__ballot(touched) == 0) return;  // nothing to add for this wavefront
  if (!touched) s = -1;
  constexpr int ACC = AccOf<SHADER>::N;
  float v[ACC];
  bool live = false;
  TriSetup ts;
  float dM[9];
  if (touched) {
    float* A = a.tacc + (int64_t)t * ACC;
    for (int q = 0; q < ACC; ++q) {
      v[q] = A[q];
      A[q] = 0.0f;
      live = live || (v[q] != 0.0f);
    }
    a.touched[t] = 0;
    if (live) {
      tri_setup(a.in, a.scenes[s], t, ts);
      // X = M^-1, G = dL/dX: dL/dM = -X^T G X^T
      const float* X = ts.Mi;
      float XtG[9];
      for (int p = 0; p < 3; ++p)
        for (int k = 0; k < 3; ++k) {
          float s2 = 0.0f;
          for (int r = 0; r < 3; ++r) s2 += X[3 * r + p] * v[A_MINV + 3 * r + k];
          XtG[3 * p + k] = s2;
        }
      for (int p = 0; p < 3; ++p)
        for (int q = 0; q < 3; ++q) {
          float s2 = 0.0f;
          for (int k = 0; k < 3; ++k) s2 += XtG[3 * p + k] * X[3 * q + k];
          dM[3 * p + q] = -s2;
        }
    }
  }
  // corner by corner, the lanes' position / normal / colour rows added
  // together (wave_row_adds), every lane taking part
  for (int k = 0; k < 3; ++k) {
    int64_t prow = -1, nrow = -1, crow = -1;
    float dp[3] = {0.0f, 0.0f, 0.0f}, dn[3] = {0.0f, 0.0f, 0.0f}, dc[3] = {0.0f, 0.0f, 0.0f};
    if (live) {
      const jr_scene& sc = a.scenes[s];
      const int32_t slot = face_slot(a.in, t, k);
      const float dclip[4] = {dM[3 * k + 0], dM[3 * k + 1], v[A_Z + k], dM[3 * k + 2]};
      prow = attr_row(a.in.pos_index, slot);
      float p[3];
      load3(a.in.positions, prow, p);
      const float ph[4] = {p[0], p[1], p[2], 1.0f};
      for (int jj = 0; jj < 4; ++jj) {
        float s2 = 0.0f;
        for (int ii = 0; ii < 4; ++ii) {
          s2 += sc.world_to_clip[4 * ii + jj] * dclip[ii];
          dW[4 * ii + jj] += dclip[ii] * ph[jj];
        }
        if (jj < 3) dp[jj] = s2;
      }
      const float dv[3] = {v[A_VAR + 3 * k], v[A_VAR + 3 * k + 1], v[A_VAR + 3 * k + 2]};
      vertex_bwd<SHADER>(a, sc, slot, dv, dW, dE, dLd, dLc, &nrow, dn, &crow, dc);
    }
    if (a.gr.d_positions) wave_row_adds(W, a.gr.d_positions, prow, dp);
    if constexpr (SHADER != JR_SHADER_DEPTH)
      if (a.gr.d_normals) wave_row_adds(W, a.gr.d_normals, nrow, dn);
    if constexpr (SHADER == JR_SHADER_GOURAUD)
      if (a.gr.d_colours) wave_row_adds(W, a.gr.d_colours, crow, dc);
  }
  if (live) {
    {
      const jr_scene& sc = a.scenes[s];
      if constexpr (SHADER == JR_SHADER_PHONG_DARBOUX) {
        // the face of slot 0 (id_to_face) whose NDC corners and uvs built the
        // tangent frame: t_m = to_cartesian(h_m), h_m = world_to_clip @ [p_m, 1]
        const int32_t face = a.in.id_to_face[face_slot(a.in, t, 0)];
        for (int m = 0; m < 3; ++m) {
          const int32_t q = a.in.face_slots[3 * (int64_t)face + m];
          const float* dt = v + A_TRI + 3 * m;
          if (dt[0] != 0.0f || dt[1] != 0.0f || dt[2] != 0.0f) {
            const int32_t prow = attr_row(a.in.pos_index, q);
            float p[3], h[4];
            load3(a.in.positions, prow, p);
            position_clip(a.in, sc.world_to_clip, q, h);
            // where(w == 0, h[:3], h[:3] / w): x / w has d/dx = 1 / w and
            // d/dw = -x / w^2
            float dh[4];
            if (h[3] == 0.0f) {
              dh[0] = dt[0];
              dh[1] = dt[1];
              dh[2] = dt[2];
              dh[3] = 0.0f;
            } else {
              float dw = 0.0f;
              for (int c = 0; c < 3; ++c) {
                dh[c] = dt[c] / h[3];
                dw += -dt[c] * h[c] / (h[3] * h[3]);
              }
              dh[3] = dw;
            }
            const float ph[4] = {p[0], p[1], p[2], 1.0f};
            for (int jj = 0; jj < 4; ++jj) {
              float s2 = 0.0f;
              for (int ii = 0; ii < 4; ++ii) {
                s2 += sc.world_to_clip[4 * ii + jj] * dh[ii];
                dW[4 * ii + jj] += dh[ii] * ph[jj];
              }
              if (jj < 3 && a.gr.d_positions)
                atomicAdd(a.gr.d_positions + 3 * (int64_t)prow + jj, s2);
            }
          }
          const float* du = v + A_TUV + 2 * m;
          if (a.gr.d_uvs && (du[0] != 0.0f || du[1] != 0.0f)) {
            const int32_t urow = attr_row(a.in.uv_index, q);
            atomicAdd(a.gr.d_uvs + 2 * (int64_t)urow, du[0]);
            atomicAdd(a.gr.d_uvs + 2 * (int64_t)urow + 1, du[1]);
          }
        }
      }
    }
  }
  if (!a.gr.d_scene) return;
  // reduce over the wave when it holds one scene, else per lane
  const int32_t s0 = __shfl(s, 0);
  const bool uniform = __ballot(s == s0 || s < 0) == __ballot(true) && s0 >= 0;
  float* base = a.gr.d_scene;
  if (uniform) {
    const int lane = threadIdx.x & 63;
    float* row = base + (int64_t)s0 * JR_DSCENE_FLOATS;
    for (int q = 0; q < 16; ++q) {
      const float w = wave_sum(dW[q]);
      const float e = wave_sum(dE[q]);
      if (lane == 0) {
        if (w != 0.0f) atomicAdd(row + JR_DSCENE_WORLD_TO_CLIP + q, w);
        if (e != 0.0f) atomicAdd(row + JR_DSCENE_WORLD_TO_EYE_NORM + q, e);
      }
    }
    for (int q = 0; q < 3; ++q) {
      const float l = wave_sum(dLd[q]);
      const float c = wave_sum(dLc[q]);
      if (lane == 0) {
        if (l != 0.0f) atomicAdd(row + JR_DSCENE_LIGHT_DIRECTION + q, l);
        if (c != 0.0f) atomicAdd(row + JR_DSCENE_LIGHT_COLOUR + q, c);
      }
    }
  } else if (s >= 0) {
    float* row = base + (int64_t)s * JR_DSCENE_FLOATS;
    for (int q = 0; q < 16; ++q) {
      if (dW[q] != 0.0f) atomicAdd(row + JR_DSCENE_WORLD_TO_CLIP + q, dW[q]);
      if (dE[q] != 0.0f) atomicAdd(row + JR_DSCENE_WORLD_TO_EYE_NORM + q, dE[q]);
    }
    for (int q = 0; q < 3; ++q) {
      if (dLd[q] != 0.0f) atomicAdd(row + JR_DSCENE_LIGHT_DIRECTION + q, dLd[q]);
      if (dLc[q] != 0.0f) atomicAdd(row + JR_DSCENE_LIGHT_COLOUR + q, dLc[q]);
    }
  }
}

// One thread per triangle: the touched ones (flag = scene + 1) are packed to
// the front of the workgroup, read their accumulators and clear them with the
// flag for the next call; wavefronts left without one return at once.  At
// most 80 VGPRs (6 wavefronts per SIMD instead of 5): the kernel waits on its
// gathers most of the time, and more wavefronts hide more of that.
template <int SHADER>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6, 8)))
void k_bwd_tris(BwdArgs a) {
  __shared__ int2 items[256];
  __shared__ int32_t wave_n[4];
  __shared__ WaveRows rows[4];
  const int32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  const int32_t flag = t < a.in.n_tris ? a.touched[t] : 0;
  const uint64_t bal = __ballot(flag != 0);
  if ((threadIdx.x & 63) == 0) wave_n[threadIdx.x >> 6] = __popcll(bal);
  __syncthreads();
  int32_t pos = __builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32),
                                          __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
  for (int w = 0; w < (int)(threadIdx.x >> 6); ++w) pos += wave_n[w];
  if (flag != 0) items[pos] = make_int2(t, flag - 1);
  __syncthreads();
  const int32_t n = wave_n[0] + wave_n[1] + wave_n[2] + wave_n[3];
  const bool has = (int32_t)threadIdx.x < n;
  const int2 it = has ? items[threadIdx.x] : make_int2(-1, -1);
  bwd_tri<SHADER>(a, it.x, it.y, has, rows[threadIdx.x >> 6]);
}

// Reads the workspace header: if it does not describe this workspace (new,
// resized, or used for another triangle count or accumulator size), asks
// k_bwd_fill to clear the flags and accumulators.
__global__ void k_bwd_prepare(uint32_t* __restrict__ hdr, int32_t n_tris, int32_t acc) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const bool ok = hdr[kHdrMagicLo] == kMagicLo && hdr[kHdrMagicHi] == kMagicHi &&
                  hdr[kHdrTris] == (uint32_t)n_tris && hdr[kHdrAcc] == (uint32_t)acc;
  hdr[kHdrFill] = ok ? 0u : 1u;
  hdr[kHdrMagicLo] = kMagicLo;
  hdr[kHdrMagicHi] = kMagicHi;
  hdr[kHdrTris] = (uint32_t)n_tris;
  hdr[kHdrAcc] = (uint32_t)acc;
}

// Zeroes n 32-bit words after the header when k_bwd_prepare asked for it.
__global__ void k_bwd_fill(const uint32_t* __restrict__ hdr, uint32_t* __restrict__ p,
                           int64_t n) {
  if (hdr[kHdrFill] == 0u) return;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0u;
}


thread_local std::string g_bwd_err;

int32_t bfail(int32_t code, const char* msg) {
  g_bwd_err = msg;
  return code;
}

template <int SHADER>
void launch_bwd(const BwdArgs& a, hipStream_t st) {
  const int64_t n_waves = (int64_t)a.T * a.n_scenes;
  hipLaunchKernelGGL(k_bwd_pixels<SHADER>, dim3((unsigned)n_waves), dim3(64), 0, st, a);
  hipLaunchKernelGGL(k_bwd_tris<SHADER>, dim3((a.in.n_tris + 255) / 256), dim3(256), 0, st, a);
}

// Workspace: header (256 B), touched flags (4 B per triangle), accumulators
// (4 acc_floats(shader) B per triangle).
struct BwdLayout {
  int64_t touched, tacc, total;
};

BwdLayout bwd_layout(int32_t n_tris, int shader) {
  const int64_t n = std::max(n_tris, 1);
  BwdLayout L;
  L.touched = 256;
  L.tacc = L.touched + (4 * n + 255) / 256 * 256;
  L.total = L.tacc + (4 * (int64_t)acc_floats(shader) * n + 255) / 256 * 256;
  return L;
}

}  // namespace jr

using namespace jr;

extern "C" {

int64_t jr_backward_workspace_size(const jr_frame* frame, int32_t n_tris) {
  if (!frame || n_tris < 0) return -1;
  return bwd_layout(n_tris, frame->shader).total;
}

const char* jr_backward_last_error(void) { return g_bwd_err.c_str(); }

}  // extern "C"

namespace jr {
namespace {

// jr_backward and jr_backward_l2: the canvas cotangent is d_canvas, or
// l2_scale * (canvas - target) when target is set.
int32_t backward_body(const jr_frame* fr, const jr_inputs* in, const jr_scene* scenes,
                      const float* d_canvas, const float* canvas, const float* target,
                      float l2_scale, const float* d_zbuffer, const int32_t* prim_id,
                      jr_grads* grads, void* workspace, int64_t workspace_bytes,
                      jr_stream_t stream_) {
  if (!fr || !in || !scenes || !grads || !prim_id)
    return bfail(JR_E_INVALID, "frame, inputs, scenes, grads and prim_id are required");
  if (fr->width <= 0 || fr->height <= 0 || fr->n_scenes <= 0 || in->n_tris < 0)
    return bfail(JR_E_INVALID, "bad frame sizes");
  if (in->clip) return bfail(JR_E_UNSUPPORTED, "backward needs positions, not clip input");
  const int sh = fr->shader;
  if (sh == JR_SHADER_VISIBILITY || sh < 0 || sh >= JR_SHADER_COUNT)
    return bfail(JR_E_UNSUPPORTED, "shader has no backward pass");
  if (!d_canvas && !target && !d_zbuffer) return bfail(JR_E_INVALID, "no upstream gradient");
  if (!in->positions) return bfail(JR_E_INVALID, "positions are required");
  if (!in->faces && (int64_t)in->n_slots < 3 * (int64_t)in->n_tris)
    return bfail(JR_E_INVALID, "without faces, n_slots must be at least 3 n_tris");
  if (sh == JR_SHADER_GOURAUD && (!in->normals || !in->colours))
    return bfail(JR_E_INVALID, "gouraud needs normals and colours");
  if ((sh == JR_SHADER_GOURAUD_TEXTURE || sh == JR_SHADER_PHONG ||
       sh == JR_SHADER_PHONG_DARBOUX || sh == JR_SHADER_PHONG_REFLECTION ||
       sh == JR_SHADER_PHONG_REFLECTION_SHADOW) &&
      (!in->normals || !in->uvs || !in->textures))
    return bfail(JR_E_INVALID, "textured shaders need normals, uvs and textures");
  if (sh == JR_SHADER_PHONG_DARBOUX && (!in->id_to_face || !in->face_slots || !in->normal_maps))
    return bfail(JR_E_INVALID, "phong_darboux needs id_to_face, face_slots and normal_maps");
  if ((sh == JR_SHADER_PHONG_REFLECTION || sh == JR_SHADER_PHONG_REFLECTION_SHADOW) &&
      (!in->tex_index || !in->texture_shapes || !in->specular_maps))
    return bfail(JR_E_INVALID, "phong_reflection needs tex_index, texture_shapes, specular_maps");
  if (sh == JR_SHADER_PHONG_REFLECTION_SHADOW &&
      (!in->shadow_maps || in->shadow_w <= 0 || in->shadow_h <= 0))
    return bfail(JR_E_INVALID, "phong_reflection_shadow needs shadow maps");
  const int64_t need = jr_backward_workspace_size(fr, in->n_tris);
  if (!workspace || workspace_bytes < need) return bfail(JR_E_WORKSPACE, "workspace too small");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream_);
  BwdArgs a;
  a.in = *in;
  a.scenes = scenes;
  a.gr = *grads;
  a.d_canvas = (sh == JR_SHADER_DEPTH) ? nullptr : d_canvas;
  a.canvas = canvas;
  a.target = (sh == JR_SHADER_DEPTH) ? nullptr : target;
  a.l2_scale = l2_scale;
  a.d_zbuffer = d_zbuffer;
  a.prim_id = prim_id;
  const BwdLayout WL = bwd_layout(in->n_tris, sh);
  char* ws = static_cast<char*>(workspace);
  a.header = reinterpret_cast<uint32_t*>(ws);
  a.touched = reinterpret_cast<int32_t*>(ws + WL.touched);
  a.tacc = reinterpret_cast<float*>(ws + WL.tacc);
  a.W = fr->width;
  a.H = fr->height;
  a.n_scenes = fr->n_scenes;
  const int64_t QW = (a.W + 1) / 2, QH = (a.H + 1) / 2, nq = QW * QH;
  const int64_t tiles_y = (QH + 7) / 8, tiles = (QW + 7) / 8 * tiles_y;
  // the leftover blocks are packed where that saves at least an eighth of the
  // wavefronts (84 x 84: 28 instead of 36, fwd+bwd 1.58 -> 1.52 ms on 1024
  // scenes); where it saves few (480 x 270: 507 instead of 510) plain tiles,
  // partial ones included, measured faster (the backward 3.88 against 3.98 ms
  // on 512 scenes).  Plain tiles are the first kind of block with FY = tiles_y
  // and every wavefront's blocks counted as a full tile's.
  const bool packed = 8 * ((nq + 63) / 64) <= 7 * tiles;
  const int64_t T = packed ? (nq + 63) / 64 : tiles;
  // block numbers (below 64 T) and wavefront numbers stay below 2^31
  if (64 * T > 0x7FFFFFFFLL || T * a.n_scenes > 0x7FFFFFFFLL)
    return bfail(JR_E_INVALID, "too many pixels");
  a.T = (int32_t)T;
  if (packed) {
    a.nq = (int32_t)nq;
    a.FX = (int32_t)(QW / 8);
    a.FY = (int32_t)(QH / 8);
    a.rw = (int32_t)(QW % 8);
    a.rh = (int32_t)(QH % 8);
    a.full_end = 64 * a.FX * a.FY;
    a.bottom_end = a.full_end + 8 * a.FX * a.rh;
  } else {
    a.nq = a.full_end = a.bottom_end = (int32_t)(64 * T);
    a.FX = (int32_t)(tiles / tiles_y);
    a.FY = (int32_t)tiles_y;
    a.rw = a.rh = 0;
  }
  magic31((uint32_t)a.T, &a.T_m, &a.T_sh);
  magic31((uint32_t)std::max(a.FY, 1), &a.fy_m, &a.fy_sh);
  magic31((uint32_t)std::max(a.rh, 1), &a.rh_m, &a.rh_sh);
  magic31((uint32_t)std::max(a.rw, 1), &a.rw_m, &a.rw_sh);
  if (in->n_tris == 0) return JR_OK;
  // flags and accumulators: everything after the header
  const int64_t nwords = (WL.total - WL.touched) / 4;
  hipLaunchKernelGGL(k_bwd_prepare, dim3(1), dim3(64), 0, st, a.header, in->n_tris,
                     acc_floats(sh));
  hipLaunchKernelGGL(k_bwd_fill, dim3((unsigned)std::min<int64_t>((nwords + 255) / 256, 8192)),
                     dim3(256), 0, st, a.header,
                     reinterpret_cast<uint32_t*>(ws + WL.touched), nwords);
  switch (sh) {
    case JR_SHADER_DEPTH: launch_bwd<JR_SHADER_DEPTH>(a, st); break;
    case JR_SHADER_GOURAUD: launch_bwd<JR_SHADER_GOURAUD>(a, st); break;
    case JR_SHADER_GOURAUD_TEXTURE: launch_bwd<JR_SHADER_GOURAUD_TEXTURE>(a, st); break;
    case JR_SHADER_PHONG: launch_bwd<JR_SHADER_PHONG>(a, st); break;
    case JR_SHADER_PHONG_DARBOUX: launch_bwd<JR_SHADER_PHONG_DARBOUX>(a, st); break;
    case JR_SHADER_PHONG_REFLECTION: launch_bwd<JR_SHADER_PHONG_REFLECTION>(a, st); break;
    default: launch_bwd<JR_SHADER_PHONG_REFLECTION_SHADOW>(a, st); break;
  }
  const hipError_t err = hipGetLastError();
  if (err != hipSuccess) return bfail(JR_E_HIP, hipGetErrorString(err));
  return JR_OK;
}

}  // namespace
}  // namespace jr

extern "C" {

int32_t jr_backward(const jr_frame* fr, const jr_inputs* in, const jr_scene* scenes,
                    const float* d_canvas, const float* d_zbuffer, const int32_t* prim_id,
                    jr_grads* grads, void* workspace, int64_t workspace_bytes,
                    jr_stream_t stream) {
  return backward_body(fr, in, scenes, d_canvas, nullptr, nullptr, 0.0f, d_zbuffer, prim_id,
                       grads, workspace, workspace_bytes, stream);
}

int32_t jr_backward_l2(const jr_frame* fr, const jr_inputs* in, const jr_scene* scenes,
                       const float* canvas, const float* target, float scale,
                       const float* d_zbuffer, const int32_t* prim_id, jr_grads* grads,
                       void* workspace, int64_t workspace_bytes, jr_stream_t stream) {
  if (!canvas || !target) return bfail(JR_E_INVALID, "canvas and target are required");
  return backward_body(fr, in, scenes, nullptr, canvas, target, scale, d_zbuffer, prim_id,
                       grads, workspace, workspace_bytes, stream);
}

}  // extern "C"
