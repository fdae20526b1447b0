// jr_shade.h -- per-fragment work of the built-in shaders, device side.
//
// Given the primitive chosen for a pixel and its perspective-correct
// barycentrics, recompute the vertex-shader varyings of its three slots,
// interpolate them (Shader.interpolate, SMOOTH: renderer/shader.py:253-290)
// and run the shader's fragment stage.  Returns the fragment's `keeps` and
// writes the RGB colour.  Formula order follows the reference files cited at
// each shader.
#pragma once

#include "../../include/jrhip.h"
#include "jr_math.h"

namespace jr {

struct Corner {
  int32_t slot[3];
};

JR_DEV int32_t attr_row(const int32_t* __restrict__ index, int32_t slot) {
  return index ? index[slot] : slot;
}

JR_DEV void load3(const float* __restrict__ base, int32_t row, float v[3]) {
  v[0] = base[3 * (int64_t)row + 0];
  v[1] = base[3 * (int64_t)row + 1];
  v[2] = base[3 * (int64_t)row + 2];
}

JR_DEV void load2(const float* __restrict__ base, int32_t row, float v[2]) {
  v[0] = base[2 * (int64_t)row + 0];
  v[1] = base[2 * (int64_t)row + 1];
}

// Interpolation.SMOOTH: dot_general(coef, values) = (b0 v0 + b1 v1) + b2 v2
// (renderer/geometry.py:104-108).
JR_DEV float interp(const float b[3], float v0, float v1, float v2) {
  float s = b[0] * v0;
  s = s + b[1] * v1;
  s = s + b[2] * v2;
  return s;
}

// Diffuse lookup `texture[u, v]` for a [tex_w][tex_h][tex_c] map whose
// indices are already in range; tex_c is 1 or 3 (broadcast to RGB).
JR_DEV void tex_rgb(const float* __restrict__ tex, int64_t base, int32_t th,
                    int32_t tc, int32_t u, int32_t v, float out[3]) {
  const int64_t o = base + ((int64_t)u * th + v) * tc;
  out[0] = tex[o];
  out[1] = tex[o + (tc > 1 ? 1 : 0)];
  out[2] = tex[o + (tc > 1 ? 2 : 0)];
}

// Gouraud vertex stage (shaders/gouraud.py:54-88): colour * light.colour *
// dot(normalise(n), normalise(light.direction)).
JR_DEV void gouraud_vertex(const jr_inputs& in, const jr_scene& sc, int32_t slot,
                           float col[3]) {
  float n[3], nn[3], ln[3], c[3];
  load3(in.normals, attr_row(in.nrm_index, slot), n);
  normalise3(n, nn);
  normalise3(sc.light_direction, ln);
  const float intensity = dot3(nn, ln);
  load3(in.colours, attr_row(in.col_index, slot), c);
  col[0] = (c[0] * sc.light_colour[0]) * intensity;
  col[1] = (c[1] * sc.light_colour[1]) * intensity;
  col[2] = (c[2] * sc.light_colour[2]) * intensity;
}

// Eye-space normal of the phong family's vertex stage
// (shaders/phong.py:74-104): apply_vec(normalise(n), world_to_eye_norm), of
// a loaded normal n.
JR_DEV void eye_normal_of(const float n[3], const jr_scene& sc, float out[3]) {
  float nn[3];
  normalise3(n, nn);
  apply_vec(sc.world_to_eye_norm, nn, out);
}
JR_DEV void eye_normal(const jr_inputs& in, const jr_scene& sc, int32_t slot,
                       float out[3]) {
  float n[3];
  load3(in.normals, attr_row(in.nrm_index, slot), n);
  eye_normal_of(n, sc, out);
}

// A slot's attribute rows; with every index array present (the usual case)
// the three index loads issue together.
struct SlotRows {
  int32_t pos, nrm, uv;
};
JR_DEV SlotRows slot_rows(const jr_inputs& in, int32_t slot) {
  SlotRows r;
  if (in.pos_index && in.nrm_index && in.uv_index) {
    r.pos = in.pos_index[slot];
    r.nrm = in.nrm_index[slot];
    r.uv = in.uv_index[slot];
  } else {
    r.pos = attr_row(in.pos_index, slot);
    r.nrm = attr_row(in.nrm_index, slot);
    r.uv = attr_row(in.uv_index, slot);
  }
  return r;
}

JR_DEV void position_clip(const jr_inputs& in, const float* M, int32_t slot,
                          float out[4]) {
  float p[3];
  load3(in.positions, attr_row(in.pos_index, slot), p);
  apply4(M, p[0], p[1], p[2], 1.0f, out);
}

// Shadow.get (renderer/shadow.py:127-153): round half away from zero, negative
// indices count from the end, anything else out of range reads +inf.
JR_DEV float shadow_get(const float* __restrict__ map, int32_t sw, int32_t sh,
                        float x, float y) {
  int32_t ix = f2i_sat(roundf(x));
  int32_t iy = f2i_sat(roundf(y));
  if (ix < 0) ix += sw;
  if (iy < 0) iy += sh;
  if (ix < 0 || ix >= sw || iy < 0 || iy >= sh) return __builtin_inff();
  return map[(int64_t)ix * sh + iy];
}

// Depth of the shadow map at light-viewport position (x, y) of a scene, read
// from a finished map (in.shadow_maps).  A raster may hand the shadow shader
// another source of the same depths.
struct MapShadowDepth {
  // a read whose result the colour does not depend on may be skipped
  JR_DEV bool required() const { return false; }
  JR_DEV float operator()(const jr_inputs& in, int scene, float x, float y) const {
    const float* map = in.shadow_maps + (int64_t)scene * in.shadow_w * in.shadow_h;
    return shadow_get(map, in.shadow_w, in.shadow_h, x, y);
  }
};

// Phong reflection fragment core shared by the two reflection shaders
// (shaders/phong_reflection.py:154-235, phong_reflection_shadow.py:173-271).
// Computes tex colour, diffuse scalar and specular scalar.
// l: normalise(light_dir_eye), the same for every fragment of a scene.
// texture_shape[texture_index] (the gather clamps the row), per triangle.
JR_DEV void texture_shape(const jr_inputs& in, const jr_scene& sc, int32_t tex_index,
                          float* s0, float* s1) {
  const int32_t row = index_clamp(tex_index, sc.texshape_count);
  *s0 = __int_as_float(in.texture_shapes[2 * (int64_t)(sc.texshape_base + row) + 0]);
  *s1 = __int_as_float(in.texture_shapes[2 * (int64_t)(sc.texshape_base + row) + 1]);
}

// s0, s1: texture_shape[texture_index] (texture_shape above).
JR_DEV void reflection_terms(const jr_inputs& in, const jr_scene& sc,
                             const float n_i[3], const float uv[2],
                             int32_t tex_index, int32_t s0, int32_t s1, const float l[3],
                             float tex[3], float* diffuse, float* specular) {
  // MergedModel.uv_repeat (renderer/model.py:303-339)
  float f0 = modf_frac(uv[0]);
  float f1 = modf_frac(uv[1]);
  f0 = (f0 < 0.0f) ? f0 + 1.0f : f0;
  f1 = (f1 < 0.0f) ? f1 + 1.0f : f1;
  const int32_t shift = (int32_t)((uint32_t)tex_index * (uint32_t)sc.tex_offset);
  const float r0 = f0 * (float)s0 + (float)shift;
  const float r1 = f1 * (float)s1;
  const int32_t u0 = f2i_sat(floorf(r0));
  const int32_t v0 = f2i_sat(floorf(r1));
  const int32_t u = index_clamp(u0, sc.tex_w);
  const int32_t v = index_clamp(v0, sc.tex_h);
  tex_rgb(in.textures, sc.tex_base, sc.tex_h, sc.tex_c, u, v, tex);
  const int32_t us = index_clamp(u0, sc.spec_w);
  const int32_t vs = index_clamp(v0, sc.spec_h);
  const float spec_exp = in.specular_maps[sc.spec_base + (int64_t)us * sc.spec_h + vs];

  float n[3];
  normalise3(n_i, n);
  const float d = dot3(n, l);
  *diffuse = max_nan(d, 0.0f);
  const float two_d = 2.0f * d;
  float r[3];
  r[0] = two_d * n[0] - l[0];
  r[1] = two_d * n[1] - l[1];
  r[2] = two_d * n[2] - l[2];
  float rn[3];
  normalise3(r, rn);
  *specular = pow_unit(max_nan(rn[2], 0.0f), spec_exp);
}

template <int SHADER>
struct Shade;

template <>
struct Shade<JR_SHADER_DEPTH> {
  static JR_DEV bool run(const jr_inputs&, const jr_scene&, const Corner&,
                         const float*, bool, float*, int) {
    return true;
  }
};

template <>
struct Shade<JR_SHADER_VISIBILITY> {
  static JR_DEV bool run(const jr_inputs&, const jr_scene&, const Corner&,
                         const float*, bool, float*, int) {
    return true;
  }
};




// shaders/phong_darboux.py:116-293
template <>
struct Shade<JR_SHADER_PHONG_DARBOUX> {
  static JR_DEV bool run(const jr_inputs& in, const jr_scene& sc,
                         const Corner& c, const float b[3], bool front,
                         float col[3], int) {
    float ne[3][3], uvv[3][2];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      eye_normal(in, sc, c.slot[k], ne[k]);
      load2(in.uvs, attr_row(in.uv_index, c.slot[k]), uvv[k]);
    }
    // triangle (NDC) and triangle_uv of the face of slot 0, not interpolated
    const int32_t face = in.id_to_face[c.slot[0]];
    float tri[3][3], tuv[3][2];
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      const int32_t q = in.face_slots[3 * (int64_t)face + m];
      float h[4];
      position_clip(in, sc.world_to_clip, q, h);
      // to_cartesian: drop w without dividing when w == 0
      const bool wz = (h[3] == 0.0f);
      tri[m][0] = wz ? h[0] : h[0] / h[3];
      tri[m][1] = wz ? h[1] : h[1] / h[3];
      tri[m][2] = wz ? h[2] : h[2] / h[3];
      load2(in.uvs, attr_row(in.uv_index, q), tuv[m]);
    }
    float n_i[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) n_i[q] = interp(b, ne[0][q], ne[1][q], ne[2][q]);
    const float uu = interp(b, uvv[0][0], uvv[1][0], uvv[2][0]);
    const float vv = interp(b, uvv[0][1], uvv[1][1], uvv[2][1]);
    const int32_t u = floor_mod(f2i_sat(floorf(uu)), sc.tex_w);
    const int32_t v = floor_mod(f2i_sat(floorf(vv)), sc.tex_h);
    float n[3];
    normalise3(n_i, n);
    float A[9], AI[9];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      A[q] = tri[1][q] - tri[0][q];
      A[3 + q] = tri[2][q] - tri[0][q];
      A[6 + q] = n[q];
    }
    inv3(A, AI);
    const float du1 = tuv[1][0] - tuv[0][0], du2 = tuv[2][0] - tuv[0][0];
    const float dv1 = tuv[1][1] - tuv[0][1], dv2 = tuv[2][1] - tuv[0][1];
    float iv[3], jv[3];
#pragma unroll
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
    // B = [normalise(i), normalise(j), n] as columns; B @ normal_map[u, v]
    const int32_t un = index_clamp(u, sc.nmap_w);
    const int32_t vn = index_clamp(v, sc.nmap_h);
    float nm[3];
    load3(in.normal_maps + sc.nmap_base, un * sc.nmap_h + vn, nm);
    float bn[3];
#pragma unroll
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
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) {
      lc[ch] = sc.light_colour[ch] * d;
      pos = pos && (lc[ch] >= 0.0f);
    }
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) col[ch] = pos ? tex[ch] * lc[ch] : 0.0f;
    return front;
  }
};

// Shaders whose fragment stage splits into a per-triangle part (the vertex
// stage outputs of the three corners, packed into N floats) and a per-pixel
// part.  A tile computes the per-triangle part once for each distinct winning
// triangle and shares it among that triangle's pixels; the result is the same
// arithmetic as running both parts for every pixel.  N == 0: not split.
template <int SHADER>
struct Varyings {
  static constexpr int N = 0;
};

// shaders/gouraud.py:58-124: per corner the lit vertex colour.
template <>
struct Varyings<JR_SHADER_GOURAUD> {
  static constexpr int N = 9;
  // corner k's entries of v (another lane may write the other corners')
  static JR_DEV void corner(const jr_inputs& in, const jr_scene& sc, const Corner& c, int k,
                            float* v) {
    gouraud_vertex(in, sc, c.slot[k], v + 3 * k);
  }
  static JR_DEV void corners(const jr_inputs& in, const jr_scene& sc, const Corner& c,
                             float v[N]) {
#pragma unroll
    for (int k = 0; k < 3; ++k) corner(in, sc, c, k, v);
  }
  static JR_DEV void scene_terms(const jr_scene&, float l[3]) { l[0] = l[1] = l[2] = 0.0f; }
  static JR_DEV bool pixel(const jr_inputs&, const jr_scene&, const float v[N], const float*,
                           const float b[3], bool front, float col[3], int) {
    bool pos = true;
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) {
      col[ch] = interp(b, v[ch], v[3 + ch], v[6 + ch]);
      pos = pos && (col[ch] >= 0.0f);
    }
    return front && pos;
  }
};

// shaders/gouraud_texture.py:68-144: per corner the light colour times the
// lambertian intensity, and the uv.
template <>
struct Varyings<JR_SHADER_GOURAUD_TEXTURE> {
  static constexpr int N = 15;
  static JR_DEV void corner(const jr_inputs& in, const jr_scene& sc, const Corner& c, int k,
                            float* v) {
    const SlotRows r = slot_rows(in, c.slot[k]);
    float nn_l[3], n[3], nn[3];
    load3(in.normals, r.nrm, n);
    load2(in.uvs, r.uv, v + 9 + 2 * k);
    normalise3(sc.light_direction, nn_l);
    normalise3(n, nn);
    const float intensity = dot3(nn, nn_l);
    v[3 * k + 0] = sc.light_colour[0] * intensity;
    v[3 * k + 1] = sc.light_colour[1] * intensity;
    v[3 * k + 2] = sc.light_colour[2] * intensity;
  }
  static JR_DEV void corners(const jr_inputs& in, const jr_scene& sc, const Corner& c,
                             float v[N]) {
#pragma unroll
    for (int k = 0; k < 3; ++k) corner(in, sc, c, k, v);
  }
  static JR_DEV void scene_terms(const jr_scene&, float l[3]) { l[0] = l[1] = l[2] = 0.0f; }
  static JR_DEV bool pixel(const jr_inputs& in, const jr_scene& sc, const float v[N],
                           const float*, const float b[3], bool front, float col[3], int) {
    const float* uvv = v + 9;
    float light[3];
    bool pos = true;
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) {
      light[ch] = interp(b, v[ch], v[3 + ch], v[6 + ch]);
      pos = pos && (light[ch] >= 0.0f);
    }
    const float uu = interp(b, uvv[0], uvv[2], uvv[4]);
    const float vv = interp(b, uvv[1], uvv[3], uvv[5]);
    const int32_t u = floor_mod(f2i_sat(floorf(uu)), sc.tex_w);
    const int32_t t = floor_mod(f2i_sat(floorf(vv)), sc.tex_h);
    float tex[3];
    tex_rgb(in.textures, sc.tex_base, sc.tex_h, sc.tex_c, u, t, tex);
    col[0] = tex[0] * light[0];
    col[1] = tex[1] * light[1];
    col[2] = tex[2] * light[2];
    return front && pos;
  }
};

// shaders/phong.py:74-164: per corner the eye-space normal and the uv.
template <>
struct Varyings<JR_SHADER_PHONG> {
  static constexpr int N = 15;
  static JR_DEV void corner(const jr_inputs& in, const jr_scene& sc, const Corner& c, int k,
                            float* v) {
    const SlotRows r = slot_rows(in, c.slot[k]);
    float n[3];
    load3(in.normals, r.nrm, n);
    load2(in.uvs, r.uv, v + 9 + 2 * k);
    eye_normal_of(n, sc, v + 3 * k);
  }
  static JR_DEV void corners(const jr_inputs& in, const jr_scene& sc, const Corner& c,
                             float v[N]) {
#pragma unroll
    for (int k = 0; k < 3; ++k) corner(in, sc, c, k, v);
  }
  static JR_DEV void scene_terms(const jr_scene& sc, float l[3]) {
    normalise3(sc.light_direction, l);
  }
  static JR_DEV bool pixel(const jr_inputs& in, const jr_scene& sc, const float v[N],
                           const float ln[3], const float b[3], bool front, float col[3],
                           int) {
    const float* ne = v;
    const float* uvv = v + 9;
    float n[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) n[q] = interp(b, ne[q], ne[3 + q], ne[6 + q]);
    const float uu = interp(b, uvv[0], uvv[2], uvv[4]);
    const float vv = interp(b, uvv[1], uvv[3], uvv[5]);
    const int32_t u = floor_mod(f2i_sat(floorf(uu)), sc.tex_w);
    const int32_t t = floor_mod(f2i_sat(floorf(vv)), sc.tex_h);
    float tex[3], nn[3];
    tex_rgb(in.textures, sc.tex_base, sc.tex_h, sc.tex_c, u, t, tex);
    normalise3(n, nn);
    const float d = dot3(nn, ln);
    float lc[3];
    bool pos = true;
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) {
      lc[ch] = sc.light_colour[ch] * d;
      pos = pos && (lc[ch] >= 0.0f);
    }
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) col[ch] = pos ? tex[ch] * lc[ch] : 0.0f;
    return front;
  }
};

// The shaders above, whole, for a pixel of the triangle with corners c.
template <int SHADER>
struct ShadeVaryings {
  static JR_DEV bool run(const jr_inputs& in, const jr_scene& sc, const Corner& c,
                         const float b[3], bool front, float col[3], int scene) {
    using V = Varyings<SHADER>;
    float v[V::N], l[3];
    V::corners(in, sc, c, v);
    V::scene_terms(sc, l);
    return V::pixel(in, sc, v, l, b, front, col, scene);
  }
};
template <>
struct Shade<JR_SHADER_GOURAUD> : ShadeVaryings<JR_SHADER_GOURAUD> {};
template <>
struct Shade<JR_SHADER_GOURAUD_TEXTURE> : ShadeVaryings<JR_SHADER_GOURAUD_TEXTURE> {};
template <>
struct Shade<JR_SHADER_PHONG> : ShadeVaryings<JR_SHADER_PHONG> {};

// shaders/phong_reflection.py:100-235: per corner the eye-space normal and
// uv; the texture index of corner 0.
template <>
struct Varyings<JR_SHADER_PHONG_REFLECTION> {
  static constexpr int N = 18;
  static constexpr int kTi = 15;     // texture index of corner 0
  static constexpr int kShape = 16;  // its texture_shape row (2 ints)
  // corner k's entries of v; corner 0 also the triangle's texture index and shape
  static JR_DEV void corner(const jr_inputs& in, const jr_scene& sc, const Corner& c, int k,
                            float* v) {
    const SlotRows r = slot_rows(in, c.slot[k]);
    float n[3];
    load3(in.normals, r.nrm, n);
    load2(in.uvs, r.uv, v + 9 + 2 * k);
    eye_normal_of(n, sc, v + 3 * k);
    if (k == 0) {
      const int32_t ti = in.tex_index[attr_row(in.tex_index_index, c.slot[0])];
      v[kTi] = __int_as_float(ti);
      texture_shape(in, sc, ti, v + kShape, v + kShape + 1);
    }
  }
  static JR_DEV void corners(const jr_inputs& in, const jr_scene& sc, const Corner& c,
                             float v[N]) {
#pragma unroll
    for (int k = 0; k < 3; ++k) corner(in, sc, c, k, v);
  }
  // per-scene terms of the pixel stage, computed once per tile
  static JR_DEV void scene_terms(const jr_scene& sc, float l[3]) {
    normalise3(sc.light_dir_eye, l);
  }
  static JR_DEV bool pixel(const jr_inputs& in, const jr_scene& sc, const float v[N],
                           const float l[3], const float b[3], bool front, float col[3],
                           int) {
    const float* ne = v;
    const float* uvv = v + 9;
    const int32_t ti = __float_as_int(v[kTi]);
    float n[3], uv[2];
#pragma unroll
    for (int q = 0; q < 3; ++q) n[q] = interp(b, ne[q], ne[3 + q], ne[6 + q]);
    uv[0] = interp(b, uvv[0], uvv[2], uvv[4]);
    uv[1] = interp(b, uvv[1], uvv[3], uvv[5]);
    float tex[3], diff, spec;
    reflection_terms(in, sc, n, uv, ti, __float_as_int(v[kShape]), __float_as_int(v[kShape + 1]), l,
                     tex, &diff, &spec);
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) {
      const float k = sc.diffuse[ch] * diff + sc.specular[ch] * spec;
      col[ch] = sc.ambient[ch] * tex[ch] + (k * sc.light_colour[ch]) * tex[ch];
    }
    return front;
  }
};

// shaders/phong_reflection_shadow.py:110-271: as above plus each corner's
// homogeneous-normalised shadow-camera clip position.
template <>
struct Varyings<JR_SHADER_PHONG_REFLECTION_SHADOW> {
  static constexpr int N = 30;
  static constexpr int kTi = 27;
  static constexpr int kShape = 28;
  static JR_DEV void corner(const jr_inputs& in, const jr_scene& sc, const Corner& c, int k,
                            float* v) {
    const SlotRows r = slot_rows(in, c.slot[k]);
    float n[3], p[3], h[4];
    load3(in.normals, r.nrm, n);
    load3(in.positions, r.pos, p);
    load2(in.uvs, r.uv, v + 9 + 2 * k);
    eye_normal_of(n, sc, v + 3 * k);
    apply4(sc.shadow_world_to_clip, p[0], p[1], p[2], 1.0f, h);
    // normalise_homogeneous: every component divided by w
    v[15 + 4 * k + 0] = h[0] / h[3];
    v[15 + 4 * k + 1] = h[1] / h[3];
    v[15 + 4 * k + 2] = h[2] / h[3];
    v[15 + 4 * k + 3] = h[3] / h[3];
    if (k == 0) {
      const int32_t ti = in.tex_index[attr_row(in.tex_index_index, c.slot[0])];
      v[kTi] = __int_as_float(ti);
      texture_shape(in, sc, ti, v + kShape, v + kShape + 1);
    }
  }
  static JR_DEV void corners(const jr_inputs& in, const jr_scene& sc, const Corner& c,
                             float v[N]) {
#pragma unroll
    for (int k = 0; k < 3; ++k) corner(in, sc, c, k, v);
  }
  static JR_DEV void scene_terms(const jr_scene& sc, float l[3]) {
    normalise3(sc.light_dir_eye, l);
  }
  template <class ShadowDepth = MapShadowDepth>
  static JR_DEV bool pixel(const jr_inputs& in, const jr_scene& sc, const float v[N],
                           const float l[3], const float b[3], bool front, float col[3],
                           int scene, const ShadowDepth& shadow_depth = ShadowDepth()) {
    const float* ne = v;
    const float* uvv = v + 9;
    const float* scd = v + 15;
    const int32_t ti = __float_as_int(v[kTi]);
    float n[3], uv[2], s4[4];
#pragma unroll
    for (int q = 0; q < 3; ++q) n[q] = interp(b, ne[q], ne[3 + q], ne[6 + q]);
    uv[0] = interp(b, uvv[0], uvv[2], uvv[4]);
    uv[1] = interp(b, uvv[1], uvv[3], uvv[5]);
#pragma unroll
    for (int q = 0; q < 4; ++q) s4[q] = interp(b, scd[q], scd[4 + q], scd[8 + q]);
    float tex[3], diff, spec;
    reflection_terms(in, sc, n, uv, ti, __float_as_int(v[kShape]), __float_as_int(v[kShape + 1]), l,
                     tex, &diff, &spec);
    float k[3];
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) k[ch] = sc.diffuse[ch] * diff + sc.specular[ch] * spec;
    // The shadow test only scales k: where all three are zero (a surface
    // facing away from the light that reflects none of it towards the eye)
    // shadow * k is a zero either way and the colour does not depend on the
    // test, so it is skipped unless its depth must be computed anyway.
    bool lit = true;
    if (shadow_depth.required() || k[0] != 0.0f || k[1] != 0.0f || k[2] != 0.0f) {
      // shadow.camera.viewport @ shadow_coord, then normalise_homogeneous
      float sv[4];
      apply4(sc.shadow_viewport, s4[0], s4[1], s4[2], s4[3], sv);
      const float sx = sv[0] / sv[3], sy = sv[1] / sv[3], sz = sv[2] / sv[3];
      const float sd = shadow_depth(in, scene, sx, sy);
      lit = sz <= sd;
    }
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) {
      const float shadow = lit ? 1.0f : 1.0f - sc.shadow_strength[ch];
      col[ch] = sc.ambient[ch] * tex[ch] + ((shadow * k[ch]) * tex[ch]) * sc.light_colour[ch];
    }
    return front;
  }
};

template <>
struct Shade<JR_SHADER_PHONG_REFLECTION> {
  static JR_DEV bool run(const jr_inputs& in, const jr_scene& sc, const Corner& c,
                         const float b[3], bool front, float col[3], int scene) {
    using V = Varyings<JR_SHADER_PHONG_REFLECTION>;
    float v[V::N], l[3];
    V::corners(in, sc, c, v);
    V::scene_terms(sc, l);
    return V::pixel(in, sc, v, l, b, front, col, scene);
  }
};

template <>
struct Shade<JR_SHADER_PHONG_REFLECTION_SHADOW> {
  static JR_DEV bool run(const jr_inputs& in, const jr_scene& sc, const Corner& c,
                         const float b[3], bool front, float col[3], int scene) {
    using V = Varyings<JR_SHADER_PHONG_REFLECTION_SHADOW>;
    float v[V::N], l[3];
    V::corners(in, sc, c, v);
    V::scene_terms(sc, l);
    return V::pixel(in, sc, v, l, b, front, col, scene);
  }
};

}  // namespace jr
