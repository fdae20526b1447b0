/*
 * jr_oracle.c -- CPU oracle of the rasterisation path.  TEST INFRASTRUCTURE
 * ONLY: tests/, the package's smoke check and bench.py's cpu_baseline leg use
 * it as the checker; the product path (libjrhip.so) never calls it.
 *
 * A brute-force restatement of jaxrenderer's `pipeline.render`
 * (renderer/pipeline.py:470-537), structured like the reference: the vertex
 * stage per slot, PerPrimitive.create per triangle, then for every pixel every
 * primitive is tested (`_per_primitive_preprocess`, renderer/pipeline.py:163-279),
 * the default primitive_chooser takes the argmin (renderer/shader.py:207-217),
 * the chosen fragment is interpolated and shaded
 * (renderer/pipeline.py:283-328), mixed (renderer/shader.py:374-396) and merged
 * (renderer/pipeline.py:401-440).  No binning, no tiling.
 *
 * Float32 arithmetic, one IEEE operation at a time (-ffp-contract=off), in the
 * reference's formula order.  Where the reference's own order is decided by
 * XLA or LAPACK (dot/reduce association, 3x3 inverse) the order is the one the
 * build documents: dots and sums left to right, LU with partial pivoting in
 * LAPACK sgetrf2 order, triangular solves in reference-BLAS strsm order.
 *
 * Interface: the jr_frame / jr_inputs / jr_scene / jr_outputs structs of
 * include/jrhip.h, with host pointers.
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/jrhip.h"

/* ---- small linear algebra in the reference's order ---------------------- */

/* Camera.apply, renderer/geometry.py:280-315: out_i = sum_j p_j M[i][j] */
static void o_apply4(const float *M, const float p[4], float out[4]) {
  for (int i = 0; i < 4; ++i) {
    float s = p[0] * M[4 * i + 0];
    s = s + p[1] * M[4 * i + 1];
    s = s + p[2] * M[4 * i + 2];
    s = s + p[3] * M[4 * i + 3];
    out[i] = s;
  }
}

static float o_dot3(const float a[3], const float b[3]) {
  float s = a[0] * b[0];
  s = s + a[1] * b[1];
  s = s + a[2] * b[2];
  return s;
}

/* geometry.normalise, renderer/geometry.py:36-47 */
static void o_normalise3(const float v[3], float out[3]) {
  float ss = v[0] * v[0];
  ss = ss + v[1] * v[1];
  ss = ss + v[2] * v[2];
  float n = sqrtf(ss);
  out[0] = v[0] / n;
  out[1] = v[1] / n;
  out[2] = v[2] / n;
}

/* Camera.apply_vec, renderer/geometry.py:346-378 */
static void o_apply_vec(const float *M, const float v[3], float out[3]) {
  float n[3], h[4], p[4], t[3];
  o_normalise3(v, n);
  p[0] = n[0]; p[1] = n[1]; p[2] = n[2]; p[3] = 0.0f;
  o_apply4(M, p, h);
  t[0] = h[0]; t[1] = h[1]; t[2] = h[2];
  o_normalise3(t, out);
}

/* jnp.linalg.det 3x3 (renderer/pipeline.py:93), JAX's _det_3x3 formula */
static float o_det3(const float a[3][3]) {
  float d = (a[0][0] * a[1][1]) * a[2][2];
  d = d + (a[0][1] * a[1][2]) * a[2][0];
  d = d + (a[0][2] * a[1][0]) * a[2][1];
  d = d - (a[0][2] * a[1][1]) * a[2][0];
  d = d - (a[0][0] * a[1][2]) * a[2][1];
  d = d - (a[0][1] * a[1][0]) * a[2][2];
  return d;
}

/* jnp.linalg.inv (renderer/pipeline.py:105): getrf (partial pivoting) then
 * solve(A, I) with the permuted identity, unit-lower then upper strsm. */
static void o_inv3(const float A[3][3], float X[3][3]) {
  float a[3][3];
  int perm[3] = {0, 1, 2};
  memcpy(a, A, sizeof(a));
  for (int j = 0; j < 3; ++j) {
    int p = j;
    float amax = fabsf(a[j][j]);
    for (int i = j + 1; i < 3; ++i)
      if (fabsf(a[i][j]) > amax) { amax = fabsf(a[i][j]); p = i; }
    if (a[p][j] != 0.0f) {
      if (p != j) {
        for (int k = 0; k < 3; ++k) { float t = a[p][k]; a[p][k] = a[j][k]; a[j][k] = t; }
        int tp = perm[p]; perm[p] = perm[j]; perm[j] = tp;
      }
      if (fabsf(a[j][j]) >= FLT_MIN) {
        float r = 1.0f / a[j][j];
        for (int i = j + 1; i < 3; ++i) a[i][j] = a[i][j] * r;
      } else {
        for (int i = j + 1; i < 3; ++i) a[i][j] = a[i][j] / a[j][j];
      }
    }
    for (int i = j + 1; i < 3; ++i)
      for (int k = j + 1; k < 3; ++k) a[i][k] = a[i][k] - a[i][j] * a[j][k];
  }
  float x[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) x[r][c] = (perm[r] == c) ? 1.0f : 0.0f;
  for (int c = 0; c < 3; ++c) {
    for (int k = 0; k < 3; ++k)
      if (x[k][c] != 0.0f)
        for (int i = k + 1; i < 3; ++i) x[i][c] = x[i][c] - x[k][c] * a[i][k];
    for (int k = 2; k >= 0; --k)
      if (x[k][c] != 0.0f) {
        x[k][c] = x[k][c] / a[k][k];
        for (int i = 0; i < k; ++i) x[i][c] = x[i][c] - x[k][c] * a[i][k];
      }
  }
  memcpy(X, x, sizeof(x));
}

/* ---- JAX scalar semantics ------------------------------------------------ */

/* astype(int32) of a float: NaN -> 0, saturating */
static int32_t o_f2i(float f) {
  if (isnan(f)) return 0;
  if (f >= 2147483648.0f) return INT32_MAX;
  if (f <= -2147483648.0f) return INT32_MIN;
  return (int32_t)f;
}

/* x[i] read: negative counts from the end, then clamp (JAX gather default) */
static int32_t o_idx(int32_t i, int32_t n) {
  if (i < 0) i += n;
  if (i < 0) i = 0;
  if (i > n - 1) i = n - 1;
  return i;
}

/* jnp integer remainder, positive divisor */
static int32_t o_mod(int32_t a, int32_t n) {
  int32_t r = a % n;
  if (r != 0 && ((r < 0) != (n < 0))) r += n;
  return r;
}

/* lax.max / jnp.maximum propagate NaN */
static float o_max(float a, float b) {
  if (isnan(a)) return a;
  if (isnan(b)) return b;
  return a > b ? a : b;
}

/* jnp.modf(x)[0] */
static float o_frac(float x) {
  float whole = (x >= 0.0f) ? floorf(x) : ceilf(x);
  return x - whole;
}

static float o_interp(const float b[3], float v0, float v1, float v2) {
  float s = b[0] * v0;
  s = s + b[1] * v1;
  s = s + b[2] * v2;
  return s;
}

/* ---- vertex stage (one slot = one gl_VertexID) --------------------------- */

static int32_t row_of(const int32_t *index, int32_t slot) {
  return index ? index[slot] : slot;
}

typedef struct {
  float clip[4];     /* gl_Position */
  float colour[3];   /* gouraud / gouraud_texture light colour */
  float normal[3];   /* eye-space normal (phong family) */
  float uv[2];
  int32_t tex_index; /* phong_reflection* */
  float shadow[4];   /* shadow_coord */
} vert_out;

static void o_vertex(const jr_inputs *in, const jr_scene *sc, int shader,
                     int32_t slot, vert_out *o) {
  if (in->clip) {
    for (int q = 0; q < 4; ++q) o->clip[q] = in->clip[4 * (int64_t)slot + q];
    return;
  }
  const float *p = in->positions + 3 * (int64_t)row_of(in->pos_index, slot);
  float ph[4] = {p[0], p[1], p[2], 1.0f};
  o_apply4(sc->world_to_clip, ph, o->clip);
  if (shader == JR_SHADER_DEPTH || shader == JR_SHADER_VISIBILITY) return;
  const float *n = in->normals + 3 * (int64_t)row_of(in->nrm_index, slot);
  float nn[3];
  o_normalise3(n, nn);
  if (shader == JR_SHADER_GOURAUD || shader == JR_SHADER_GOURAUD_TEXTURE) {
    /* shaders/gouraud.py:66-86, gouraud_texture.py:76-99 */
    float ld[3];
    o_normalise3(sc->light_direction, ld);
    float intensity = o_dot3(nn, ld);
    if (shader == JR_SHADER_GOURAUD) {
      const float *c = in->colours + 3 * (int64_t)row_of(in->col_index, slot);
      for (int q = 0; q < 3; ++q) o->colour[q] = (c[q] * sc->light_colour[q]) * intensity;
    } else {
      for (int q = 0; q < 3; ++q) o->colour[q] = sc->light_colour[q] * intensity;
    }
  } else {
    /* phong family: Camera.apply_vec(normalise(n), world_to_eye_norm) */
    o_apply_vec(sc->world_to_eye_norm, nn, o->normal);
  }
  if (shader != JR_SHADER_GOURAUD) {
    const float *uv = in->uvs + 2 * (int64_t)row_of(in->uv_index, slot);
    o->uv[0] = uv[0];
    o->uv[1] = uv[1];
  }
  if (shader == JR_SHADER_PHONG_REFLECTION || shader == JR_SHADER_PHONG_REFLECTION_SHADOW)
    o->tex_index = in->tex_index[row_of(in->tex_index_index, slot)];
  if (shader == JR_SHADER_PHONG_REFLECTION_SHADOW) {
    /* normalise_homogeneous(shadow.camera.to_clip(position)) */
    float h[4];
    o_apply4(sc->shadow_world_to_clip, ph, h);
    for (int q = 0; q < 4; ++q) o->shadow[q] = h[q] / h[3];
  }
}

/* ---- fragment stage ------------------------------------------------------- */

static void tex3(const jr_inputs *in, const jr_scene *sc, int32_t u, int32_t v, float out[3]) {
  const float *t = in->textures + sc->tex_base + ((int64_t)u * sc->tex_h + v) * sc->tex_c;
  out[0] = t[0];
  out[1] = t[sc->tex_c > 1 ? 1 : 0];
  out[2] = t[sc->tex_c > 1 ? 2 : 0];
}

/* phong_reflection*.fragment texture / diffuse / specular terms
 * (shaders/phong_reflection.py:166-220; renderer/model.py:303-339) */
static void o_reflection(const jr_inputs *in, const jr_scene *sc, const float n_i[3],
                         const float uv[2], int32_t ti, float tex[3], float *diff,
                         float *spec) {
  int32_t row = o_idx(ti, sc->texshape_count);
  int32_t s0 = in->texture_shapes[2 * (int64_t)(sc->texshape_base + row)];
  int32_t s1 = in->texture_shapes[2 * (int64_t)(sc->texshape_base + row) + 1];
  float f0 = o_frac(uv[0]), f1 = o_frac(uv[1]);
  if (f0 < 0.0f) f0 = f0 + 1.0f;
  if (f1 < 0.0f) f1 = f1 + 1.0f;
  int32_t shift = (int32_t)((uint32_t)ti * (uint32_t)sc->tex_offset);
  float r0 = f0 * (float)s0 + (float)shift;
  float r1 = f1 * (float)s1;
  int32_t u0 = o_f2i(floorf(r0)), v0 = o_f2i(floorf(r1));
  tex3(in, sc, o_idx(u0, sc->tex_w), o_idx(v0, sc->tex_h), tex);
  float se = in->specular_maps[sc->spec_base + (int64_t)o_idx(u0, sc->spec_w) * sc->spec_h +
                               o_idx(v0, sc->spec_h)];
  float n[3], l[3], r[3], rn[3];
  o_normalise3(n_i, n);
  o_normalise3(sc->light_dir_eye, l);
  float d = o_dot3(n, l);
  *diff = o_max(d, 0.0f);
  float two_d = 2.0f * d;
  for (int q = 0; q < 3; ++q) r[q] = two_d * n[q] - l[q];
  o_normalise3(r, rn);
  *spec = powf(o_max(rn[2], 0.0f), se);
}

/* Runs Shader.interpolate + Shader.fragment for the chosen primitive.
 * Returns PerFragment.keeps; writes the mixer colour. */
static int o_fragment(const jr_inputs *in, const jr_scene *sc, int shader, int scene,
                      const int32_t slot[3], const float b[3], int front, float col[3]) {
  vert_out v[3];
  col[0] = col[1] = col[2] = 0.0f;
  if (shader == JR_SHADER_DEPTH || shader == JR_SHADER_VISIBILITY) return 1;
  for (int k = 0; k < 3; ++k) o_vertex(in, sc, shader, slot[k], &v[k]);
  if (shader == JR_SHADER_GOURAUD) { /* shaders/gouraud.py:90-124 */
    int pos = 1;
    for (int q = 0; q < 3; ++q) {
      col[q] = o_interp(b, v[0].colour[q], v[1].colour[q], v[2].colour[q]);
      pos = pos && (col[q] >= 0.0f);
    }
    return front && pos;
  }
  float uv[2];
  uv[0] = o_interp(b, v[0].uv[0], v[1].uv[0], v[2].uv[0]);
  uv[1] = o_interp(b, v[0].uv[1], v[1].uv[1], v[2].uv[1]);
  if (shader == JR_SHADER_GOURAUD_TEXTURE) { /* shaders/gouraud_texture.py:103-144 */
    float lc[3], tex[3];
    int pos = 1;
    for (int q = 0; q < 3; ++q) {
      lc[q] = o_interp(b, v[0].colour[q], v[1].colour[q], v[2].colour[q]);
      pos = pos && (lc[q] >= 0.0f);
    }
    int32_t u = o_mod(o_f2i(floorf(uv[0])), sc->tex_w);
    int32_t w = o_mod(o_f2i(floorf(uv[1])), sc->tex_h);
    tex3(in, sc, u, w, tex);
    for (int q = 0; q < 3; ++q) col[q] = tex[q] * lc[q];
    return front && pos;
  }
  float n_i[3];
  for (int q = 0; q < 3; ++q) n_i[q] = o_interp(b, v[0].normal[q], v[1].normal[q], v[2].normal[q]);
  if (shader == JR_SHADER_PHONG || shader == JR_SHADER_PHONG_DARBOUX) {
    int32_t u = o_mod(o_f2i(floorf(uv[0])), sc->tex_w);
    int32_t w = o_mod(o_f2i(floorf(uv[1])), sc->tex_h);
    float n[3], ld[3], tex[3], lc[3];
    o_normalise3(n_i, n);
    if (shader == JR_SHADER_PHONG_DARBOUX) { /* shaders/phong_darboux.py:208-293 */
      int32_t face = in->id_to_face[slot[0]];
      float tri[3][3], tuv[3][2];
      for (int m = 0; m < 3; ++m) {
        int32_t q = in->face_slots[3 * (int64_t)face + m];
        const float *p = in->positions + 3 * (int64_t)row_of(in->pos_index, q);
        float ph[4] = {p[0], p[1], p[2], 1.0f}, h[4];
        o_apply4(sc->world_to_clip, ph, h);
        for (int c = 0; c < 3; ++c) tri[m][c] = (h[3] == 0.0f) ? h[c] : h[c] / h[3];
        const float *t = in->uvs + 2 * (int64_t)row_of(in->uv_index, q);
        tuv[m][0] = t[0];
        tuv[m][1] = t[1];
      }
      float A[3][3], AI[3][3];
      for (int c = 0; c < 3; ++c) {
        A[0][c] = tri[1][c] - tri[0][c];
        A[1][c] = tri[2][c] - tri[0][c];
        A[2][c] = n[c];
      }
      o_inv3(A, AI);
      float du1 = tuv[1][0] - tuv[0][0], du2 = tuv[2][0] - tuv[0][0];
      float dv1 = tuv[1][1] - tuv[0][1], dv2 = tuv[2][1] - tuv[0][1];
      float iv[3], jv[3], ni[3], nj[3], bn[3], n2[3];
      for (int r = 0; r < 3; ++r) {
        float s = AI[r][0] * du1;
        s = s + AI[r][1] * du2;
        s = s + AI[r][2] * 0.0f;
        iv[r] = s;
        float t = AI[r][0] * dv1;
        t = t + AI[r][1] * dv2;
        t = t + AI[r][2] * 0.0f;
        jv[r] = t;
      }
      o_normalise3(iv, ni);
      o_normalise3(jv, nj);
      const float *nm = in->normal_maps + sc->nmap_base +
                        3 * ((int64_t)o_idx(u, sc->nmap_w) * sc->nmap_h + o_idx(w, sc->nmap_h));
      for (int r = 0; r < 3; ++r) {
        float s = ni[r] * nm[0];
        s = s + nj[r] * nm[1];
        s = s + n[r] * nm[2];
        bn[r] = s;
      }
      o_normalise3(bn, n2);
      memcpy(n, n2, sizeof(n));
    }
    o_normalise3(sc->light_direction, ld);
    float d = o_dot3(n, ld);
    int pos = 1;
    for (int q = 0; q < 3; ++q) {
      lc[q] = sc->light_colour[q] * d;
      pos = pos && (lc[q] >= 0.0f);
    }
    tex3(in, sc, u, w, tex);
    for (int q = 0; q < 3; ++q) col[q] = pos ? tex[q] * lc[q] : 0.0f;
    return front;
  }
  /* phong_reflection / phong_reflection_shadow: texture index of vertex 0 */
  float tex[3], diff, spec;
  o_reflection(in, sc, n_i, uv, v[0].tex_index, tex, &diff, &spec);
  if (shader == JR_SHADER_PHONG_REFLECTION) {
    for (int q = 0; q < 3; ++q) {
      float k = sc->diffuse[q] * diff + sc->specular[q] * spec;
      col[q] = sc->ambient[q] * tex[q] + (k * sc->light_colour[q]) * tex[q];
    }
    return front;
  }
  /* shadow lookup, shaders/phong_reflection_shadow.py:196-209 and
   * Shadow.get, renderer/shadow.py:127-153 */
  float s4[4], sv[4];
  for (int q = 0; q < 4; ++q) s4[q] = o_interp(b, v[0].shadow[q], v[1].shadow[q], v[2].shadow[q]);
  o_apply4(sc->shadow_viewport, s4, sv);
  float sx = sv[0] / sv[3], sy = sv[1] / sv[3], sz = sv[2] / sv[3];
  int32_t ix = o_f2i(roundf(sx)), iy = o_f2i(roundf(sy));
  if (ix < 0) ix += in->shadow_w;
  if (iy < 0) iy += in->shadow_h;
  float sd = INFINITY;
  if (ix >= 0 && ix < in->shadow_w && iy >= 0 && iy < in->shadow_h)
    sd = in->shadow_maps[(int64_t)scene * in->shadow_w * in->shadow_h +
                         (int64_t)ix * in->shadow_h + iy];
  int lit = sz <= sd;
  for (int q = 0; q < 3; ++q) {
    float shadow = lit ? 1.0f : 1.0f - sc->shadow_strength[q];
    float k = sc->diffuse[q] * diff + sc->specular[q] * spec;
    col[q] = sc->ambient[q] * tex[q] + ((shadow * k) * tex[q]) * sc->light_colour[q];
  }
  return front;
}

/* ---- the pipeline --------------------------------------------------------- */

typedef struct {
  float minv[3][3];
  float z[3];
  float det;
  int keep;
  int32_t slot[3];
} prim;

/* jr_forward_cpu, also counting per pixel the primitives that pass
 * keep & inside & front-facing (the depth complexity the chooser resolves)
 * into complexity[scene][width][height] when it is not NULL. */
int32_t jr_forward_cpu_counted(const jr_frame *fr, const jr_inputs *in, const jr_scene *scenes,
                               jr_outputs *out, int32_t *complexity) {
  if (!fr || !in || !scenes || !out || !out->zbuffer) return JR_E_INVALID;
  const int W = fr->width, H = fr->height, shader = fr->shader;
  for (int s = 0; s < fr->n_scenes; ++s) {
    const jr_scene *sc = &scenes[s];
    const int32_t F = sc->tri_end - sc->tri_begin;
    prim *P = (prim *)malloc(sizeof(prim) * (size_t)(F > 0 ? F : 1));
    /* vertex stage + PerPrimitive.create (renderer/pipeline.py:76-113) */
    for (int32_t f = 0; f < F; ++f) {
      const int32_t t = sc->tri_begin + f;
      float m[3][3];
      for (int k = 0; k < 3; ++k) {
        vert_out v;
        P[f].slot[k] = in->faces[3 * (int64_t)t + k];
        o_vertex(in, sc, JR_SHADER_DEPTH, P[f].slot[k], &v);
        m[k][0] = v.clip[0];
        m[k][1] = v.clip[1];
        m[k][2] = v.clip[3];
        P[f].z[k] = v.clip[2];
      }
      P[f].det = o_det3(m);
      P[f].keep = fabsf(P[f].det) > 1e-6f;
      o_inv3(m, P[f].minv);
    }
    const float *vp = sc->viewport;
#pragma omp parallel for schedule(dynamic, 1)
    for (int i = 0; i < W; ++i) {
      for (int j = 0; j < H; ++j) {
        const int64_t pix = (int64_t)s * W * H + (int64_t)i * H + j;
        /* _per_primitive_preprocess for every primitive, then the default
         * primitive_chooser: argmin (first minimum, first NaN) of
         * where(keeps & front, z, inf) */
        const float x = ((float)i - vp[3]) / vp[0];
        const float y = ((float)j - vp[7]) / vp[5];
        int32_t idx = 0;
        int32_t n_cover = 0;
        float best = 0.0f;
        for (int32_t f = 0; f < F; ++f) {
          float e[3];
          for (int k = 0; k < 3; ++k) {
            float v = x * P[f].minv[0][k];
            v = v + y * P[f].minv[1][k];
            v = v + P[f].minv[2][k];
            e[k] = v;
          }
          int inside = e[0] >= 0.0f && e[1] >= 0.0f && e[2] >= 0.0f;
          float z = e[0] * P[f].z[0];
          z = z + e[1] * P[f].z[1];
          z = z + e[2] * P[f].z[2];
          z = z * vp[10] + vp[11];
          int front = P[f].det >= 0.0f;
          n_cover += P[f].keep && inside && front;
          float depth = (P[f].keep && inside && front) ? z : INFINITY;
          if (f == 0) {
            best = depth;
            idx = 0;
          } else if (!isnan(best) && (isnan(depth) || depth < best)) {
            best = depth;
            idx = f;
          }
        }
        if (complexity) complexity[pix] = n_cover;
        int keep = 0;
        float z = 0.0f, b[3] = {0.0f, 0.0f, 0.0f}, col[3] = {0.0f, 0.0f, 0.0f};
        if (F > 0) {
          const prim *p = &P[idx];
          float e[3];
          for (int k = 0; k < 3; ++k) {
            float v = x * p->minv[0][k];
            v = v + y * p->minv[1][k];
            v = v + p->minv[2][k];
            e[k] = v;
          }
          int inside = e[0] >= 0.0f && e[1] >= 0.0f && e[2] >= 0.0f;
          float wr = e[0] + e[1];
          wr = wr + e[2];
          z = e[0] * p->z[0];
          z = z + e[1] * p->z[1];
          z = z + e[2] * p->z[2];
          z = z * vp[10] + vp[11];
          for (int k = 0; k < 3; ++k) b[k] = e[k] / wr;
          int keeps = p->keep && inside;
          int fk = o_fragment(in, sc, shader, s, p->slot, b, p->det >= 0.0f, col);
          keep = keeps && fk; /* _interpolate_and_fragment_shading + mix */
        }
        if (keep) {
          out->zbuffer[pix] = fr->apply_offset ? z + fr->depth_offset : z;
          if (out->canvas)
            for (int q = 0; q < 3; ++q) out->canvas[3 * pix + q] = col[q];
        } else if (fr->init_buffers) {
          out->zbuffer[pix] = fr->apply_offset ? fr->zbuffer_default + fr->depth_offset
                                               : fr->zbuffer_default;
          if (out->canvas)
            for (int q = 0; q < 3; ++q) out->canvas[3 * pix + q] = fr->colour_default[q];
        } else if (fr->apply_offset) {
          out->zbuffer[pix] = out->zbuffer[pix] + fr->depth_offset;
        }
        if (out->display && out->canvas) {
          /* transpose_for_display((clamp(0, canvas, 1) * 255).astype(uint8)),
           * renderer/utils.py:72-96 and notebooks/84x84/A100.ipynb cell 7 */
          uint8_t *d = out->display + 3 * (((int64_t)s * H + (H - 1 - j)) * W + i);
          for (int q = 0; q < 3; ++q) {
            float v = out->canvas[3 * pix + q];
            if (isnan(v)) {
              d[q] = 0;
            } else {
              v = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
              d[q] = (uint8_t)(int32_t)(v * 255.0f);
            }
          }
        }
        if (out->prim_id) out->prim_id[pix] = keep ? sc->tri_begin + idx : -1;
        if (out->bary)
          for (int q = 0; q < 3; ++q) out->bary[3 * pix + q] = keep ? b[q] : 0.0f;
      }
    }
    free(P);
  }
  return JR_OK;
}

int32_t jr_forward_cpu(const jr_frame *fr, const jr_inputs *in, const jr_scene *scenes,
                       jr_outputs *out) {
  return jr_forward_cpu_counted(fr, in, scenes, out, NULL);
}

/* Inverse of one 3x3 matrix (row-major), exported for the oracle's own tests. */
void jr_oracle_inv3(const float *a, float *x) {
  float A[3][3], X[3][3];
  memcpy(A, a, sizeof(A));
  o_inv3(A, X);
  memcpy(x, X, sizeof(X));
}

float jr_oracle_det3(const float *a) {
  float A[3][3];
  memcpy(A, a, sizeof(A));
  return o_det3(A);
}
