// jr_math.h -- scalar float32 math of the rasterisation path, in the order the
// reference evaluates it.  Device-only (included by the .hip sources); the CPU
// oracle restates the same formulas separately in oracle/jr_oracle.c.
//
// Everything here must be compiled with -ffp-contract=off and without
// fast-math: the primitive-id and depth buffers are compared bit-for-bit
// against the oracle.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define JR_DEV __device__ __forceinline__

namespace jr {

// ---- canonical small linear algebra -------------------------------------

// Camera.apply (renderer/geometry.py:280-315): out_i = sum_j p_j * M[i][j],
// summed left to right.
JR_DEV void apply4(const float* __restrict__ M, float p0, float p1, float p2,
                   float p3, float out[4]) {
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    float s = p0 * M[4 * i + 0];
    s = s + p1 * M[4 * i + 1];
    s = s + p2 * M[4 * i + 2];
    s = s + p3 * M[4 * i + 3];
    out[i] = s;
  }
}

// jnp.dot of two 3-vectors, left to right.
JR_DEV float dot3(const float a[3], const float b[3]) {
  float s = a[0] * b[0];
  s = s + a[1] * b[1];
  s = s + a[2] * b[2];
  return s;
}

// geometry.normalise (renderer/geometry.py:36-47): v / ||v||_2 with
// ||v|| = sqrt(sum v_i^2), each component divided.
JR_DEV void normalise3(const float v[3], float out[3]) {
  float ss = v[0] * v[0];
  ss = ss + v[1] * v[1];
  ss = ss + v[2] * v[2];
  const float n = sqrtf(ss);
  out[0] = v[0] / n;
  out[1] = v[1] / n;
  out[2] = v[2] / n;
}

// Camera.apply_vec (renderer/geometry.py:346-378): normalise, w = 0, apply,
// drop w, normalise.
JR_DEV void apply_vec(const float* __restrict__ M, const float v[3],
                      float out[3]) {
  float n[3];
  normalise3(v, n);
  float h[4];
  apply4(M, n[0], n[1], n[2], 0.0f, h);
  const float t[3] = {h[0], h[1], h[2]};
  normalise3(t, out);
}

// jnp.linalg.det for 3x3: JAX's closed form _det_3x3 evaluated left to right
// (renderer/pipeline.py:93).
JR_DEV float det3(const float a[9]) {
  float d = (a[0] * a[4]) * a[8];
  d = d + (a[1] * a[5]) * a[6];
  d = d + (a[2] * a[3]) * a[7];
  d = d - (a[2] * a[4]) * a[6];
  d = d - (a[0] * a[5]) * a[7];
  d = d - (a[1] * a[3]) * a[8];
  return d;
}

// jnp.linalg.inv (renderer/pipeline.py:105; renderer/model.py:526 for the 4x4
// object transforms): LU with partial pivoting in the operation order of
// LAPACK sgetrf2 (first maximal |pivot|, reciprocal scaling when |pivot| >=
// FLT_MIN), then the permuted identity solved with the reference-BLAS strsm
// loops (lower unit, then upper), skipping zero entries.  A and X are N x N
// row-major.
template <int N>
JR_DEV void inv_lu(const float* A, float* X) {
  // every index below is a compile-time constant (the pivot row is chosen by
  // conditional swaps with each candidate row), so the matrix stays in
  // registers
  float a[N * N];
#pragma unroll
  for (int i = 0; i < N * N; ++i) a[i] = A[i];
  int perm[N];
#pragma unroll
  for (int i = 0; i < N; ++i) perm[i] = i;
#pragma unroll
  for (int j = 0; j < N; ++j) {
    int p = j;
    float amax = fabsf(a[N * j + j]);
    float apj = a[N * j + j];
#pragma unroll
    for (int i = j + 1; i < N; ++i) {
      const float v = fabsf(a[N * i + j]);
      if (v > amax) {
        amax = v;
        p = i;
        apj = a[N * i + j];
      }
    }
    const bool nz = apj != 0.0f;
#pragma unroll
    for (int i = j + 1; i < N; ++i) {
      const bool sw = nz && p == i;  // at most one candidate row swaps with row j
#pragma unroll
      for (int k = 0; k < N; ++k) {
        const float t = a[N * i + k], u = a[N * j + k];
        a[N * i + k] = sw ? u : t;
        a[N * j + k] = sw ? t : u;
      }
      const int tp = perm[i], up = perm[j];
      perm[i] = sw ? up : tp;
      perm[j] = sw ? tp : up;
    }
    if (nz) {
      const float piv = a[N * j + j];
      if (fabsf(piv) >= 1.17549435e-38f) {
        const float r = 1.0f / piv;
#pragma unroll
        for (int i = j + 1; i < N; ++i) a[N * i + j] = a[N * i + j] * r;
      } else {
#pragma unroll
        for (int i = j + 1; i < N; ++i) a[N * i + j] = a[N * i + j] / piv;
      }
    }
#pragma unroll
    for (int i = j + 1; i < N; ++i)
#pragma unroll
      for (int k = j + 1; k < N; ++k)
        a[N * i + k] = a[N * i + k] - a[N * i + j] * a[N * j + k];
  }
  float x[N * N];
#pragma unroll
  for (int r = 0; r < N; ++r)
#pragma unroll
    for (int c = 0; c < N; ++c) x[N * r + c] = (perm[r] == c) ? 1.0f : 0.0f;
#pragma unroll
  for (int c = 0; c < N; ++c) {
#pragma unroll
    for (int k = 0; k < N; ++k) {
      if (x[N * k + c] != 0.0f) {
#pragma unroll
        for (int i = k + 1; i < N; ++i)
          x[N * i + c] = x[N * i + c] - x[N * k + c] * a[N * i + k];
      }
    }
#pragma unroll
    for (int k = N - 1; k >= 0; --k) {
      if (x[N * k + c] != 0.0f) {
        x[N * k + c] = x[N * k + c] / a[N * k + k];
#pragma unroll
        for (int i = 0; i < k; ++i)
          x[N * i + c] = x[N * i + c] - x[N * k + c] * a[N * i + k];
      }
    }
  }
#pragma unroll
  for (int i = 0; i < N * N; ++i) X[i] = x[i];
}

JR_DEV void inv3(const float A[9], float X[9]) { inv_lu<3>(A, X); }

// ---- scalar conversions with JAX semantics --------------------------------

// float32 -> int32 as XLA converts: NaN -> 0, saturate out-of-range values.
JR_DEV int32_t f2i_sat(float f) {
  if (!(f == f)) return 0;
  if (f >= 2147483648.0f) return 2147483647;
  if (f <= -2147483648.0f) return (int32_t)(-2147483647 - 1);
  return (int32_t)f;
}

// (clamp(0, v, 1) * 255).astype(uint8): lax.clamp keeps NaN, the product
// stays NaN and XLA's float -> integer conversion maps NaN to 0; in range the
// conversion truncates toward zero.
JR_DEV uint8_t display_u8(float v) {
  if (!(v == v)) return 0;
  const float c = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
  return (uint8_t)(int32_t)(c * 255.0f);
}

// NumPy/JAX integer index: negative values count from the end, then the
// gather clamps to [0, n-1] (JAX's default out-of-bounds mode for reads).
JR_DEV int32_t index_clamp(int32_t i, int32_t n) {
  if (i < 0) i += n;
  if (i < 0) i = 0;
  if (i > n - 1) i = n - 1;
  return i;
}

// Python/jnp integer `%` with a positive divisor: result in [0, n).
JR_DEV int32_t floor_mod(int32_t a, int32_t n) {
  int32_t r = a % n;
  if (r != 0 && ((r < 0) != (n < 0))) r += n;
  return r;
}

// lax.max / jnp.maximum: NaN propagates.
JR_DEV float max_nan(float a, float b) {
  if (a != a) return a;
  if (b != b) return b;
  return a > b ? a : b;
}

// jnp.power(x, y) of the specular term, x = max(cos, 0) in [+0, 1] or NaN.
// For y > 0 and a normal x it is 2^(y log2 x) from the hardware base-2
// logarithm and exponential (each within about 1 ulp) whenever
// t = y log2 x <= 0.5: the result is then at most 1.42 and its relative
// error about (1 + |t|) 1.2e-7 (the product's rounding and the logarithm's
// error are scaled by |t| ln 2), an absolute error below 3e-7, where colours
// are held to 1e-5.  x = +0 with y > 0 gives +0, as powf does.  Everything
// else (y <= 0 or NaN, subnormal or NaN x, larger results) is powf.
JR_DEV float pow_unit(float x, float y) {
  if (y > 0.0f) {
    if (x == 0.0f) return 0.0f;
    if (x >= 1.17549435e-38f) {
      const float t = y * __builtin_amdgcn_logf(x);
      if (t <= 0.5f) return __builtin_amdgcn_exp2f(t);
    }
  }
  return powf(x, y);
}

// jnp.modf fractional part: x - (x >= 0 ? floor(x) : ceil(x)).
JR_DEV float modf_frac(float x) {
  const float whole = (x >= 0.0f) ? floorf(x) : ceilf(x);
  return x - whole;
}

// Total order key of a depth for the argmin: NaN first (jnp.argmin returns the
// first NaN), -0 == +0, otherwise the IEEE order.
JR_DEV uint32_t depth_key(float z) {
  // z + 0 is +0 for z = -0 (round to nearest) and z itself otherwise; a
  // negative value has all its bits flipped, any other gets the sign bit set
  const float p = z + 0.0f;
  const uint32_t u = __float_as_uint(p);
  const uint32_t flip = (uint32_t)((int32_t)u >> 31) | 0x80000000u;
  return (p != p) ? 0u : (u ^ flip);
}

constexpr uint32_t kInfKey = 0xFF800000u;  // depth_key(+inf)

// The depth of a key from depth_key (not 0: NaN has no single inverse; -0
// comes back as +0).
JR_DEV float key_depth(uint32_t k) {
  return __uint_as_float((k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k);
}

}  // namespace jr
