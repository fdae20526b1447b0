// jr_forward.hip -- the forward rasterisation pass on CDNA4 (gfx950).
//
// Replaces the body of renderer/pipeline.py:render (vertex stage, per-primitive
// setup, the scan over rows x vmap over pixels x vmap over primitives, chooser,
// interpolate, fragment, mix and merge) with six kernel stages:
//
//   1. k_clear   per-pixel keys, tile counters, per-scene sample tables and the
//                scene of every 64-triangle wavefront.
//   2. k_setup   per triangle: clip-space corners (vertex stage positions),
//                det / keep / LU inverse (PerPrimitive.create,
//                renderer/pipeline.py:76-113) and a conservative pixel range;
//                small triangles resolved into the key buffer, the others
//                counted into the tiles they may cover.
//   3. k_wide    the tile counts of triangles spanning many tiles.
//   4. k_scan_*  exclusive scan of the per-tile counts (all scenes at once).
//   5. k_bin     triangle ids into the tile lists (k_wide for the wide ones).
//   6. k_raster  one group per (scene, tile): stages the tile's triangle
//                records in LDS, keeps the nearest (depth, id) key per pixel
//                (the default primitive_chooser: argmin with first-index ties),
//                then shades the winner and writes the merged buffers.
//
// Visibility is decided with exactly the reference's per-pixel arithmetic
// (renderer/pipeline.py:170-279); binning only removes (triangle, tile) pairs
// that provably fail that arithmetic at every pixel of the tile, so the result
// equals the brute-force O(W*H*F) evaluation bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/jrhip.h"
#include "jr_common.h"
#include "jr_math.h"
#include "jr_shade.h"

namespace jr {

// Conservative pixel range of a triangle.  The pixel test accepts (x, y) only
// when every float edge value is >= 0.  The float sample position and the
// three rounded operations keep each computed edge within about
// 5.1u(X|a| + Y|b| + |c|) of the exact edge at the exact sample position
// (u = 2^-24), so every accepted pixel's exact sample lies in the region
// R = {a_k x + b_k y + c_k >= -delta_k} with delta_k = 8u(X|a| + Y|b| + |c|).
// The range is the bounding box of R within the viewport's sample rectangle,
// mapped to pixel indices with a 1/64-pixel allowance.
//
// When the three shifted lines meet pairwise in three points that each satisfy
// the third inequality with room to spare, R is exactly the triangle they
// span, whose box comes from three 2x2 solves in float, widened by an explicit
// bound on their rounding (u = 2^-24 per operation; only systems far from
// singular, |det| > 1e-3 of the product of the rows' 1-norms, and no
// coefficient of magnitude in (0, 1e-12), so no intermediate is subnormal).
// Otherwise (R unbounded, empty or ill-conditioned, or coefficients of 1e18
// and more: about one triangle in a hundred) the range is the box of the
// viewport rectangles that the exact rectangle test cannot reject
// (fallback_range), and the exact per-tile test of the binning keeps only the
// tiles that may hold a covered pixel.
JR_DEV bool shifted_triangle_box_f(const float a[3], const float b[3], const float c[3],
                                   float* xmin, float* xmax, float* ymin, float* ymax) {
  constexpr float g = 8.0f / 16777216.0f;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float m[3] = {fabsf(a[k]), fabsf(b[k]), fabsf(c[k])};
    for (int q = 0; q < 3; ++q)
      if (m[q] != 0.0f && !(m[q] >= 1e-12f)) return false;
  }
  float lo_x = INFINITY, hi_x = -INFINITY, lo_y = INFINITY, hi_y = -INFINITY;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const int p = k, q = (k + 1) % 3, r = (k + 2) % 3;
    const float det = a[p] * b[q] - a[q] * b[p];
    const float scale = (fabsf(a[p]) + fabsf(b[p])) * (fabsf(a[q]) + fabsf(b[q]));
    if (!(fabsf(det) > 1e-3f * scale)) return false;
    const float inv = 1.0f / det;
    const float nx = b[p] * c[q] - b[q] * c[p];
    const float ny = a[q] * c[p] - a[p] * c[q];
    const float x = nx * inv, y = ny * inv;
    const float cond = (fabsf(a[p] * b[q]) + fabsf(a[q] * b[p])) * fabsf(inv);
    const float ex = g * ((fabsf(b[p] * c[q]) + fabsf(b[q] * c[p])) * fabsf(inv) +
                          fabsf(x) * cond + fabsf(x));
    const float ey = g * ((fabsf(a[q] * c[p]) + fabsf(a[p] * c[q])) * fabsf(inv) +
                          fabsf(y) * cond + fabsf(y));
    const float v = a[r] * x + b[r] * y + c[r];
    const float mag = fabsf(a[r] * x) + fabsf(b[r] * y) + fabsf(c[r]);
    if (!(v > fabsf(a[r]) * ex + fabsf(b[r]) * ey + g * mag)) return false;
    lo_x = fminf(lo_x, x - ex);
    hi_x = fmaxf(hi_x, x + ex);
    lo_y = fminf(lo_y, y - ey);
    hi_y = fmaxf(hi_y, y + ey);
  }
  if (!(lo_x <= hi_x) || !(lo_y <= hi_y) || !(hi_x - lo_x < 1e30f) || !(hi_y - lo_y < 1e30f))
    return false;
  *xmin = lo_x;
  *xmax = hi_x;
  *ymin = lo_y;
  *ymax = hi_y;
  return true;
}

// Pixels whose exact NDC sample lies in [xmin, xmax] x [ymin, ymax], with a
// 1/64-pixel allowance for the rounding of the box corners.
JR_DEV Rect pixel_box(double xmin, double xmax, double ymin, double ymax, const float* vp,
                      const Frame& f) {
  Rect r = {1, 0, 1, 0};
  const double vx = vp[0], vx3 = vp[3], vy = vp[5], vy3 = vp[7];
  constexpr double kSlack = 1.0 / 64.0;
  double i0 = ceil(xmin * vx + vx3 - kSlack), i1 = floor(xmax * vx + vx3 + kSlack);
  double j0 = ceil(ymin * vy + vy3 - kSlack), j1 = floor(ymax * vy + vy3 + kSlack);
  i0 = fmax(i0, 0.0);
  j0 = fmax(j0, 0.0);
  i1 = fmin(i1, (double)(f.W - 1));
  j1 = fmin(j1, (double)(f.H - 1));
  if (!(i0 <= i1) || !(j0 <= j1)) return r;
  r.tx0 = (int32_t)i0;
  r.tx1 = (int32_t)i1;
  r.ty0 = (int32_t)j0;
  r.ty1 = (int32_t)j1;
  return r;
}

JR_DEV bool rect_may_cover(const float4& r0, const float4& r1, const float4& r2,
                           bool sane, const float* vp, const float* __restrict__ sx,
                           const Frame& f, int32_t i0, int32_t i1, int32_t j0, int32_t j1);

// The range of a triangle whose region R has no usable bound (about one
// triangle in a hundred in a Brax-like frame: edge-on or crossing the eye
// plane): the whole viewport, unless the exact rectangle test finds that no
// pixel of it can pass (half of them cover no pixel at all).
JR_DEV Rect fallback_range(const float4& r0, const float4& r1, const float4& r2, bool sane,
                           const float* vp, const float* __restrict__ sx, const Frame& f) {
  if (!rect_may_cover(r0, r1, r2, sane, vp, sx, f, 0, f.W - 1, 0, f.H - 1))
    return Rect{1, 0, 1, 0};
  return Rect{0, f.W - 1, 0, f.H - 1};
}

// sx: the scene's sample table
JR_DEV Rect pixel_range(const float4& r0, const float4& r1, const float4& r2,
                        bool sane, const float* vp, const float* __restrict__ sx,
                        const Frame& f) {
  if (!sane || !(vp[0] > 0.0f) || !(vp[5] > 0.0f))
    return fallback_range(r0, r1, r2, sane, vp, sx, f);
  const double X0 = pix_x(0, vp), X1 = pix_x(f.W - 1, vp);
  const double Y0 = pix_y(0, vp), Y1 = pix_y(f.H - 1, vp);
  if (!(X0 <= X1) || !(Y0 <= Y1)) return fallback_range(r0, r1, r2, sane, vp, sx, f);
  const double XM = fmax(fabs(X0), fabs(X1)), YM = fmax(fabs(Y0), fabs(Y1));
  const float4 rs[3] = {r0, r1, r2};
  Rect r = {1, 0, 1, 0};
  double xmin, xmax, ymin, ymax;
  // float attempt: the shift is rounded up by a further u |C| (the bound above
  // has 2.9u to spare), so the float lines enclose the same region R
  float Af[3], Bf[3], Cf[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    Af[k] = rs[k].x;
    Bf[k] = rs[k].y;
    const float d = (8.0f / 16777216.0f) *
                    ((float)XM * fabsf(rs[k].x) + (float)YM * fabsf(rs[k].y) + fabsf(rs[k].z));
    Cf[k] = rs[k].z + d * 1.0001f + 1e-37f;
  }
  float fxmin, fxmax, fymin, fymax;
  if (shifted_triangle_box_f(Af, Bf, Cf, &fxmin, &fxmax, &fymin, &fymax)) {
    xmin = fmax((double)fxmin, X0);
    xmax = fmin((double)fxmax, X1);
    ymin = fmax((double)fymin, Y0);
    ymax = fmin((double)fymax, Y1);
    if (!(xmin <= xmax) || !(ymin <= ymax)) return r;
    return pixel_box(xmin, xmax, ymin, ymax, vp, f);
  }
  // R unbounded, empty or ill-conditioned: the fallback range; the exact
  // per-tile test then keeps only the tiles that may hold a covered pixel
  return fallback_range(r0, r1, r2, sane, vp, sx, f);
}

JR_DEV Rect tiles_of(const Rect& p, const Frame& f) {
  if (p.tx0 > p.tx1) return p;
  Rect t = {p.tx0 >> f.tw_sh, p.tx1 >> f.tw_sh, p.ty0 >> f.th_sh, p.ty1 >> f.th_sh};
  return t;
}

// Triangles whose pixel range holds at most this many pixels skip the tile
// bins: their setup thread evaluates the per-pixel test itself and folds the
// result into the per-pixel key buffer with a 64-bit atomic minimum.
constexpr int32_t kSmallPixels = 64;

// (depth key, primitive id) packed so that the unsigned minimum is the
// chooser's argmin with first-index ties.
JR_DEV uint64_t pack_key(uint32_t key, int32_t id) {
  return ((uint64_t)key << 32) | (uint32_t)id;
}

// Exact rejection of a rectangle of pixels (a tile, or a whole viewport).
// With finite coefficients every computed edge value is monotone in i and in
// j (IEEE rounding is monotone), so its maximum over a rectangle sits at a
// corner pixel.  With an infinite coefficient a (no NaN), x*a
// only depends on the sign class of x (-inf, NaN at x == 0, +inf) and the
// finite terms beside it no longer matter, so one pixel per sign class
// decides.  Testing the pixels {i0, i1, i with x == 0} x {j0, j1, j with
// y == 0} therefore finds a passing value whenever any pixel of the tile has
// one; a rectangle where some edge fails at all of them holds no pixel that
// passes the reference's inside test.
JR_DEV int32_t zero_pixel(float v3, int32_t lo, int32_t hi) {
  // the pixel whose NDC coordinate is exactly 0, if it lies in [lo, hi]
  const float r = floorf(v3);
  if (r != v3 || !(r >= (float)lo) || !(r <= (float)hi)) return -1;
  return (int32_t)r;
}

JR_DEV bool rect_may_cover(const float4& r0, const float4& r1, const float4& r2,
                           bool sane, const float* vp, const float* __restrict__ sx,
                           const Frame& f, int32_t i0, int32_t i1, int32_t j0, int32_t j1) {
  // sx: the scene's sample table (pix_x for i < W, then pix_y)
  float xs[3] = {sx[i0], sx[i1], 0.0f};
  float ys[3] = {sx[f.W + j0], sx[f.W + j1], 0.0f};
  int nx = 2, ny = 2;
  if (!sane) {
    if (!(vp[0] > 0.0f) || !(vp[5] > 0.0f)) return true;
    const int32_t iz = zero_pixel(vp[3], i0, i1), jz = zero_pixel(vp[7], j0, j1);
    if (iz >= 0) xs[nx++] = sx[iz];
    if (jz >= 0) ys[ny++] = sx[f.W + jz];
  }
  const float4 rs[3] = {r0, r1, r2};
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    bool pass = false;
    for (int p = 0; p < nx; ++p)
      for (int q = 0; q < ny; ++q) pass = pass || (edge(rs[k], xs[p], ys[q]) >= 0.0f);
    if (!pass) return false;
  }
  return true;
}

// rect_may_cover for the pixels of tile (tx, ty)
JR_DEV bool tile_may_cover(const float4& r0, const float4& r1, const float4& r2,
                           bool sane, const float* vp, const float* __restrict__ sx,
                           const Frame& f, int32_t tx, int32_t ty) {
  const int32_t i0 = tx * f.TW, i1 = min(i0 + f.TW, f.W) - 1;
  const int32_t j0 = ty * f.TH, j1 = min(j0 + f.TH, f.H) - 1;
  return rect_may_cover(r0, r1, r2, sane, vp, sx, f, i0, i1, j0, j1);
}

// Triangles whose tile range spans more than this many tiles are binned by a
// whole wavefront (one lane per tile) instead of by their own setup thread.
constexpr int32_t kWideTiles = 16;

// Offset in words from the wide-list length to k_bin's list length in the
// workspace's misc words: 128 bytes apart, so their atomics do not share a
// cache line.
constexpr int32_t kBinListAt = 32;

// Per-triangle binning word written by k_setup: kMaskNone (culled, small,
// or no tile may hold a covered pixel: nothing to bin and no tile range
// stored), kMaskWide (binned by k_wide), else bit q set for each counted tile
// (tx0 + q / h, ty0 + q % h) of the stored range, h = ty1 - ty0 + 1 (at most
// kWideTiles bits).
constexpr uint32_t kMaskNone = 0u;
constexpr uint32_t kMaskWide = ~0u;

JR_DEV bool has_nan_column(const float4& r) {
  return (r.x != r.x) || (r.y != r.y) || (r.z != r.z);
}

// Small-triangle keys of the narrow form.  They live in the workspace's key
// buffer as 64-bit words (tag << 32) | depth key, tag = ~epoch of the binning
// that wrote them, and stay there from call to call: each binning of the
// workspace takes the next epoch, whose smaller tag wins every atomic minimum
// against an older word, and a word with another tag reads as "no
// candidate".  So the buffer is not cleared per call, only when the key
// state does not describe it: a new or foreign workspace, another pixel
// count or buffer offset, any other binning on the workspace in between (the
// state sits at the workspace's start, where every layout keeps it, and the
// wide form's k_clear marks it stale), or the epoch running out.  The fill
// value ~0 has tag 0xFFFFFFFF, which no epoch >= 1 produces.
constexpr int kKsMagic = 0, kKsPixLo = 1, kKsPixHi = 2, kKsOffLo = 3, kKsOffHi = 4,
              kKsEpoch = 5, kKsFill = 6;
constexpr uint32_t kKeyMagic = 0x6b657973u;

JR_DEV uint64_t tagged_key(uint32_t key, uint32_t tag) {
  return ((uint64_t)tag << 32) | key;
}
JR_DEV uint32_t untag_key(uint64_t k, uint32_t tag) {
  return (uint32_t)(k >> 32) == tag ? (uint32_t)k : ~0u;
}

// Before a narrow binning: the next epoch, or a full fill of the key buffer
// (by k_clear) with epoch 1 when the state does not describe the buffer.
__global__ void k_keys_prepare(uint32_t* __restrict__ ks, int64_t n_pix, int64_t keys_at) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const uint32_t lo = (uint32_t)n_pix, hi = (uint32_t)((uint64_t)n_pix >> 32);
  const uint32_t olo = (uint32_t)keys_at, ohi = (uint32_t)((uint64_t)keys_at >> 32);
  const uint32_t e = ks[kKsEpoch];
  const bool ok = ks[kKsMagic] == kKeyMagic && ks[kKsPixLo] == lo && ks[kKsPixHi] == hi &&
                  ks[kKsOffLo] == olo && ks[kKsOffHi] == ohi && e >= 1u && e < 0xFFFFFFFEu;
  ks[kKsFill] = ok ? 0u : 1u;
  ks[kKsEpoch] = ok ? e + 1u : 1u;
  ks[kKsMagic] = kKeyMagic;
  ks[kKsPixLo] = lo;
  ks[kKsPixHi] = hi;
  ks[kKsOffLo] = olo;
  ks[kKsOffHi] = ohi;
}

// Clears the tile counters (n32 int32), sets the n64 per-pixel keys to "no
// candidate" (the narrow form: only when k_keys_prepare asked for a fill; the
// wide form: always, and the key state no longer describes the buffer) and
// fills every scene's sample table: the NDC coordinates pix_x(i) for i < W
// followed by pix_y(j) for j < H.
__global__ void k_clear(int32_t* __restrict__ p32, int64_t n32, uint64_t* __restrict__ p64,
                        int64_t n64, uint32_t* __restrict__ ks,
                        const jr_scene* __restrict__ scenes, Frame f,
                        float* __restrict__ samples, int32_t* __restrict__ wave_scene,
                        int32_t n_tris) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t first = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  // wave_scene[w] = the scene of triangle 64 w (the last scene whose
  // tri_begin <= 64 w): one wavefront per scene writes the entries whose
  // first triangle lies in [tri_begin of the scene, tri_begin of the next)
  {
    const int32_t n_w = (n_tris + 63) >> 6;
    const int lane = threadIdx.x & 63;
    for (int64_t s = first >> 6; s < f.n_scenes; s += stride >> 6) {
      const int32_t lo = s == 0 ? 0 : max(scenes[s].tri_begin, 0);
      const int32_t hi = s + 1 < f.n_scenes ? max(scenes[s + 1].tri_begin, 0) : n_tris;
      const int32_t w1 = min((hi + 63) >> 6, n_w);
      for (int32_t w = ((lo + 63) >> 6) + lane; w < w1; w += 64) wave_scene[w] = (int32_t)s;
    }
  }
  for (int64_t i = first; i < n32; i += stride) p32[i] = 0;
  if (f.narrow) {
    if (ks[kKsFill] == 0u) n64 = 0;
  } else if (first == 0) {
    ks[kKsMagic] = 0u;
  }
  for (int64_t i = first; i < n64; i += stride) p64[i] = ~0ull;
  const int32_t per = f.W + f.H;
  for (int64_t i = first; i < (int64_t)f.n_scenes * per; i += stride) {
    const int32_t s = (int32_t)(i / per), q = (int32_t)(i - (int64_t)s * per);
    const float* vp = scenes[s].viewport;
    samples[i] = q < f.W ? pix_x(q, vp) : pix_y(q - f.W, vp);
  }
}

// Stage 1: vertex positions -> clip space, PerPrimitive.create; small
// triangles are resolved into the key buffer here, large ones counted into
// the tiles they may cover.
//
// The pixels of a wavefront's small triangles are spread over its 64 lanes:
// a wavefront prefix sum of the triangles' pixel-box areas numbers all of
// their (triangle, pixel) pairs, and each pass of the loop gives one pair to
// every lane (the owning triangle found by a binary search over the prefix
// sums in LDS).  A wavefront therefore takes sum(areas) / 64 steps instead of
// max(area) steps of one lane per triangle.
// True when the window depth ((e0 z0 + e1 z1) + e2 z2) vp22 + vp23 stays
// below 1e37 in magnitude for every sample of the pixel box: then no product
// overflows and, with finite z_k, the depth is finite (never NaN) wherever
// the pixel test passes.  |e_k| <= |a_k| X + |b_k| Y + |c_k| over the box.
JR_DEV bool depth_bounded(const float4& r0, const float4& r1, const float4& r2,
                          const Rect& px, const float* vp, const float* sx, const Frame& f) {
  const float X = fmaxf(fabsf(sx[px.tx0]), fabsf(sx[px.tx1]));
  const float Y = fmaxf(fabsf(sx[f.W + px.ty0]), fabsf(sx[f.W + px.ty1]));
  const float4 rs[3] = {r0, r1, r2};
  float b = 0.0f;
#pragma unroll
  for (int k = 0; k < 3; ++k)
    b += (fabsf(rs[k].x) * X + fabsf(rs[k].y) * Y + fabsf(rs[k].z)) * 1.001f * fabsf(rs[k].w);
  b = b * fabsf(vp[10]) + fabsf(vp[11]);
  return b < 1e37f;
}

// Orders a tile's LDS staging against its use: a workgroup barrier, or, when
// each wavefront of the workgroup works on a tile of its own, a wavefront
// barrier (a wavefront's LDS operations complete in issue order, so keeping
// the compiler from moving memory operations across the point is enough).
template <bool WAVE>
JR_DEV void tile_sync() {
  if (WAVE) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

// max(v, v moved by the DPP pattern CTRL into the rows of ROW_MASK); lanes
// without a source compare against -1.
template <int CTRL, int ROW_MASK>
JR_DEV int32_t dpp_max_step(int32_t v) {
  return max(v, __builtin_amdgcn_update_dpp(-1, v, CTRL, ROW_MASK, 0xf, false));
}

// For lane k owning items [first_k, first_k + count_k) (first ascending over
// the lanes, count 0 allowed): the owner of item base + lane, i.e. the last
// lane whose first item is <= it.  Lanes whose range starts inside the window
// mark their start in the wavefront's 64-entry LDS scratch `mark`; lane 0
// starts from the last lane starting at or before base (a ballot), and an
// inclusive prefix maximum over the lanes (DPP: row shifts 1, 2, 4, 8, then
// the row broadcasts 15 and 31) carries each owner forward.  Call with every
// lane of the wavefront active.
JR_DEV int32_t wave_owner(int32_t first, int32_t count, int32_t base, int8_t* mark,
                          int lane) {
  mark[lane] = -1;
  tile_sync<true>();
  if (count > 0 && first > base && first < base + 64) mark[first - base] = (int8_t)lane;
  tile_sync<true>();
  int32_t v = mark[lane];
  const uint64_t le = __ballot(first <= base);
  if (lane == 0) v = 63 - __clzll((long long)le);
  v = dpp_max_step<0x111, 0xf>(v);
  v = dpp_max_step<0x112, 0xf>(v);
  v = dpp_max_step<0x114, 0xf>(v);
  v = dpp_max_step<0x118, 0xf>(v);
  v = dpp_max_step<0x142, 0xa>(v);
  v = dpp_max_step<0x143, 0xc>(v);
  return v;
}

constexpr int kSetupThreads = 256;

// Per-lane data of k_setup's two cooperative loops.  A live triangle is
// small (its pixels are tested one per lane-step), or counted into at most
// kWideTiles tiles (one tile test per lane-step), or neither; box is its pixel
// box (tx0, ty0, height, scene) or its tile range (tx0, ty0, height in tiles,
// scene | sane << 30) accordingly.
struct SmallStage {
  float4 r[3][kSetupThreads];
  int4 box[kSetupThreads];
  int16_t off[kSetupThreads];   // exclusive prefix sum of the areas per wavefront
  union {
    struct {
      int32_t tri[kSetupThreads];
      int16_t toff[kSetupThreads];    // exclusive prefix sum of the tile counts
      uint32_t tmask[kSetupThreads];  // tiles that passed the test (bit q)
    } t;
    // after the tile tests: the sample table of the workgroup's one scene,
    // over toff and tmask (and over tri in the narrow form, which keeps no
    // triangle ids in its keys)
    struct {
      int32_t tri_keep[kSetupThreads];
      float x[(kSetupThreads * 6) / 4];
    } smp_wide;
    float smp_narrow[(kSetupThreads * 10) / 4];
  } u;
};
constexpr int kSmpWide = (kSetupThreads * 6) / 4;
constexpr int kSmpNarrow = (kSetupThreads * 10) / 4;

// A triangle handed from the first half of k_setup to the second: its clip
// corners (x, y, z, w) and (id, scene, det bits).
struct SetupItem {
  float4 c[3];
  int4 meta;
};

union SetupStage {
  SetupItem item[kSetupThreads];
  SmallStage sm;
};

__global__ __launch_bounds__(kSetupThreads) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_setup(
    jr_inputs in, const jr_scene* __restrict__ scenes, Frame f,
    float4* __restrict__ rec, Rect* __restrict__ rects,
    int32_t* __restrict__ counts, uint64_t* __restrict__ keys,
    int2* __restrict__ wide, int32_t* __restrict__ wide_count,
    const float* __restrict__ samples, const uint32_t* __restrict__ ks,
    uint32_t* __restrict__ masks, const int32_t* __restrict__ wave_scene,
    int2* __restrict__ blist) {
  __shared__ SetupStage stage;
  __shared__ int32_t wave_live[kSetupThreads / 64];
  __shared__ int32_t bl_n[kSetupThreads / 64 + 1];  // list entries per wavefront, then the base
  // wave_owner scratch, 64 per wavefront, in bytes: with 4-byte entries the
  // workgroup's LDS would pass 20 KiB, an eighth of the CU's 160 KiB, and 7
  // workgroups would fit on a CU instead of 8
  __shared__ int8_t mark[kSetupThreads];
  // lowest and highest scene among the workgroup's small triangles
  __shared__ int32_t wg_scene[2];
  SmallStage& sm = stage.sm;
  if (threadIdx.x == 0) {
    wg_scene[0] = 0x7FFFFFFF;
    wg_scene[1] = -1;
  }
  const int lane = threadIdx.x & 63;
  const int wbase = threadIdx.x & ~63;

  // First half, every triangle: corners, det, keep / front.  Only a kept
  // front-facing triangle can win a pixel; besides those only a scene's
  // first triangle (argmin's choice when nothing covers a pixel) has its
  // record read.  Only these "live" triangles go on to the inverse and the
  // pixel range; the others just record that they have nothing to bin.  The
  // live ones are packed to the front of the workgroup, so the second half
  // runs on about half as many wavefronts.
  {
    const int32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    bool live = false;
    float clip[3][4];
    float det = 0.0f;
    int32_t s = 0;
    if (__builtin_amdgcn_readfirstlane(t & ~63) < in.n_tris)
      s = scene_of_wave(scenes, f.n_scenes, wave_scene, t, in.n_tris);
    if (t < in.n_tris) {
      const jr_scene& sc = scenes[s];
      // a wavefront of one scene (nearly every one) reads the camera matrix
      // with scalar loads
      const int32_t s_u = __builtin_amdgcn_readfirstlane(s);
      if (__ballot(s != s_u) == 0)
        tri_corners(in, scenes[s_u], t, clip);
      else
        tri_corners(in, sc, t, clip);
      float M[9];
      det = tri_det(clip, M);
      const bool keep = fabsf(det) > 1e-6f, front = det >= 0.0f;
      live = (keep && front) || t == sc.tri_begin;
      if (!live) masks[t] = kMaskNone;  // nothing to bin, nothing else read
    }
    const uint64_t bal = __ballot(live);
    if (lane == 0) wave_live[threadIdx.x >> 6] = __popcll(bal);
    __syncthreads();
    int32_t pos = __builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32),
                                            __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
    for (int w = 0; w < (int)(threadIdx.x >> 6); ++w) pos += wave_live[w];
    if (live) {
      SetupItem& it = stage.item[pos];
#pragma unroll
      for (int k = 0; k < 3; ++k)
        it.c[k] = make_float4(clip[k][0], clip[k][1], clip[k][2], clip[k][3]);
      it.meta = make_int4(t, s, (int32_t)__float_as_uint(det), 0);
    }
    __syncthreads();
  }
  int32_t n_live = 0;
#pragma unroll
  for (int w = 0; w < kSetupThreads / 64; ++w) n_live += wave_live[w];
  const bool valid = (int32_t)threadIdx.x < n_live;
  TriSetup ts;
  int32_t t = 0, s = 0;
  if (valid) {
    const SetupItem& it = stage.item[threadIdx.x];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const float4 c = it.c[k];
      ts.clip[k][0] = c.x;
      ts.clip[k][1] = c.y;
      ts.clip[k][2] = c.z;
      ts.clip[k][3] = c.w;
    }
    t = it.meta.x;
    s = it.meta.y;
    ts.det = __uint_as_float((uint32_t)it.meta.z);
  }
  __syncthreads();  // the items' storage becomes the small-triangle stage

  // Second half, live triangles: inverse, records, pixel range; small
  // triangles are resolved into the key buffer here, the others counted into
  // the tiles they may cover.
  int32_t n_small = 0, n_tile = 0;
  if (valid) {
    const jr_scene& sc = scenes[s];
    tri_det(ts.clip, ts.M);
    tri_finish(ts);
    const bool keep = ts.keep, front = ts.front, sane = ts.sane;
    const float4 r0 = rec_col(ts, 0);
    const float4 r1 = rec_col(ts, 1);
    const float4 r2 = rec_col(ts, 2);
    Rect r = {1, 0, 1, 0};
    Rect px = r;
    // a NaN coefficient makes that edge value NaN at every pixel: never inside
    const bool nan_col = has_nan_column(r0) || has_nan_column(r1) || has_nan_column(r2);
    if (keep && front && !nan_col)
      px = pixel_range(r0, r1, r2, sane, sc.viewport, samples + (int64_t)s * (f.W + f.H), f);
    const uint32_t flags = (keep ? kKeep : 0u) | (front ? kFront : 0u) | (sane ? kSane : 0u);
    const int64_t area = (px.tx0 > px.tx1)
        ? 0 : (int64_t)(px.tx1 - px.tx0 + 1) * (px.ty1 - px.ty0 + 1);
    // the narrow form reconstructs a small winner's depth from its key, so its
    // small triangles must have a finite depth at every pixel of the box
    const bool small = area > 0 && area <= kSmallPixels &&
                       (!f.narrow || depth_bounded(r0, r1, r2, px, sc.viewport,
                                                   samples + (int64_t)s * (f.W + f.H), f));
    // the record is read for binned triangles, for pixel winners and for the
    // scene's first triangle; the narrow form finishes its small winners
    // from their keys alone, and a triangle whose pixel range is empty is
    // never binned (its mask is empty, so not even an overflowing tile's scan
    // reads it) and never wins
    if ((area > 0 && !(f.narrow && small)) || t == sc.tri_begin) {
      rec[4 * (int64_t)t + 0] = r0;
      rec[4 * (int64_t)t + 1] = r1;
      rec[4 * (int64_t)t + 2] = r2;
      rec[4 * (int64_t)t + 3] = make_float4(ts.det, __uint_as_float(flags), 0.0f, 0.0f);
    }
    if (area > 0 && !small) r = tiles_of(px, f);
    if (small) {
      n_small = (int32_t)area;
      atomicMin(&wg_scene[0], s);
      atomicMax(&wg_scene[1], s);
      sm.r[0][threadIdx.x] = r0;
      sm.r[1][threadIdx.x] = r1;
      sm.r[2][threadIdx.x] = r2;
      sm.box[threadIdx.x] = make_int4(px.tx0, px.ty0, px.ty1 - px.ty0 + 1, s);
      sm.u.t.tri[threadIdx.x] = t;
    }
    const int64_t n_t =
        (r.tx0 > r.tx1) ? 0 : (int64_t)(r.tx1 - r.tx0 + 1) * (r.ty1 - r.ty0 + 1);
    if (n_t > 0) rects[t] = r;
    if (n_t > kWideTiles) {
      // counted by k_wide (one lane per tile)
      wide[atomicAdd(wide_count, 1)] = make_int2(t, s);
      masks[t] = kMaskWide;
    } else if (n_t == 0) {
      masks[t] = kMaskNone;
    } else {
      // counted below, one (triangle, tile) pair per lane-step
      n_tile = (int32_t)n_t;
      sm.r[0][threadIdx.x] = r0;
      sm.r[1][threadIdx.x] = r1;
      sm.r[2][threadIdx.x] = r2;
      sm.box[threadIdx.x] =
          make_int4(r.tx0, r.ty0, r.ty1 - r.ty0 + 1, s | (sane ? (1 << 30) : 0));
    }
  }
  int32_t incl = n_small, incl_t = n_tile;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int32_t v = __shfl_up(incl, d, 64);
    const int32_t vt = __shfl_up(incl_t, d, 64);
    if (lane >= d) {
      incl += v;
      incl_t += vt;
    }
  }
  const int32_t total = __shfl(incl, 63, 64);
  const int32_t total_t = __shfl(incl_t, 63, 64);
  sm.off[threadIdx.x] = (int16_t)(incl - n_small);
  sm.u.t.toff[threadIdx.x] = (int16_t)(incl_t - n_tile);
  sm.u.t.tmask[threadIdx.x] = 0u;
  __syncthreads();
  // the tile test of every (triangle, tile) pair of the wavefront's counted
  // triangles: count the tiles that may hold a covered pixel and remember
  // which ones (bit q: tile (tx0 + q / h, ty0 + q % h)) for k_bin
  for (int32_t base = 0; base < total_t; base += 64) {
    const int32_t item = base + lane;
    const int o = wbase + wave_owner(incl_t - n_tile, n_tile, base, mark + wbase, lane);
    if (item >= total_t) continue;
    const int4 box = sm.box[o];
    const int32_t q = item - sm.u.t.toff[o];
    const int32_t so = box.w & 0x3FFFFFFF;
    const int32_t dq = div_small(q, box.z);
    const int32_t tx = box.x + dq, ty = box.y + (q - dq * box.z);
    if (tile_may_cover(sm.r[0][o], sm.r[1][o], sm.r[2][o], (box.w >> 30) != 0,
                       scenes[so].viewport, samples + (int64_t)so * (f.W + f.H), f, tx, ty)) {
      atomicAdd(counts + (int64_t)so * f.T + tx * f.tiles_y + ty, 1);
      atomicOr(&sm.u.t.tmask[o], 1u << q);
    }
  }
  __syncthreads();
  // the counted triangles with a tile to bin go on k_bin's list as
  // (triangle, scene), the workgroup's entries together: one atomic per
  // workgroup on the list length (its own cache line, kBinListAt words past
  // wide_count)
  {
    const uint32_t tm = n_tile > 0 ? sm.u.t.tmask[threadIdx.x] : 0u;
    if (n_tile > 0) masks[t] = tm;
    const uint64_t bb = __ballot(tm != 0u);
    if (lane == 0) bl_n[threadIdx.x >> 6] = (int32_t)__popcll(bb);
    __syncthreads();
    if (threadIdx.x == 0) {
      int32_t total = 0;
#pragma unroll
      for (int w = 0; w < kSetupThreads / 64; ++w) total += bl_n[w];
      bl_n[kSetupThreads / 64] = total > 0 ? atomicAdd(wide_count + kBinListAt, total) : 0;
    }
    __syncthreads();
    if (tm != 0u) {
      int32_t at = bl_n[kSetupThreads / 64] + (int32_t)__popcll(bb & ((1ull << lane) - 1ull));
      for (int w = 0; w < (int)(threadIdx.x >> 6); ++w) at += bl_n[w];
      blist[at] = make_int2(t, s);
    }
  }
  // The sample table of the workgroup's small triangles' scene goes to LDS
  // when they all share one scene and it fits: the loop below then reads no
  // global memory, and its key atomics never hold up a load (a wavefront's
  // memory operations complete in order).
  const int32_t s_lo = wg_scene[0], s_hi = wg_scene[1];
  const int32_t per = f.W + f.H;
  const bool staged = s_lo == s_hi && per <= (f.narrow ? kSmpNarrow : kSmpWide);
  float* smp = f.narrow ? sm.u.smp_narrow : sm.u.smp_wide.x;
  float vz0 = 0.0f, vz1 = 0.0f;
  if (staged) {
    __syncthreads();  // every tmask read before the table overwrites it
    const float* src = samples + (int64_t)s_lo * per;
    for (int32_t k = threadIdx.x; k < per; k += kSetupThreads) smp[k] = src[k];
    vz0 = scenes[s_lo].viewport[10];
    vz1 = scenes[s_lo].viewport[11];
    __syncthreads();
  }
  const uint32_t tag = f.narrow ? ~ks[kKsEpoch] : 0u;  // this binning's key tag
  // the tile raster's exact per-pixel arithmetic for every (triangle, pixel)
  auto small_pairs = [&](auto from_lds) {
    constexpr bool LDS = decltype(from_lds)::value;
    for (int32_t base = 0; base < total; base += 64) {
      const int32_t item = base + lane;
      const int o = wbase + wave_owner(incl - n_small, n_small, base, mark + wbase, lane);
      if (item >= total) continue;
      const int4 box = sm.box[o];
      const int32_t q = item - sm.off[o];
      const int32_t di = div_small(q, box.z);
      const int32_t i = box.x + di, j = box.y + (q - di * box.z);
      const float4 q0 = sm.r[0][o], q1 = sm.r[1][o], q2 = sm.r[2][o];
      float x, y;
      if constexpr (LDS) {
        x = smp[i];
        y = smp[f.W + j];
      } else {
        const float* sx = samples + (int64_t)box.w * per;
        x = sx[i];
        y = sx[f.W + j];
      }
      const float e0 = edge(q0, x, y), e1 = edge(q1, x, y), e2 = edge(q2, x, y);
      if (e0 >= 0.0f && e1 >= 0.0f && e2 >= 0.0f) {
        float v0 = vz0, v1 = vz1;
        if constexpr (!LDS) {
          const float* vp = scenes[box.w].viewport;
          v0 = vp[10];
          v1 = vp[11];
        }
        const uint32_t key = depth_key(win_depth(q0, q1, q2, e0, e1, e2, v0, v1));
        const int64_t pix = ((int64_t)box.w * f.W + i) * f.H + j;
        if (f.narrow)
          atomicMin((unsigned long long*)(keys + pix), (unsigned long long)tagged_key(key, tag));
        else
          atomicMin((unsigned long long*)(keys + pix),
                    (unsigned long long)pack_key(key, sm.u.t.tri[o]));
      }
    }
  };
  if (staged)
    small_pairs(std::true_type{});
  else
    small_pairs(std::false_type{});
}

// Counts (BIN = false) or bins (BIN = true) the wide triangles: one wavefront
// per triangle, its lanes striding over the triangle's tiles.  The counting
// pass also keeps each triangle's tile-test results as a bitmap (one word per
// 32 tiles), so the binning pass reads bits instead of testing the tiles
// again: entry w of the wide list owns the fixed words [w * per, (w + 1) *
// per), per = ceil(T / 32), while they fit in the pool; later entries test
// again.
constexpr int32_t kWideBitWords = 1 << 20;

template <bool BIN>
__global__ __launch_bounds__(256) void k_wide(
    const jr_scene* __restrict__ scenes, Frame f, const float4* __restrict__ rec,
    const Rect* __restrict__ rects, const int2* __restrict__ wide,
    const int32_t* __restrict__ wide_count, int32_t* __restrict__ counts,
    const int32_t* __restrict__ offsets, int32_t* __restrict__ cursor,
    int32_t* __restrict__ bins, const float* __restrict__ samples,
    uint32_t* __restrict__ wbits) {
  const int lane = threadIdx.x & 63;
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t n_waves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  const int32_t n = *wide_count;
  for (int64_t w = wave; w < n; w += n_waves) {
    const int2 ts = wide[w];  // (triangle, scene)
    const int32_t t = ts.x, s = ts.y;
    const Rect r = rects[t];
    const int32_t h = r.ty1 - r.ty0 + 1;
    const int32_t n_t = (r.tx1 - r.tx0 + 1) * h;
    const int64_t tbase = (int64_t)s * f.T;
    const int32_t per = (f.T + 31) >> 5;
    const int32_t base = (w < kWideBitWords / per) ? (int32_t)w * per : -1;
    float4 r0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), r1 = r0, r2 = r0;
    bool sane = false;
    const float* vp = scenes[s].viewport;
    if (!BIN || base < 0) {
      r0 = rec[4 * (int64_t)t];
      r1 = rec[4 * (int64_t)t + 1];
      r2 = rec[4 * (int64_t)t + 2];
      sane = (__float_as_uint(rec[4 * (int64_t)t + 3].y) & kSane) != 0;
    }
    for (int32_t q0 = 0; q0 < n_t; q0 += 64) {
      const int32_t q = q0 + lane;
      bool pass = false;
      if (q < n_t) {
        const int32_t dq = div_small(q, h);
        const int32_t tx = r.tx0 + dq, ty = r.ty0 + (q - dq * h);
        if (BIN && base >= 0)
          pass = ((wbits[base + (q >> 5)] >> (q & 31)) & 1u) != 0u;
        else
          pass = tile_may_cover(r0, r1, r2, sane, vp, samples + (int64_t)s * (f.W + f.H), f,
                                tx, ty);
        if (pass) {
          const int64_t gt = tbase + tx * f.tiles_y + ty;
          if (BIN) {
            const int64_t pos = (int64_t)offsets[gt] + atomicAdd(cursor + gt, 1);
            if (pos < f.capacity) bins[pos] = t;
          } else {
            atomicAdd(counts + gt, 1);
          }
        }
      }
      if (!BIN) {
        const uint64_t b = __ballot(pass);
        if (base >= 0 && lane == 0) {
          wbits[base + (q0 >> 5)] = (uint32_t)b;
          if (q0 + 32 < n_t) wbits[base + (q0 >> 5) + 1] = (uint32_t)(b >> 32);
        }
      }
    }
  }
}

// Stage 2: exclusive scan of n counts -> offsets (two levels, 2048 per block).
constexpr int kScanThreads = 256;
constexpr int kScanPer = 8;
constexpr int kScanBlock = kScanThreads * kScanPer;

JR_DEV int32_t block_excl_scan(int32_t v, int32_t* lds, int32_t* total) {
  const int tid = threadIdx.x;
  lds[tid] = v;
  __syncthreads();
  for (int off = 1; off < kScanThreads; off <<= 1) {
    const int32_t add = (tid >= off) ? lds[tid - off] : 0;
    __syncthreads();
    lds[tid] += add;
    __syncthreads();
  }
  const int32_t incl = lds[tid];
  *total = lds[kScanThreads - 1];
  __syncthreads();
  return incl - v;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_local(
    const int32_t* __restrict__ counts, int32_t* __restrict__ offsets,
    int32_t* __restrict__ block_sums, int64_t n) {
  __shared__ int32_t lds[kScanThreads];
  const int64_t base = (int64_t)blockIdx.x * kScanBlock + threadIdx.x * kScanPer;
  int32_t v[kScanPer];
  int32_t sum = 0;
#pragma unroll
  for (int q = 0; q < kScanPer; ++q) {
    v[q] = (base + q < n) ? counts[base + q] : 0;
    sum += v[q];
  }
  int32_t total;
  int32_t run = block_excl_scan(sum, lds, &total);
#pragma unroll
  for (int q = 0; q < kScanPer; ++q) {
    if (base + q < n) offsets[base + q] = run;
    run += v[q];
  }
  if (threadIdx.x == 0) block_sums[blockIdx.x] = total;
}

// Scans the block sums in place (one group, any count) and stores the grand
// total in *total_out.
__global__ __launch_bounds__(kScanThreads) void k_scan_top(
    int32_t* __restrict__ block_sums, int64_t nblk, int32_t* __restrict__ total_out) {
  __shared__ int32_t lds[kScanThreads];
  int32_t carry = 0;
  for (int64_t c = 0; c < nblk; c += kScanThreads) {
    const int64_t i = c + threadIdx.x;
    const int32_t v = (i < nblk) ? block_sums[i] : 0;
    int32_t total;
    const int32_t ex = block_excl_scan(v, lds, &total);
    if (i < nblk) block_sums[i] = carry + ex;
    carry += total;
  }
  if (threadIdx.x == 0) *total_out = carry;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_add(
    int32_t* __restrict__ offsets, const int32_t* __restrict__ block_sums,
    int64_t n) {
  const int64_t base = (int64_t)blockIdx.x * kScanBlock;
  const int32_t add = block_sums[blockIdx.x];
  for (int q = threadIdx.x; q < kScanBlock; q += kScanThreads)
    if (base + q < n) offsets[base + q] += add;
}

// Stage 3: write triangle ids into the tile lists.
// One (triangle, tile) pair per lane-step: the wavefront's binned triangles'
// set mask bits are numbered by a prefix sum of their counts and spread over
// the lanes, so every step issues up to 64 independent slot claims instead of
// each lane claiming its tiles one after another.  All staging is
// wavefront-local.
__global__ __launch_bounds__(256) void k_bin(
    Frame f, const Rect* __restrict__ rects,
    const int32_t* __restrict__ offsets, int32_t* __restrict__ cursor,
    int32_t* __restrict__ bins, const uint32_t* __restrict__ masks,
    const int2* __restrict__ blist, const int32_t* __restrict__ blist_n) {
  __shared__ int4 own[256];  // tile range (tx0, ty0, height in tiles) and s * T
  __shared__ int32_t own_t[256];
  __shared__ uint32_t own_m[256];
  __shared__ int32_t own_off[256];  // exclusive prefix sum of the bit counts
  __shared__ int8_t mark[256];      // wave_owner scratch, 64 per wavefront
  const int lane = threadIdx.x & 63;
  const int wbase = threadIdx.x & ~63;
  const int32_t n = *blist_n;
  // the wavefronts stride over k_setup's list, 64 entries at a time
  const int32_t wave = (int32_t)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int32_t n_waves = (int32_t)((gridDim.x * blockDim.x) >> 6);
  for (int32_t first = wave * 64; first < n; first += n_waves * 64) {
    const int32_t i = first + lane;
    int32_t n_b = 0;
    if (i < n) {
      const int2 ts = blist[i];  // (triangle, scene)
      const uint32_t mask = masks[ts.x];  // the tiles k_setup counted
      const Rect r = rects[ts.x];
      own[threadIdx.x] = make_int4(r.tx0, r.ty0, r.ty1 - r.ty0 + 1, ts.y * f.T);
      own_t[threadIdx.x] = ts.x;
      own_m[threadIdx.x] = mask;
      n_b = __popc(mask);
    }
    int32_t incl = n_b;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const int32_t v = __shfl_up(incl, d, 64);
      if (lane >= d) incl += v;
    }
    const int32_t total = __shfl(incl, 63, 64);
    own_off[threadIdx.x] = incl - n_b;
    tile_sync<true>();
    for (int32_t base = 0; base < total; base += 64) {
      const int32_t item = base + lane;
      const int o = wbase + wave_owner(incl - n_b, n_b, base, mark + wbase, lane);
      if (item >= total) continue;
      uint32_t m = own_m[o];
      for (int32_t k = item - own_off[o]; k > 0; --k) m &= m - 1u;  // k-th set bit
      const int q = __ffs(m) - 1;
      const int4 box = own[o];
      const int32_t dq = div_small(q, box.z);
      const int32_t tx = box.x + dq, ty = box.y + (q - dq * box.z);
      const int64_t gt = (int64_t)box.w + tx * f.tiles_y + ty;
      const int64_t pos = (int64_t)offsets[gt] + atomicAdd(cursor + gt, 1);
      if (pos < f.capacity) bins[pos] = own_t[o];
    }
    tile_sync<true>();  // the entries' staging is rewritten by the next step
  }
}

// True when k_setup counted tile (tx, ty) for triangle t (any tile of its
// range for a wide triangle).
JR_DEV bool lists_tile(const Rect* __restrict__ rects, const uint32_t* __restrict__ masks,
                       int32_t t, int32_t tx, int32_t ty) {
  const uint32_t m = masks[t];
  if (m == kMaskNone) return false;
  const Rect r = rects[t];
  if (tx < r.tx0 || tx > r.tx1 || ty < r.ty0 || ty > r.ty1) return false;
  if (m == kMaskWide) return true;
  const int32_t q = (tx - r.tx0) * (r.ty1 - r.ty0 + 1) + (ty - r.ty0);
  return ((m >> q) & 1u) != 0u;
}

// The shadow pass's binning as a phong_reflection_shadow raster reads it to
// compute the shadow depths it needs itself (jr_shadow_source): the shadow
// pass's records, tile lists and sample tables, its per-texel small-triangle
// keys (narrow form), its frame and scene records, and the texel buffer each
// depth found is written to.
struct ShadowLookup {
  const float4* rec;
  const int32_t* counts;
  const int32_t* offsets;
  const int32_t* bins;
  const Rect* rects;
  const uint32_t* masks;
  const float* samples;
  const uint64_t* keys;  // its tagged small-triangle keys
  const uint32_t* ks;    // its key state (the current epoch)
  const jr_scene* scenes;
  float* texels;
  Frame f;
  float zdef, depth_offset;
};

// Depth of texel round(x, y) of scene s (Shadow.get, renderer/shadow.py:127-153:
// half away from zero, negative indices from the end, +inf out of range) as
// the shadow pass's raster would have written it: its small-triangle key,
// then the texel's tile list (every triangle of the scene that lists the tile
// when the list overflowed) with the raster's edge, depth-key and tie rules,
// the chooser's primitive-0 fallback, the default depth where nothing covers
// the texel, plus the offset.  The depth is stored in the texel buffer.
// Entries outside the scene's triangles are skipped, so a workspace that no
// binning of this frame wrote gives wrong depths but no stray reads.
JR_DEV float shadow_depth_on_demand(const ShadowLookup& L, int32_t s, float x, float y) {
  const Frame& f = L.f;
  int32_t ix = f2i_sat(roundf(x));
  int32_t iy = f2i_sat(roundf(y));
  if (ix < 0) ix += f.W;
  if (iy < 0) iy += f.H;
  if (ix < 0 || ix >= f.W || iy < 0 || iy >= f.H) return __builtin_inff();
  const jr_scene& sc = L.scenes[s];
  const int32_t b0 = sc.tri_begin, b1 = sc.tri_end;
  const int64_t pix = (int64_t)s * f.W * f.H + (int64_t)ix * f.H + iy;
  const float* smp = L.samples + (int64_t)s * (f.W + f.H);
  const float px = smp[ix], py = smp[f.W + iy];
  const float vp22 = sc.viewport[10], vp23 = sc.viewport[11];
  uint32_t best_key = 0xFFFFFFFFu;
  int32_t best_id = 0x7FFFFFFF;
  const uint32_t k0 = untag_key(L.keys[pix], ~L.ks[kKsEpoch]);
  if (k0 != ~0u) {
    best_key = k0;
    best_id = -1;  // small winner: its id is not kept, no equal key displaces it
  }
  const int32_t tx = ix >> f.tw_sh, ty = iy >> f.th_sh;
  const int64_t gt = (int64_t)s * f.T + (int64_t)tx * f.tiles_y + ty;
  const int32_t cnt = L.counts[gt];
  const int64_t start = L.offsets[gt];
  const bool overflow = start < 0 || cnt < 0 || start + cnt > f.capacity;
  const int32_t n = overflow ? b1 - b0 : cnt;
  // two list entries per step: their ids, then their records, are loaded
  // together (the (key, id) minimum does not depend on the visiting order)
  for (int32_t m0 = 0; m0 < n; m0 += 2) {
    int32_t tt[2];
    float4 ra[2], rb[2], rc[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int32_t m = m0 + u;
      int32_t t = -1;
      if (m < n) t = overflow ? b0 + m : L.bins[start + m];
      if (t < b0 || t >= b1) t = -1;
      tt[u] = t;
    }
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      ra[u] = rb[u] = rc[u] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      if (tt[u] >= 0) {
        const int64_t tq = 4 * (int64_t)tt[u];
        ra[u] = L.rec[tq];
        rb[u] = L.rec[tq + 1];
        rc[u] = L.rec[tq + 2];
      }
    }
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int32_t t = tt[u];
      if (t < 0) continue;
      if (overflow && !lists_tile(L.rects, L.masks, t, tx, ty)) continue;
      const float e0 = edge(ra[u], px, py), e1 = edge(rb[u], px, py), e2 = edge(rc[u], px, py);
      if (e0 >= 0.0f && e1 >= 0.0f && e2 >= 0.0f) {
        const uint32_t key = depth_key(win_depth(ra[u], rb[u], rc[u], e0, e1, e2, vp22, vp23));
        if (key < best_key || (key == best_key && t < best_id)) {
          best_key = key;
          best_id = t;
        }
      }
    }
  }
  float d;
  if (best_id == -1 ||
      (best_key < kInfKey && best_key != 0u && best_key != 0x80000000u)) {
    d = key_depth(best_key) + L.depth_offset;
  } else {
    // no finite winner (argmin's index 0 decides with its own keep & inside),
    // or a NaN / zero key whose depth the key does not give back
    int32_t w = -1;
    bool cover = false;
    if (b1 > b0) {
      if (best_key >= kInfKey) {
        w = b0;
        const float4 r0 = L.rec[4 * (int64_t)w], r1 = L.rec[4 * (int64_t)w + 1],
                     r2 = L.rec[4 * (int64_t)w + 2];
        const uint32_t fl = __float_as_uint(L.rec[4 * (int64_t)w + 3].y);
        cover = (fl & kKeep) && edge(r0, px, py) >= 0.0f && edge(r1, px, py) >= 0.0f &&
                edge(r2, px, py) >= 0.0f;
      } else {
        w = best_id;
        cover = true;
      }
    }
    if (cover) {
      const float4 r0 = L.rec[4 * (int64_t)w], r1 = L.rec[4 * (int64_t)w + 1],
                   r2 = L.rec[4 * (int64_t)w + 2];
      const float e0 = edge(r0, px, py), e1 = edge(r1, px, py), e2 = edge(r2, px, py);
      d = win_depth(r0, r1, r2, e0, e1, e2, vp22, vp23) + L.depth_offset;
    } else {
      d = L.zdef + L.depth_offset;
    }
  }
  if (L.texels) L.texels[pix] = d;
  return d;
}

// The shadow shader's depth source in the raster of a call with a shadow
// source: the shadow pass's binning.
struct LookupShadowDepth {
  const ShadowLookup* L;
  // kept texels are read back by the backward: every one a pixel projects to
  // is computed
  JR_DEV bool required() const { return L->texels != nullptr; }
  JR_DEV float operator()(const jr_inputs&, int scene, float x, float y) const {
    return shadow_depth_on_demand(*L, scene, x, y);
  }
};

// Raster instance of phong_reflection_shadow for calls with a shadow source:
// its own template instance, so the instance that reads a finished map does
// not carry the lookup's code and registers.
constexpr int kShaderShadowOnDemand = JR_SHADER_COUNT;
template <>
struct Varyings<kShaderShadowOnDemand> : Varyings<JR_SHADER_PHONG_REFLECTION_SHADOW> {};

// Stage 4: per-tile visibility, shading and merge.
struct RasterArgs {
  jr_inputs in;
  const jr_scene* scenes;
  jr_outputs out;
  Frame f;
  const float4* rec;
  const int32_t* counts;
  const int32_t* offsets;
  const int32_t* bins;
  const uint64_t* keys;
  const uint32_t* ks;    // key state (the narrow form's epoch)
  const float* samples;  // per scene: pix_x(i) for i < W, then pix_y(j)
  const Rect* rects;     // per-triangle tile ranges and binning words (k_setup)
  const uint32_t* masks;
  int32_t* work_counter;
  int32_t chunk;  // tiles claimed per work-queue step (persistent form)
  int32_t xcd_scenes;  // one-group-per-tile form: scenes kept on one XCD each
  int32_t tiles_per_group;  // tiles each group of the non-persistent form takes
  int32_t pixels_per_lane;  // 1, or 4 (16x16 and 8x32 tiles): pixels per thread
  int32_t init_buffers, apply_offset;
  float zdef, cdef0, cdef1, cdef2, depth_offset;
  int32_t shadow_on_demand;  // phong_reflection_shadow: depths from `sl`
  ShadowLookup sl;
};

// Chooser fallback, shading, merge and stores of pixel (i, j) of scene s
// whose running minimum over all candidates is (best_key, best_id).

// The primitive a pixel shows: the chooser's winner, or, when no candidate
// has a finite (or NaN) depth, argmin's index 0 of the scene
// (renderer/shader.py:207-217), whose own keep & in_triangle then decides
// (renderer/pipeline.py:275).  With cover set, also the barycentrics, the
// window depth and the primitive's flags.
struct PixelHit {
  int32_t w;
  bool cover;
  uint32_t fl;
  float z;
  float b[3];
};

JR_DEV PixelHit resolve_pixel(const RasterArgs& a, const jr_scene& sc, float x, float y,
                              uint32_t best_key, int32_t best_id) {
  PixelHit h;
  h.fl = 0u;
  h.z = 0.0f;
  h.b[0] = h.b[1] = h.b[2] = 0.0f;
  if (sc.tri_end <= sc.tri_begin) {
    h.w = -1;  // a scene without primitives draws nothing
    h.cover = false;
  } else if (best_key >= kInfKey) {
    h.w = sc.tri_begin;
    const float4 r0 = a.rec[4 * (int64_t)h.w], r1 = a.rec[4 * (int64_t)h.w + 1],
                 r2 = a.rec[4 * (int64_t)h.w + 2];
    const uint32_t fl = __float_as_uint(a.rec[4 * (int64_t)h.w + 3].y);
    const float e0 = edge(r0, x, y), e1 = edge(r1, x, y), e2 = edge(r2, x, y);
    h.cover = (fl & kKeep) && e0 >= 0.0f && e1 >= 0.0f && e2 >= 0.0f;
  } else {
    h.w = best_id;
    h.cover = true;
  }
  if (h.cover) {
    const float4 r0 = a.rec[4 * (int64_t)h.w], r1 = a.rec[4 * (int64_t)h.w + 1],
                 r2 = a.rec[4 * (int64_t)h.w + 2];
    h.fl = __float_as_uint(a.rec[4 * (int64_t)h.w + 3].y);
    const float e0 = edge(r0, x, y), e1 = edge(r1, x, y), e2 = edge(r2, x, y);
    float wr = e0 + e1;
    wr = wr + e2;
    h.b[0] = e0 / wr;
    h.b[1] = e1 / wr;
    h.b[2] = e2 / wr;
    h.z = win_depth(r0, r1, r2, e0, e1, e2, sc.viewport[10], sc.viewport[11]);
  }
  return h;
}


// merge (renderer/pipeline.py:401-440), the shadow-map offset
// (renderer/shadow.py:116) and the optional id / barycentric / display outputs
JR_DEV void store_pixel(const RasterArgs& a, int32_t s, int32_t i, int32_t j,
                        const PixelHit& h, bool keep, const float col[3]) {
  const Frame& f = a.f;
  const int64_t pix = (int64_t)s * f.W * f.H + (int64_t)i * f.H + j;
  if (keep) {
    a.out.zbuffer[pix] = a.apply_offset ? h.z + a.depth_offset : h.z;
    if (a.out.canvas) {
      a.out.canvas[3 * pix + 0] = col[0];
      a.out.canvas[3 * pix + 1] = col[1];
      a.out.canvas[3 * pix + 2] = col[2];
    }
  } else {
    if (a.init_buffers) {
      a.out.zbuffer[pix] = a.apply_offset ? a.zdef + a.depth_offset : a.zdef;
      if (a.out.canvas) {
        a.out.canvas[3 * pix + 0] = a.cdef0;
        a.out.canvas[3 * pix + 1] = a.cdef1;
        a.out.canvas[3 * pix + 2] = a.cdef2;
      }
    } else if (a.apply_offset) {
      a.out.zbuffer[pix] = a.out.zbuffer[pix] + a.depth_offset;
    }
  }
  if (a.out.display) {
    // the merged canvas value, transposed and flipped for display
    float c0 = col[0], c1 = col[1], c2 = col[2];
    if (!keep) {
      if (a.init_buffers) {
        c0 = a.cdef0;
        c1 = a.cdef1;
        c2 = a.cdef2;
      } else {
        c0 = a.out.canvas[3 * pix + 0];
        c1 = a.out.canvas[3 * pix + 1];
        c2 = a.out.canvas[3 * pix + 2];
      }
    }
    uint8_t* d = a.out.display + 3 * (((int64_t)s * f.H + (f.H - 1 - j)) * f.W + i);
    d[0] = display_u8(c0);
    d[1] = display_u8(c1);
    d[2] = display_u8(c2);
  }
  if (a.out.prim_id) a.out.prim_id[pix] = keep ? h.w : -1;
  if (a.out.bary) {
    a.out.bary[3 * pix + 0] = keep ? h.b[0] : 0.0f;
    a.out.bary[3 * pix + 1] = keep ? h.b[1] : 0.0f;
    a.out.bary[3 * pix + 2] = keep ? h.b[2] : 0.0f;
  }
}

template <int SHADER>
JR_DEV void finish_pixel(const RasterArgs& a, const jr_scene& sc, int32_t s, int32_t i,
                         int32_t j, float x, float y, uint32_t best_key, int32_t best_id) {
  const Frame& f = a.f;
  if (best_id == -1) {
    // narrow form, small winner: its depth is the key's (finite by
    // construction); the depth shader keeps every fragment
    const int64_t pix = (int64_t)s * f.W * f.H + (int64_t)i * f.H + j;
    a.out.zbuffer[pix] = key_depth(best_key) + a.depth_offset;
    return;
  }
  if ((SHADER == JR_SHADER_DEPTH || SHADER == JR_SHADER_VISIBILITY) && !a.out.bary &&
      best_key < kInfKey && best_key != 0u && best_key != 0x80000000u) {
    // a depth-only winner whose key is neither NaN nor zero gives its depth
    // back exactly; no barycentrics are needed, and these shaders keep every
    // fragment
    const int64_t pix = (int64_t)s * f.W * f.H + (int64_t)i * f.H + j;
    const float z = key_depth(best_key);
    a.out.zbuffer[pix] = a.apply_offset ? z + a.depth_offset : z;
    if (a.out.prim_id) a.out.prim_id[pix] = best_id;
    return;
  }
  const PixelHit h = resolve_pixel(a, sc, x, y, best_key, best_id);
  bool keep = false;
  float col[3] = {0.0f, 0.0f, 0.0f};
  if (h.cover)
    keep = Shade<SHADER>::run(a.in, sc, corner_slots(a.in, h.w), h.b, (h.fl & kFront) != 0,
                              col, s);
  store_pixel(a, s, i, j, h, keep, col);
}


// Finish of a tile for a shader with a per-triangle stage: every distinct
// covered primitive of the tile gets its corners' vertex-stage outputs
// computed once (by one thread, into LDS: the triangle records' storage,
// free once visibility is decided), and its pixels read them from there.
// Pixels of primitives beyond the table's capacity compute their own.
template <int SHADER, int NT, int PPL, bool WAVE>
JR_DEV void finish_tile_shared(const RasterArgs& a, const jr_scene& sc, int32_t s, int tid,
                               const int32_t* pi, const int32_t* pj, const float* px,
                               const float* py, const bool* inside, const uint32_t* best_key,
                               const int32_t* best_id, float4* lds_rec, int32_t* lds_id) {
  using V = Varyings<SHADER>;
  constexpr int NS = V::N | 1;  // odd stride: no LDS bank conflicts
  constexpr int CAP0 = (48 * NT) / (4 * NS);
  constexpr int CAP = CAP0 < (NT - 1) / 2 ? CAP0 : (NT - 1) / 2;
  float* vary = reinterpret_cast<float*>(lds_rec);
  int32_t* tab = lds_id;
  int32_t* dense = lds_id + CAP;        // the occupied entries of tab, packed
  int32_t* n_dense = lds_id + NT - 1;   // their count
  PixelHit h[PPL];
#pragma unroll
  for (int q = 0; q < PPL; ++q) {
    if (inside[q]) {
      h[q] = resolve_pixel(a, sc, px[q], py[q], best_key[q], best_id[q]);
    } else {
      h[q].w = -1;
      h[q].cover = false;
    }
  }
  tile_sync<WAVE>();  // the visibility loop's reads of the records are done
  for (int e = tid; e < CAP; e += NT) tab[e] = -1;
  if (tid == 0) *n_dense = 0;
  tile_sync<WAVE>();
  int32_t slot[PPL];
#pragma unroll
  for (int q = 0; q < PPL; ++q) slot[q] = h[q].cover ? table_insert(tab, CAP, h[q].w) : -1;
  tile_sync<WAVE>();
  for (int e = tid; e < CAP; e += NT)
    if (tab[e] >= 0) dense[atomicAdd(n_dense, 1)] = e;
  tile_sync<WAVE>();
  // one lane per corner of each distinct winning triangle
  const int32_t nd = *n_dense;
  for (int32_t idx = tid; idx < 3 * nd; idx += NT) {
    const int32_t d = idx / 3, k = idx - 3 * d;
    const int32_t e = dense[d];
    V::corner(a.in, sc, corner_slots(a.in, tab[e]), k, vary + e * NS);
  }
  tile_sync<WAVE>();
  float l[3];
  V::scene_terms(sc, l);
#pragma unroll
  for (int q = 0; q < PPL; ++q) {
    if (!inside[q]) continue;
    bool keep = false;
    float col[3] = {0.0f, 0.0f, 0.0f};
    if (h[q].cover) {
      float v[V::N];
      if (slot[q] >= 0) {
#pragma unroll
        for (int k = 0; k < V::N; ++k) v[k] = vary[slot[q] * NS + k];
      } else {
        V::corners(a.in, sc, corner_slots(a.in, h[q].w), v);
      }
      if constexpr (SHADER == kShaderShadowOnDemand) {
        const LookupShadowDepth depth{&a.sl};
        keep = V::pixel(a.in, sc, v, l, h[q].b, (h[q].fl & kFront) != 0, col, s, depth);
      } else {
        keep = V::pixel(a.in, sc, v, l, h[q].b, (h[q].fl & kFront) != 0, col, s);
      }
    }
    store_pixel(a, s, pi[q], pj[q], h[q], keep, col);
  }
}

// One screen tile of TW x TH pixels with TW * TH / PPL threads: thread tid
// owns the pixels tid + q * (TW * TH / PPL), q < PPL.  The tile's triangle
// list is staged through LDS once and tested against all of a thread's pixels.
template <int SHADER, int TW, int TH, int PPL, bool WAVE = false>
JR_DEV void raster_tile(const RasterArgs& a, int64_t gt, float4* lds_rec,
                        int32_t* lds_id) {
  constexpr int NT = TW * TH / PPL;
  static_assert(!WAVE || NT == 64, "a wavefront-owned tile has one lane per 4 pixels");
  const Frame& f = a.f;
  const int32_t s = udiv31((int32_t)gt, f.T_m, f.T_sh);
  const int32_t tile = (int32_t)gt - s * f.T;
  const int32_t tx = udiv31(tile, f.ty_m, f.ty_sh), ty = tile - tx * f.tiles_y;
  const int tid = WAVE ? (int)(threadIdx.x & 63) : (int)threadIdx.x;
  const jr_scene& sc = a.scenes[s];
  const float* vp = sc.viewport;
  const float vp22 = vp[10], vp23 = vp[11];
  int32_t pi[PPL], pj[PPL];
  float px[PPL], py[PPL];
  bool inside[PPL];
  uint32_t best_key[PPL];
  int32_t best_id[PPL];
  // with NT a multiple of TH all of a thread's pixels share one j (and y)
  constexpr bool kOneRow = (NT % TH) == 0;
  const float* sx = a.samples + (int64_t)s * (f.W + f.H);
  const uint32_t ktag = f.narrow ? ~a.ks[kKsEpoch] : 0u;
#pragma unroll
  for (int q = 0; q < PPL; ++q) {
    const int p = tid + q * NT;
    pi[q] = tx * TW + p / TH;
    pj[q] = ty * TH + (kOneRow ? tid % TH : p % TH);
    inside[q] = (pi[q] < f.W) && (pj[q] < f.H);
    px[q] = inside[q] ? sx[pi[q]] : 0.0f;
    py[q] = inside[q] ? sx[f.W + pj[q]] : 0.0f;
    // start from the small triangles already resolved by k_setup (in the
    // narrow form their id is not kept: -1, which no equal key displaces)
    best_key[q] = 0xFFFFFFFFu;
    best_id[q] = 0x7FFFFFFF;
    if (inside[q]) {
      const int64_t pix = (int64_t)s * f.W * f.H + (int64_t)pi[q] * f.H + pj[q];
      if (f.narrow) {
        const uint32_t k0 = untag_key(a.keys[pix], ktag);
        if (k0 != ~0u) {
          best_key[q] = k0;
          best_id[q] = -1;
        }
      } else {
        const uint64_t k0 = a.keys[pix];
        if (k0 != ~0ull) {
          best_key[q] = (uint32_t)(k0 >> 32);
          best_id[q] = (int32_t)(uint32_t)k0;
        }
      }
    }
  }

  const int32_t cnt = a.counts[gt];
  const int64_t start = a.offsets[gt];
  const bool overflow = start + cnt > f.capacity;
  // overflowing tile lists fall back to every triangle of the scene
  const int32_t n_items = overflow ? (sc.tri_end - sc.tri_begin) : cnt;
  for (int32_t base = 0; base < n_items; base += NT) {
    const int32_t n = min(NT, n_items - base);
    tile_sync<WAVE>();
    if (tid < n) {
      // binned triangles are kept and front-facing; the scan of an
      // overflowing tile takes the scene's triangles that list this tile (the
      // small ones are already in the key buffer, the culled ones never win)
      const int32_t t = overflow ? (sc.tri_begin + base + tid) : a.bins[start + base + tid];
      const bool use = !overflow || lists_tile(a.rects, a.masks, t, tx, ty);
      if (use) {
        lds_rec[3 * tid + 0] = a.rec[4 * (int64_t)t];
        lds_rec[3 * tid + 1] = a.rec[4 * (int64_t)t + 1];
        lds_rec[3 * tid + 2] = a.rec[4 * (int64_t)t + 2];
      }
      lds_id[tid] = use ? t : -1;
    }
    tile_sync<WAVE>();
    for (int m = 0; m < n; ++m) {
      const int32_t t = __builtin_amdgcn_readfirstlane(lds_id[m]);
      if (t < 0) continue;
      const float4 r0 = lds_rec[3 * m + 0], r1 = lds_rec[3 * m + 1],
                   r2 = lds_rec[3 * m + 2];
      // y * b_k is one rounded product, shared by pixels of the same row
      const float yb0 = py[0] * r0.y, yb1 = py[0] * r1.y, yb2 = py[0] * r2.y;
#pragma unroll
      for (int q = 0; q < PPL; ++q) {
        float e0, e1, e2;
        if (kOneRow) {
          e0 = px[q] * r0.x;
          e0 = e0 + yb0;
          e0 = e0 + r0.z;
          e1 = px[q] * r1.x;
          e1 = e1 + yb1;
          e1 = e1 + r1.z;
          e2 = px[q] * r2.x;
          e2 = e2 + yb2;
          e2 = e2 + r2.z;
        } else {
          e0 = edge(r0, px[q], py[q]);
          e1 = edge(r1, px[q], py[q]);
          e2 = edge(r2, px[q], py[q]);
        }
        if (e0 >= 0.0f && e1 >= 0.0f && e2 >= 0.0f) {
          const uint32_t key = depth_key(win_depth(r0, r1, r2, e0, e1, e2, vp22, vp23));
          // a 64-bit (key, id) comparison instead is slower here
          if (key < best_key[q] || (key == best_key[q] && t < best_id[q])) {
            best_key[q] = key;
            best_id[q] = t;
          }
        }
      }
    }
  }
  if constexpr (Varyings<SHADER>::N > 0) {
    finish_tile_shared<SHADER, NT, PPL, WAVE>(a, sc, s, tid, pi, pj, px, py, inside, best_key,
                                              best_id, lds_rec, lds_id);
  } else {
#pragma unroll
    for (int q = 0; q < PPL; ++q)
      if (inside[q])
        finish_pixel<SHADER>(a, sc, s, pi[q], pj[q], px[q], py[q], best_key[q], best_id[q]);
  }
}

// Workgroups are dispatched to the 8 XCDs round-robin by group id, and each
// XCD has its own L2.  With many scenes, group b takes a tile of a scene
// s == b (mod 8), so every tile of a scene runs on one XCD and the scene's
// triangle records and attributes are fetched into one L2 only.  The grid is
// 8 * ceil(n_scenes / 8) * T groups; the padding groups exit at once.
constexpr int kXcds = 8;

// (scene, tile) item of the k-th tile of XCD x's list: the tiles of scenes
// x, x + 8, x + 16, ... in order (-1 past its end).
JR_DEV int64_t xcd_item(const RasterArgs& a, int32_t x, int64_t k) {
  const int64_t n_x = (int64_t)((a.f.n_scenes - x + kXcds - 1) / kXcds) * a.f.T;
  if (k >= n_x) return -1;
  const int64_t sl = udiv31((int32_t)k, a.f.T_m, a.f.T_sh);
  return (sl * kXcds + x) * a.f.T + (k - sl * a.f.T);
}

// At most 72 registers (7 wavefronts per SIMD): the shading instances would
// otherwise take 95 and run 5 per SIMD; the occupancy pays for the few spills.
template <int SHADER, int TW, int TH, int PPL>
__global__ __launch_bounds__(TW* TH / PPL) __attribute__((amdgpu_waves_per_eu(7, 8)))
void k_raster(RasterArgs a) {
  __shared__ float4 lds_rec[3 * TW * TH / PPL];
  __shared__ int32_t lds_id[TW * TH / PPL];
  int64_t gt = blockIdx.x;
  if (a.xcd_scenes) {
    gt = xcd_item(a, (int32_t)(blockIdx.x % kXcds), blockIdx.x / kXcds);
    if (gt < 0) return;
  }
  raster_tile<SHADER, TW, TH, PPL>(a, gt, lds_rec, lds_id);
}

// Several consecutive 256-pixel tiles per workgroup, one per wavefront, side
// by side (four pixels per lane).
template <int SHADER, int TW, int TH, int PPL>
__global__ __launch_bounds__(256) void k_raster_waves(RasterArgs a) {
  __shared__ float4 lds_rec[4][3 * 64];
  __shared__ int32_t lds_id[4][64];
  const int w = threadIdx.x >> 6;
  const int64_t b = blockIdx.x;
  int64_t gt;
  if (a.xcd_scenes) {
    gt = xcd_item(a, (int32_t)(b % kXcds), (b / kXcds) * a.tiles_per_group + w);
    if (gt < 0) return;
  } else {
    gt = b * a.tiles_per_group + w;
    if (gt >= (int64_t)a.f.n_scenes * a.f.T) return;
  }
  raster_tile<SHADER, TW, TH, PPL, true>(a, gt, lds_rec[w], lds_id[w]);
}

// Several consecutive tiles per workgroup (a separate kernel: the loop raises
// the register count of the single-tile form).
template <int SHADER, int TW, int TH, int PPL>
__global__ __launch_bounds__(TW* TH / PPL) void k_raster_multi(RasterArgs a) {
  __shared__ float4 lds_rec[3 * TW * TH / PPL];
  __shared__ int32_t lds_id[TW * TH / PPL];
  const int64_t b = blockIdx.x;
  const int64_t n_items = (int64_t)a.f.n_scenes * a.f.T;
  for (int32_t q = 0; q < a.tiles_per_group; ++q) {
    int64_t gt;
    if (a.xcd_scenes) {
      gt = xcd_item(a, (int32_t)(b % kXcds), (b / kXcds) * a.tiles_per_group + q);
      if (gt < 0) break;
    } else {
      gt = b * a.tiles_per_group + q;
      if (gt >= n_items) break;
    }
    raster_tile<SHADER, TW, TH, PPL>(a, gt, lds_rec, lds_id);
  }
}

// Persistent form: a fixed grid drains the (scene, tile) queue.  Every group
// exits once the shared counter passes the item count.
template <int SHADER, int TW, int TH, int PPL>
__global__ __launch_bounds__(TW* TH / PPL) void k_raster_persistent(RasterArgs a,
                                                                    int64_t n_items) {
  __shared__ float4 lds_rec[3 * TW * TH / PPL];
  __shared__ int32_t lds_id[TW * TH / PPL];
  __shared__ int32_t item;
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) item = atomicAdd(a.work_counter, a.chunk);
    __syncthreads();
    const int32_t first = item;
    if (first >= n_items) break;
    const int32_t last = (int32_t)min((int64_t)first + a.chunk, n_items);
    for (int32_t gt = first; gt < last; ++gt) raster_tile<SHADER, TW, TH, PPL>(a, gt, lds_rec, lds_id);
  }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
thread_local std::string g_err;

// Optional per-stage timing: when enabled, every jr_forward records five HIP
// events on its own stream (start, after setup, after scan, after binning,
// after raster); jr_stage_times reads them back.
std::mutex g_timing_mu;
bool g_timing = false;
std::vector<std::array<hipEvent_t, JR_STAGES + 1>> g_timed;

struct StageTimer {
  hipStream_t st;
  std::array<hipEvent_t, JR_STAGES + 1> ev{};
  bool on = false;
  // only calls that run every stage are timed
  StageTimer(hipStream_t s, bool whole) : st(s) {
    std::lock_guard<std::mutex> g(g_timing_mu);
    on = g_timing && whole;
    if (!on) return;
    for (auto& e : ev)
      if (hipEventCreate(&e) != hipSuccess) e = nullptr;
    mark(0);
  }
  void mark(int i) {
    if (on && ev[i]) (void)hipEventRecord(ev[i], st);
  }
  ~StageTimer() {
    if (!on) return;
    std::lock_guard<std::mutex> g(g_timing_mu);
    g_timed.push_back(ev);
  }
};

int32_t fail(int32_t code, const std::string& msg) {
  g_err = msg;
  return code;
}

constexpr int64_t kAlign = 256;
int64_t align_up(int64_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

struct WsLayout {
  int64_t rec, rects, counts, cursor, offsets, block_sums, misc, keys, kstate, wide, samples,
      masks, wave_scene, wbits, bins, blist, total;
  int64_t n_tiles, nblk, n_pix;
};


int32_t log2_exact(int32_t v) {
  int32_t l = 0;
  while ((1 << l) < v) ++l;
  return l;
}

// The kernels' view of a frame; `narrow`: the depth-only form whose
// small-triangle keys stay in the zbuffer.
Frame dev_frame(const jr_frame* fr, bool narrow) {
  Frame f;
  f.W = fr->width;
  f.H = fr->height;
  f.TW = fr->tile_w;
  f.TH = fr->tile_h;
  f.tiles_x = (f.W + f.TW - 1) / f.TW;
  f.tiles_y = (f.H + f.TH - 1) / f.TH;
  f.T = f.tiles_x * f.tiles_y;
  f.tw_sh = log2_exact(f.TW);
  f.th_sh = log2_exact(f.TH);
  magic31((uint32_t)f.T, &f.T_m, &f.T_sh);
  magic31((uint32_t)f.tiles_y, &f.ty_m, &f.ty_sh);
  f.n_scenes = fr->n_scenes;
  f.capacity = fr->bin_capacity;
  f.narrow = narrow;
  return f;
}

// A depth pass with an offset and no id outputs keeps 32-bit depth keys in
// its zbuffer (an offset of -0 would keep the sign of a -0 depth, which a key
// cannot tell apart).
bool narrow_form(const jr_frame* fr, bool ids) {
  uint32_t ob;
  std::memcpy(&ob, &fr->depth_offset, 4);
  return fr->shader == JR_SHADER_DEPTH && !ids && fr->apply_offset && fr->init_buffers &&
         ob != 0x80000000u;
}

WsLayout ws_layout(const jr_frame* fr, int32_t n_tris) {
  WsLayout L{};
  const int64_t tiles_x = (fr->width + fr->tile_w - 1) / fr->tile_w;
  const int64_t tiles_y = (fr->height + fr->tile_h - 1) / fr->tile_h;
  L.n_tiles = tiles_x * tiles_y * (int64_t)fr->n_scenes;
  L.nblk = (L.n_tiles + kScanBlock - 1) / kScanBlock;
  int64_t o = 0;
  L.kstate = o;     o += align_up(64);   // key state (k_keys_prepare): first in every layout
  L.rec = o;        o += align_up(64 * (int64_t)std::max(n_tris, 1));
  L.rects = o;      o += align_up(16 * (int64_t)std::max(n_tris, 1));
  L.counts = o;     o += align_up(4 * L.n_tiles);
  L.cursor = o;     o += align_up(4 * L.n_tiles);
  L.misc = o;       o += align_up(256);  // [0] total pairs, [1] work counter, [2] wide count,
                                         // [2 + kBinListAt] k_bin's list length
  L.offsets = o;    o += align_up(4 * (L.n_tiles + 1));
  L.block_sums = o; o += align_up(4 * std::max<int64_t>(L.nblk, 1));
  L.n_pix = (int64_t)fr->width * fr->height * fr->n_scenes;
  L.keys = o;       o += align_up(8 * L.n_pix);
  L.wide = o;       o += align_up(8 * (int64_t)std::max(n_tris, 1));
  L.samples = o;    o += align_up(4 * (int64_t)(fr->width + fr->height) * fr->n_scenes);
  L.masks = o;      o += align_up(4 * (int64_t)std::max(n_tris, 1));
  L.wave_scene = o; o += align_up(4 * (((int64_t)std::max(n_tris, 1) + 63) / 64));
  L.wbits = o;      o += align_up(4 * (int64_t)kWideBitWords);
  L.bins = o;       o += align_up(4 * std::max<int64_t>(fr->bin_capacity, 1));
  L.blist = o;      o += align_up(8 * (int64_t)std::max(n_tris, 1));
  L.total = o;
  return L;
}

int32_t check_frame(const jr_frame* fr) {
  if (!fr) return fail(JR_E_INVALID, "frame is NULL");
  if (fr->width <= 0 || fr->height <= 0 || fr->n_scenes <= 0)
    return fail(JR_E_INVALID, "width, height and n_scenes must be positive");
  if (fr->shader < 0 || fr->shader >= JR_SHADER_COUNT)
    return fail(JR_E_INVALID, "unknown shader id");
  const bool ok_tile = (fr->tile_w == 8 && fr->tile_h == 8) ||
                       (fr->tile_w == 16 && fr->tile_h == 16) ||
                       (fr->tile_w == 8 && fr->tile_h == 32) ||
                       (fr->tile_w == 4 && fr->tile_h == 16);
  if (!ok_tile) return fail(JR_E_INVALID, "unsupported tile shape");
  const int64_t n_tiles = (int64_t)((fr->width + fr->tile_w - 1) / fr->tile_w) *
                          ((fr->height + fr->tile_h - 1) / fr->tile_h) * fr->n_scenes;
  if (n_tiles > 0x7FFFFFFFLL) return fail(JR_E_INVALID, "more than 2^31 - 1 tiles");
  if (n_tiles / fr->n_scenes >= (1LL << 22))
    return fail(JR_E_INVALID, "2^22 or more tiles per scene");
  if (fr->bin_capacity < 0 || fr->bin_capacity > 0x7FFFFFFFLL)
    return fail(JR_E_INVALID, "bin_capacity out of range");
  if (fr->part != JR_PART_ALL && fr->part != JR_PART_BIN && fr->part != JR_PART_RASTER)
    return fail(JR_E_INVALID, "part must be JR_PART_ALL, JR_PART_BIN or JR_PART_RASTER");
  if (fr->tiles_per_group < 0 || fr->tiles_per_group > 64)
    return fail(JR_E_INVALID, "tiles_per_group must be in [0, 64]");
  const bool ok_ppl = fr->pixels_per_lane <= 1 ||
                      (fr->pixels_per_lane == 4 && fr->tile_w * fr->tile_h == 256);
  if (!ok_ppl)
    return fail(JR_E_INVALID, "pixels_per_lane must be 0/1, or 4 with 16x16 or 8x32 tiles");
  return JR_OK;
}

template <int SHADER, int TW, int TH, int PPL>
void launch_raster_t(const RasterArgs& a, int64_t n_items, bool persistent,
                     hipStream_t st) {
  constexpr int NT = TW * TH / PPL;
  if (persistent) {
    int dev = 0, cus = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    const int per_cu = 2048 / NT;  // up to 32 waves per CU
    const int64_t grid = std::min<int64_t>(n_items, (int64_t)std::max(cus, 1) * per_cu);
    // claim tiles in chunks: one queue atomic per chunk, about eight chunks
    // per group
    RasterArgs b = a;
    b.chunk = (int32_t)std::max<int64_t>(1, std::min<int64_t>(32, n_items / (grid * 8)));
    hipLaunchKernelGGL((k_raster_persistent<SHADER, TW, TH, PPL>), dim3((unsigned)grid),
                       dim3(NT), 0, st, b, n_items);
  } else {
    RasterArgs b = a;
    const int64_t tpg = std::max(1, a.tiles_per_group);
    b.tiles_per_group = (int32_t)tpg;
    int64_t grid = (n_items + tpg - 1) / tpg;
    b.xcd_scenes = a.f.n_scenes >= 8 * kXcds;
    if (b.xcd_scenes) {
      const int64_t per_xcd = (int64_t)((a.f.n_scenes + kXcds - 1) / kXcds) * a.f.T;
      grid = kXcds * ((per_xcd + tpg - 1) / tpg);
    }
    if (tpg == 1) {
      hipLaunchKernelGGL((k_raster<SHADER, TW, TH, PPL>), dim3((unsigned)grid), dim3(NT), 0,
                         st, b);
      return;
    }
    if constexpr (NT == 64) {
      if (tpg <= 4) {
        hipLaunchKernelGGL((k_raster_waves<SHADER, TW, TH, PPL>), dim3((unsigned)grid),
                           dim3((unsigned)(64 * tpg)), 0, st, b);
        return;
      }
    }
    hipLaunchKernelGGL((k_raster_multi<SHADER, TW, TH, PPL>), dim3((unsigned)grid),
                       dim3(NT), 0, st, b);
  }
}

template <int SHADER>
void launch_raster_s(const RasterArgs& a, int64_t n_items, bool persistent,
                     hipStream_t st) {
  const int tw = a.f.TW, th = a.f.TH, ppl = a.pixels_per_lane;
  if (tw == 8 && th == 8) launch_raster_t<SHADER, 8, 8, 1>(a, n_items, persistent, st);
  else if (tw == 16 && th == 16 && ppl == 4)
    launch_raster_t<SHADER, 16, 16, 4>(a, n_items, persistent, st);
  else if (tw == 16 && th == 16) launch_raster_t<SHADER, 16, 16, 1>(a, n_items, persistent, st);
  else if (tw == 8 && th == 32 && ppl == 4)
    launch_raster_t<SHADER, 8, 32, 4>(a, n_items, persistent, st);
  else if (tw == 8 && th == 32) launch_raster_t<SHADER, 8, 32, 1>(a, n_items, persistent, st);
  else launch_raster_t<SHADER, 4, 16, 1>(a, n_items, persistent, st);
}

void launch_raster(int shader, const RasterArgs& a, int64_t n_items,
                   bool persistent, hipStream_t st) {
  switch (shader) {
    case JR_SHADER_DEPTH: launch_raster_s<JR_SHADER_DEPTH>(a, n_items, persistent, st); break;
    case JR_SHADER_GOURAUD: launch_raster_s<JR_SHADER_GOURAUD>(a, n_items, persistent, st); break;
    case JR_SHADER_GOURAUD_TEXTURE: launch_raster_s<JR_SHADER_GOURAUD_TEXTURE>(a, n_items, persistent, st); break;
    case JR_SHADER_PHONG: launch_raster_s<JR_SHADER_PHONG>(a, n_items, persistent, st); break;
    case JR_SHADER_PHONG_DARBOUX: launch_raster_s<JR_SHADER_PHONG_DARBOUX>(a, n_items, persistent, st); break;
    case JR_SHADER_PHONG_REFLECTION: launch_raster_s<JR_SHADER_PHONG_REFLECTION>(a, n_items, persistent, st); break;
    case JR_SHADER_PHONG_REFLECTION_SHADOW:
      if (a.shadow_on_demand)
        launch_raster_s<kShaderShadowOnDemand>(a, n_items, persistent, st);
      else
        launch_raster_s<JR_SHADER_PHONG_REFLECTION_SHADOW>(a, n_items, persistent, st);
      break;
    default: launch_raster_s<JR_SHADER_VISIBILITY>(a, n_items, persistent, st); break;
  }
}

// A shadow source must describe a depth pass in the narrow form, of this
// call's scenes and shadow-map size, whose workspace holds its binning.
int32_t check_shadow_source(const jr_frame* fr, const jr_inputs* in) {
  const jr_shadow_source* src = in->shadow_source;
  if (fr->shader != JR_SHADER_PHONG_REFLECTION_SHADOW)
    return fail(JR_E_INVALID, "shadow_source is read only by phong_reflection_shadow");
  if (!src->frame || !src->scenes || !src->workspace)
    return fail(JR_E_INVALID, "shadow_source: frame, scenes and workspace are required");
  const jr_frame* sf = src->frame;
  int32_t rc = check_frame(sf);
  if (rc) return rc;
  if (!narrow_form(sf, false))
    return fail(JR_E_INVALID, "shadow_source: the shadow pass must be a depth pass with "
                              "init_buffers, apply_offset and an offset other than -0");
  if (sf->n_scenes != fr->n_scenes || sf->width != in->shadow_w || sf->height != in->shadow_h)
    return fail(JR_E_INVALID, "shadow_source: scene count or map size differs from this call");
  if (src->n_tris < 0 || src->workspace_bytes < ws_layout(sf, src->n_tris).total)
    return fail(JR_E_WORKSPACE, "shadow_source: workspace smaller than its frame needs");
  return JR_OK;
}

int32_t check_inputs(const jr_frame* fr, const jr_inputs* in, const jr_outputs* out) {
  if (!in || !out) return fail(JR_E_INVALID, "inputs/outputs NULL");
  if (in->n_tris < 0 || in->n_slots < 0) return fail(JR_E_INVALID, "negative sizes");
  if (!out->zbuffer) return fail(JR_E_INVALID, "zbuffer is required");
  if (in->n_tris == 0) return JR_OK;
  if (!in->positions && !in->clip) return fail(JR_E_INVALID, "positions (or clip) are required");
  if (!in->faces && (int64_t)in->n_slots < 3 * (int64_t)in->n_tris)
    return fail(JR_E_INVALID, "without faces, n_slots must be at least 3 n_tris");
  if (in->clip && !(fr->shader == JR_SHADER_VISIBILITY || fr->shader == JR_SHADER_DEPTH))
    return fail(JR_E_INVALID, "clip input is only valid with the depth/visibility shaders");
  const int sh = fr->shader;
  const bool needs_canvas = !(sh == JR_SHADER_DEPTH || sh == JR_SHADER_VISIBILITY);
  if (needs_canvas && !out->canvas) return fail(JR_E_INVALID, "shader writes a canvas but canvas is NULL");
  if (out->display && !needs_canvas)
    return fail(JR_E_INVALID, "display output needs a shader that writes a canvas");
  if (sh == JR_SHADER_GOURAUD && (!in->normals || !in->colours))
    return fail(JR_E_INVALID, "gouraud needs normals and colours");
  if ((sh == JR_SHADER_GOURAUD_TEXTURE || sh == JR_SHADER_PHONG || sh == JR_SHADER_PHONG_DARBOUX ||
       sh == JR_SHADER_PHONG_REFLECTION || sh == JR_SHADER_PHONG_REFLECTION_SHADOW) &&
      (!in->normals || !in->uvs || !in->textures))
    return fail(JR_E_INVALID, "textured shaders need normals, uvs and textures");
  if ((sh == JR_SHADER_PHONG_REFLECTION || sh == JR_SHADER_PHONG_REFLECTION_SHADOW) &&
      (!in->tex_index || !in->texture_shapes || !in->specular_maps))
    return fail(JR_E_INVALID, "phong_reflection needs tex_index, texture_shapes, specular_maps");
  if (sh == JR_SHADER_PHONG_DARBOUX && (!in->id_to_face || !in->face_slots || !in->normal_maps))
    return fail(JR_E_INVALID, "phong_darboux needs id_to_face, face_slots, normal_maps");
  if (sh == JR_SHADER_PHONG_REFLECTION_SHADOW &&
      ((!in->shadow_maps && !in->shadow_source) || in->shadow_w <= 0 || in->shadow_h <= 0))
    return fail(JR_E_INVALID, "phong_reflection_shadow needs shadow maps");
  return in->shadow_source ? check_shadow_source(fr, in) : JR_OK;
}

}  // namespace jr

using namespace jr;

extern "C" {

int32_t jr_version(void) { return JR_ABI_VERSION; }

const char* jr_last_error(void) { return g_err.c_str(); }

int64_t jr_sizeof(int32_t which) {
  switch (which) {
    case 0: return sizeof(jr_scene);
    case 1: return sizeof(jr_inputs);
    case 2: return sizeof(jr_outputs);
    case 3: return sizeof(jr_frame);
    case 4: return sizeof(jr_grads);
    case 5: return sizeof(jr_pose_args);
    case 6: return sizeof(jr_shadow_source);
    case 7: return sizeof(jr_pack_args);
    default: return -1;
  }
}

int64_t jr_workspace_size(const jr_frame* frame, int32_t n_tris) {
  if (check_frame(frame) != JR_OK || n_tris < 0) return -1;
  return ws_layout(frame, n_tris).total;
}

}  // extern "C"

namespace jr {

// The body of jr_forward and jr_forward_noalloc: validated arguments, a
// workspace of at least ws_layout(fr, n_tris).total bytes.  `timed`: record
// the stage events when jr_stage_timing is on (creating events allocates, so
// the no-allocation entry never times).
int32_t forward_body(const jr_frame* fr, const jr_inputs* in, const jr_scene* scenes,
                     jr_outputs* out, char* ws, const WsLayout& L, hipStream_t st,
                     bool timed) {
  const Frame f = dev_frame(fr, narrow_form(fr, out->prim_id || out->bary));

  float4* rec = reinterpret_cast<float4*>(ws + L.rec);
  Rect* rects = reinterpret_cast<Rect*>(ws + L.rects);
  int32_t* counts = reinterpret_cast<int32_t*>(ws + L.counts);
  int32_t* cursor = reinterpret_cast<int32_t*>(ws + L.cursor);
  int32_t* misc = reinterpret_cast<int32_t*>(ws + L.misc);
  int32_t* offsets = reinterpret_cast<int32_t*>(ws + L.offsets);
  int32_t* block_sums = reinterpret_cast<int32_t*>(ws + L.block_sums);
  int32_t* bins = reinterpret_cast<int32_t*>(ws + L.bins);
  uint64_t* keys = reinterpret_cast<uint64_t*>(ws + L.keys);
  uint32_t* kstate = reinterpret_cast<uint32_t*>(ws + L.kstate);
  int2* wide = reinterpret_cast<int2*>(ws + L.wide);
  float* samples = reinterpret_cast<float*>(ws + L.samples);
  uint32_t* masks = reinterpret_cast<uint32_t*>(ws + L.masks);
  int32_t* wave_scene = reinterpret_cast<int32_t*>(ws + L.wave_scene);
  int32_t* wide_count = misc + 2;
  uint32_t* wbits = reinterpret_cast<uint32_t*>(ws + L.wbits);
  int2* blist = reinterpret_cast<int2*>(ws + L.blist);
  const unsigned wide_grid = 8192;  // 32768 wavefronts stride over the wide list

  // a binning-only call is timed too (its raster stage then takes no time)
  StageTimer timer(st, timed && fr->part != JR_PART_RASTER);
  const int n_tris = in->n_tris;
  if (fr->part != JR_PART_RASTER) {
  // counts, cursor and misc are contiguous: clear them, and the key buffer,
  // with one fill kernel (a kernel rather than hipMemsetAsync so the sequence
  // can be graph-captured); the narrow form's tagged keys are filled only
  // when their state asks for it
  const int64_t nz = (L.offsets - L.counts) / 4;
  // (the narrow form fills its keys only now and then: a smaller grid)
  const int64_t nclr = std::max(std::max(nz, f.narrow ? std::min<int64_t>(L.n_pix, 1 << 20)
                                                      : L.n_pix),
                                (int64_t)(f.W + f.H) * f.n_scenes);
  if (f.narrow)
    hipLaunchKernelGGL(k_keys_prepare, dim3(1), dim3(64), 0, st, kstate, L.n_pix, L.keys);
  hipLaunchKernelGGL(k_clear, dim3((unsigned)std::min<int64_t>((nclr + 255) / 256, 8192)),
                     dim3(256), 0, st, counts, nz, keys, L.n_pix, kstate, scenes, f, samples,
                     wave_scene, in->n_tris);
  timer.mark(1);
  if (n_tris > 0)
    hipLaunchKernelGGL(k_setup, dim3((n_tris + kSetupThreads - 1) / kSetupThreads),
                       dim3(kSetupThreads), 0, st, *in,
                       scenes, f, rec, rects, counts, keys, wide, wide_count, samples, kstate,
                       masks, wave_scene, blist);
  timer.mark(2);
  if (n_tris > 0)
    hipLaunchKernelGGL(k_wide<false>, dim3(wide_grid), dim3(256), 0, st, scenes, f, rec,
                       rects, wide, wide_count, counts, offsets, cursor, bins, samples, wbits);
  timer.mark(3);
  hipLaunchKernelGGL(k_scan_local, dim3((unsigned)std::max<int64_t>(L.nblk, 1)),
                     dim3(kScanThreads), 0, st, counts, offsets, block_sums, L.n_tiles);
  hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kScanThreads), 0, st, block_sums,
                     L.nblk, misc);
  hipLaunchKernelGGL(k_scan_add, dim3((unsigned)std::max<int64_t>(L.nblk, 1)),
                     dim3(kScanThreads), 0, st, offsets, block_sums, L.n_tiles);
  timer.mark(4);
  if (n_tris > 0) {
    hipLaunchKernelGGL(k_bin, dim3((unsigned)std::min((n_tris + 255) / 256, 2048)), dim3(256), 0,
                       st, f, rects, offsets, cursor, bins, masks, blist,
                       wide_count + kBinListAt);
    hipLaunchKernelGGL(k_wide<true>, dim3(wide_grid), dim3(256), 0, st, scenes, f, rec,
                       rects, wide, wide_count, counts, offsets, cursor, bins, samples, wbits);
  }
  timer.mark(5);
  }
  if (fr->part == JR_PART_BIN) {
    timer.mark(6);
    const hipError_t err = hipGetLastError();
    if (err != hipSuccess)
      return fail(JR_E_HIP, std::string("launch failed: ") + hipGetErrorString(err));
    return JR_OK;
  }
  RasterArgs a;
  a.in = *in;
  a.scenes = scenes;
  a.out = *out;
  a.f = f;
  a.rec = rec;
  a.counts = counts;
  a.offsets = offsets;
  a.bins = bins;
  a.rects = rects;
  a.masks = masks;
  a.keys = keys;
  a.ks = kstate;
  a.samples = samples;
  a.work_counter = misc + 1;
  a.chunk = 1;
  a.xcd_scenes = 0;
  a.tiles_per_group = fr->tiles_per_group;
  a.pixels_per_lane = fr->pixels_per_lane > 1 ? fr->pixels_per_lane : 1;
  a.init_buffers = fr->init_buffers;
  a.apply_offset = fr->apply_offset;
  a.zdef = fr->zbuffer_default;
  a.cdef0 = fr->colour_default[0];
  a.cdef1 = fr->colour_default[1];
  a.cdef2 = fr->colour_default[2];
  a.depth_offset = fr->depth_offset;
  a.shadow_on_demand = 0;
  if (in->shadow_source) {
    // validated by check_shadow_source
    const jr_shadow_source& src = *in->shadow_source;
    const WsLayout SL = ws_layout(src.frame, src.n_tris);
    const char* sw = static_cast<const char*>(src.workspace);
    a.sl.rec = reinterpret_cast<const float4*>(sw + SL.rec);
    a.sl.counts = reinterpret_cast<const int32_t*>(sw + SL.counts);
    a.sl.offsets = reinterpret_cast<const int32_t*>(sw + SL.offsets);
    a.sl.bins = reinterpret_cast<const int32_t*>(sw + SL.bins);
    a.sl.rects = reinterpret_cast<const Rect*>(sw + SL.rects);
    a.sl.masks = reinterpret_cast<const uint32_t*>(sw + SL.masks);
    a.sl.samples = reinterpret_cast<const float*>(sw + SL.samples);
    a.sl.keys = reinterpret_cast<const uint64_t*>(sw + SL.keys);
    a.sl.ks = reinterpret_cast<const uint32_t*>(sw + SL.kstate);
    a.sl.scenes = src.scenes;
    a.sl.texels = src.texels;
    a.sl.f = dev_frame(src.frame, true);
    a.sl.zdef = src.frame->zbuffer_default;
    a.sl.depth_offset = src.frame->depth_offset;
    a.shadow_on_demand = 1;
  }
  if (n_tris == 0) {
    // no primitive: every pixel keeps its old value (plus the offset)
    if (!fr->init_buffers && !fr->apply_offset) {
      timer.mark(6);
      return JR_OK;
    }
  }
  launch_raster(fr->shader, a, L.n_tiles, fr->persistent != 0, st);
  timer.mark(6);
  const hipError_t err = hipGetLastError();
  if (err != hipSuccess) return fail(JR_E_HIP, std::string("launch failed: ") + hipGetErrorString(err));
  return JR_OK;
}

// Scratch of jr_forward_noalloc: one array per GPU, resident from the
// library's load until the process ends.  Host code never passes the array's
// name to a kernel (on the host it names a stand-in, not device memory): the
// device address of each GPU's copy comes from hipGetSymbolAddress, taken
// once per device with that device current, and is kept in g_scratch_ptr.
static __device__ __align__(256) unsigned char scratch[JR_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

int32_t scratch_of_current_device(char** ptr) {
  int dev = -1;
  hipError_t err = hipGetDevice(&dev);
  if (err != hipSuccess) return fail(JR_E_HIP, std::string("hipGetDevice: ") + hipGetErrorString(err));
  if (dev < 0 || dev >= kMaxDevices) return fail(JR_E_INVALID, "device ordinal beyond 63");
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (err != hipSuccess || !p)
      return fail(JR_E_HIP, std::string("hipGetSymbolAddress(scratch): ") +
                                (err != hipSuccess ? hipGetErrorString(err) : "NULL address"));
    g_scratch_ptr[dev] = p;
  }
  *ptr = static_cast<char*>(g_scratch_ptr[dev]);
  return JR_OK;
}

}  // namespace jr

extern "C" {

int32_t jr_forward(const jr_frame* fr, const jr_inputs* in, const jr_scene* scenes,
                   jr_outputs* out, void* workspace, int64_t workspace_bytes,
                   jr_stream_t stream_) {
  int32_t rc = check_frame(fr);
  if (rc) return rc;
  rc = check_inputs(fr, in, out);
  if (rc) return rc;
  if (!scenes) return fail(JR_E_INVALID, "scenes is NULL");
  const WsLayout L = ws_layout(fr, in->n_tris);
  if (!workspace || workspace_bytes < L.total)
    return fail(JR_E_WORKSPACE, "workspace smaller than jr_workspace_size()");
  return forward_body(fr, in, scenes, out, static_cast<char*>(workspace), L,
                      reinterpret_cast<hipStream_t>(stream_), true);
}

int32_t jr_forward_noalloc(const jr_frame* fr, const jr_inputs* in, const jr_scene* scenes,
                           jr_outputs* out, jr_stream_t stream_) {
  int32_t rc = check_frame(fr);
  if (rc) return rc;
  rc = check_inputs(fr, in, out);
  if (rc) return rc;
  if (!scenes) return fail(JR_E_INVALID, "scenes is NULL");
  if (fr->part != JR_PART_ALL)
    return fail(JR_E_INVALID, "jr_forward_noalloc runs whole passes only (part = JR_PART_ALL)");
  const WsLayout L = ws_layout(fr, in->n_tris);
  if (L.total > (int64_t)JR_SCRATCH_BYTES)
    return fail(JR_E_WORKSPACE,
                "frame needs " + std::to_string(L.total) + " bytes of workspace; "
                "jr_forward_noalloc's scratch holds JR_SCRATCH_BYTES = " +
                std::to_string((long long)JR_SCRATCH_BYTES) + " (use jr_forward)");
  char* ws = nullptr;
  rc = scratch_of_current_device(&ws);
  if (rc) return rc;
  return forward_body(fr, in, scenes, out, ws, L, reinterpret_cast<hipStream_t>(stream_),
                      false);
}

int32_t jr_stage_timing(int32_t enable) {
  std::lock_guard<std::mutex> g(g_timing_mu);
  g_timing = enable != 0;
  return JR_OK;
}

int32_t jr_stage_times(float* ms, int32_t max_calls) {
  std::vector<std::array<hipEvent_t, JR_STAGES + 1>> calls;
  {
    std::lock_guard<std::mutex> g(g_timing_mu);
    calls.swap(g_timed);
  }
  int32_t n = 0;
  for (auto& ev : calls) {
    if (n < max_calls && ms) {
      (void)hipEventSynchronize(ev[JR_STAGES]);
      for (int k = 0; k < JR_STAGES; ++k) {
        float t = -1.0f;
        if (ev[k] && ev[k + 1] && hipEventElapsedTime(&t, ev[k], ev[k + 1]) != hipSuccess)
          t = -1.0f;
        ms[JR_STAGES * n + k] = t;
      }
      ++n;
    }
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e);
  }
  return n;
}

int32_t jr_binned_pairs(const void* workspace, const jr_frame* frame, int32_t n_tris,
                        int64_t* pairs_host, jr_stream_t stream_) {
  int32_t rc = check_frame(frame);
  if (rc) return rc;
  if (!workspace || !pairs_host) return fail(JR_E_INVALID, "NULL argument");
  const WsLayout L = ws_layout(frame, n_tris);
  int32_t v = 0;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream_);
  hipError_t err = hipMemcpyAsync(&v, static_cast<const char*>(workspace) + L.misc,
                                  sizeof(int32_t), hipMemcpyDeviceToHost, st);
  if (err == hipSuccess) err = hipStreamSynchronize(st);
  if (err != hipSuccess) return fail(JR_E_HIP, hipGetErrorString(err));
  *pairs_host = v;
  return JR_OK;
}

}  // extern "C"
