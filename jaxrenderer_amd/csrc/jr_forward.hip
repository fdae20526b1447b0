// jr_forward.hip -- the forward rasterisation pass on CDNA4 (gfx950).
//
// Replaces the body of renderer/pipeline.py:render (vertex stage, per-primitive
// setup, the scan over rows x vmap over pixels x vmap over primitives, chooser,
// interpolate, fragment, mix and merge) with five kernel stages:
//
//   1. k_setup   one thread per triangle: clip-space corners (vertex stage
//                positions), det / keep / LU inverse (PerPrimitive.create,
//                renderer/pipeline.py:76-113), a conservative screen-tile range
//                and a count of the tiles whose pixels it may cover.
//   2. k_scan_*  exclusive scan of the per-tile counts (all scenes at once).
//   3. k_bin     one thread per triangle writes its id into every tile list.
//   4. k_raster  one wavefront-sized group per (scene, tile): stages the tile's
//                triangle records in LDS, keeps the nearest (depth, id) key per
//                pixel (the default primitive_chooser: argmin with first-index
//                ties), then shades the winner and writes the merged buffers
//                with stores along the contiguous height axis.
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
#include <vector>

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
  int32_t tx0, tx1, ty0, ty1;  // inclusive tile range, empty when tx0 > tx1
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
  int32_t n_scenes;
  int64_t capacity;
};

JR_DEV int32_t scene_of_tri(const jr_scene* __restrict__ scenes, int32_t n,
                            int32_t t) {
  int32_t lo = 0, hi = n - 1;
  while (lo < hi) {
    const int32_t mid = (lo + hi + 1) >> 1;
    if (scenes[mid].tri_begin <= t)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// Conservative tile range of a triangle.  The pixel test accepts (x, y) only
// when every float edge value is >= 0; float rounding keeps each computed edge
// within 3u(|x a| + |y b| + |c|) of its exact value (u = 2^-24), so every
// accepted pixel lies in the region {a x + b y + c >= -delta_k} with
// delta_k = 8u(X|a| + Y|b| + |c|).  That region, intersected with the
// viewport's sample rectangle, is clipped in double precision and its bounding
// box, widened by one pixel, gives the tile range.
JR_DEV Rect tile_range(const float4& r0, const float4& r1, const float4& r2,
                       bool sane, const float* vp, const Frame& f) {
  Rect full = {0, f.tiles_x - 1, 0, f.tiles_y - 1};
  if (!sane || !(vp[0] > 0.0f) || !(vp[5] > 0.0f)) return full;
  const double X0 = pix_x(0, vp), X1 = pix_x(f.W - 1, vp);
  const double Y0 = pix_y(0, vp), Y1 = pix_y(f.H - 1, vp);
  if (!(X0 <= X1) || !(Y0 <= Y1)) return full;
  const double XM = fmax(fabs(X0), fabs(X1)), YM = fmax(fabs(Y0), fabs(Y1));
  double px[8], py[8], qx[8], qy[8];
  px[0] = X0; py[0] = Y0;
  px[1] = X1; py[1] = Y0;
  px[2] = X1; py[2] = Y1;
  px[3] = X0; py[3] = Y1;
  int n = 4;
  const float4 rs[3] = {r0, r1, r2};
  for (int k = 0; k < 3 && n > 0; ++k) {
    const double a = rs[k].x, b = rs[k].y, c = rs[k].z;
    const double delta =
        (8.0 / 16777216.0) * (XM * fabs(a) + YM * fabs(b) + fabs(c)) + 1e-300;
    int m = 0;
    for (int v = 0; v < n; ++v) {
      const int w = (v + 1 == n) ? 0 : v + 1;
      const double dv = a * px[v] + b * py[v] + c + delta;
      const double dw = a * px[w] + b * py[w] + c + delta;
      if (dv >= 0.0) {
        qx[m] = px[v];
        qy[m] = py[v];
        ++m;
      }
      if ((dv >= 0.0) != (dw >= 0.0)) {
        const double t = dv / (dv - dw);
        qx[m] = px[v] + t * (px[w] - px[v]);
        qy[m] = py[v] + t * (py[w] - py[v]);
        ++m;
      }
    }
    n = m;
    for (int v = 0; v < n; ++v) {
      px[v] = qx[v];
      py[v] = qy[v];
    }
  }
  Rect r = {1, 0, 1, 0};
  if (n == 0) return r;
  double xmin = px[0], xmax = px[0], ymin = py[0], ymax = py[0];
  for (int v = 1; v < n; ++v) {
    xmin = fmin(xmin, px[v]);
    xmax = fmax(xmax, px[v]);
    ymin = fmin(ymin, py[v]);
    ymax = fmax(ymax, py[v]);
  }
  const double vx = vp[0], vx3 = vp[3], vy = vp[5], vy3 = vp[7];
  double i0 = floor(xmin * vx + vx3) - 1.0, i1 = ceil(xmax * vx + vx3) + 1.0;
  double j0 = floor(ymin * vy + vy3) - 1.0, j1 = ceil(ymax * vy + vy3) + 1.0;
  i0 = fmax(i0, 0.0);
  j0 = fmax(j0, 0.0);
  i1 = fmin(i1, (double)(f.W - 1));
  j1 = fmin(j1, (double)(f.H - 1));
  if (!(i0 <= i1) || !(j0 <= j1)) return r;
  r.tx0 = (int32_t)i0 / f.TW;
  r.tx1 = (int32_t)i1 / f.TW;
  r.ty0 = (int32_t)j0 / f.TH;
  r.ty1 = (int32_t)j1 / f.TH;
  return r;
}

// Exact tile rejection: every computed edge value is monotone in i and in j
// (IEEE rounding is monotone), so its maximum over a tile sits at a corner
// pixel.  A tile whose four corners all give e_k < 0 for some k holds no pixel
// that passes the reference's inside test.
JR_DEV bool tile_may_cover(const float4& r0, const float4& r1, const float4& r2,
                           bool sane, const float* vp, const Frame& f,
                           int32_t tx, int32_t ty) {
  if (!sane) return true;
  const int32_t i0 = tx * f.TW, i1 = min(i0 + f.TW, f.W) - 1;
  const int32_t j0 = ty * f.TH, j1 = min(j0 + f.TH, f.H) - 1;
  const float xa = pix_x(i0, vp), xb = pix_x(i1, vp);
  const float ya = pix_y(j0, vp), yb = pix_y(j1, vp);
  const float4 rs[3] = {r0, r1, r2};
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float e00 = edge(rs[k], xa, ya), e01 = edge(rs[k], xa, yb);
    const float e10 = edge(rs[k], xb, ya), e11 = edge(rs[k], xb, yb);
    if (e00 < 0.0f && e01 < 0.0f && e10 < 0.0f && e11 < 0.0f) return false;
  }
  return true;
}

__global__ void k_zero(int32_t* __restrict__ p, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0;
}

// Stage 1: vertex positions -> clip space, PerPrimitive.create, tile counts.
__global__ __launch_bounds__(256) void k_setup(
    jr_inputs in, const jr_scene* __restrict__ scenes, Frame f,
    float4* __restrict__ rec, Rect* __restrict__ rects,
    int32_t* __restrict__ counts) {
  const int32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= in.n_tris) return;
  const int32_t s = scene_of_tri(scenes, f.n_scenes, t);
  const jr_scene& sc = scenes[s];
  float clip[3][4];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const int32_t slot = in.faces[3 * (int64_t)t + k];
    if (in.clip) {
#pragma unroll
      for (int q = 0; q < 4; ++q) clip[k][q] = in.clip[4 * (int64_t)slot + q];
    } else {
      position_clip(in, sc.world_to_clip, slot, clip[k]);
    }
  }
  // matrix with rows (x, y, w) of the three vertices
  float M[9];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    M[3 * k + 0] = clip[k][0];
    M[3 * k + 1] = clip[k][1];
    M[3 * k + 2] = clip[k][3];
  }
  const float det = det3(M);
  const bool keep = fabsf(det) > 1e-6f;
  const bool front = det >= 0.0f;
  float Mi[9];
  inv3(M, Mi);
  bool sane = true;
#pragma unroll
  for (int i = 0; i < 9; ++i) sane = sane && (fabsf(Mi[i]) < 1e18f);
  uint32_t flags = (keep ? kKeep : 0u) | (front ? kFront : 0u) | (sane ? kSane : 0u);
  const float4 r0 = make_float4(Mi[0], Mi[3], Mi[6], clip[0][2]);
  const float4 r1 = make_float4(Mi[1], Mi[4], Mi[7], clip[1][2]);
  const float4 r2 = make_float4(Mi[2], Mi[5], Mi[8], clip[2][2]);
  rec[4 * (int64_t)t + 0] = r0;
  rec[4 * (int64_t)t + 1] = r1;
  rec[4 * (int64_t)t + 2] = r2;
  rec[4 * (int64_t)t + 3] = make_float4(det, __uint_as_float(flags), 0.0f, 0.0f);
  Rect r = {1, 0, 1, 0};
  if (keep && front) r = tile_range(r0, r1, r2, sane, sc.viewport, f);
  rects[t] = r;
  int32_t* cnt = counts + (int64_t)s * f.T;
  for (int32_t tx = r.tx0; tx <= r.tx1; ++tx)
    for (int32_t ty = r.ty0; ty <= r.ty1; ++ty)
      if (tile_may_cover(r0, r1, r2, sane, sc.viewport, f, tx, ty))
        atomicAdd(cnt + tx * f.tiles_y + ty, 1);
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
__global__ __launch_bounds__(256) void k_bin(
    jr_inputs in, const jr_scene* __restrict__ scenes, Frame f,
    const float4* __restrict__ rec, const Rect* __restrict__ rects,
    const int32_t* __restrict__ offsets, int32_t* __restrict__ cursor,
    int32_t* __restrict__ bins) {
  const int32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= in.n_tris) return;
  const Rect r = rects[t];
  if (r.tx0 > r.tx1) return;
  const int32_t s = scene_of_tri(scenes, f.n_scenes, t);
  const float* vp = scenes[s].viewport;
  const float4 r0 = rec[4 * (int64_t)t], r1 = rec[4 * (int64_t)t + 1],
               r2 = rec[4 * (int64_t)t + 2];
  const bool sane = (__float_as_uint(rec[4 * (int64_t)t + 3].y) & kSane) != 0;
  const int64_t tbase = (int64_t)s * f.T;
  for (int32_t tx = r.tx0; tx <= r.tx1; ++tx)
    for (int32_t ty = r.ty0; ty <= r.ty1; ++ty)
      if (tile_may_cover(r0, r1, r2, sane, vp, f, tx, ty)) {
        const int64_t gt = tbase + tx * f.tiles_y + ty;
        const int64_t pos = (int64_t)offsets[gt] + atomicAdd(cursor + gt, 1);
        if (pos < f.capacity) bins[pos] = t;
      }
}

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
  int32_t* work_counter;
  int32_t init_buffers, apply_offset;
  float zdef, cdef0, cdef1, cdef2, depth_offset;
};

template <int SHADER, int TW, int TH>
JR_DEV void raster_tile(const RasterArgs& a, int64_t gt, float4* lds_rec,
                        int32_t* lds_id) {
  constexpr int NT = TW * TH;
  const Frame& f = a.f;
  const int32_t s = (int32_t)(gt / f.T);
  const int32_t tile = (int32_t)(gt - (int64_t)s * f.T);
  const int32_t tx = tile / f.tiles_y, ty = tile - tx * f.tiles_y;
  const int tid = threadIdx.x;
  const int32_t i = tx * TW + tid / TH;
  const int32_t j = ty * TH + tid % TH;
  const bool inside = (i < f.W) && (j < f.H);
  const jr_scene& sc = a.scenes[s];
  const float* vp = sc.viewport;
  const float x = pix_x(i, vp), y = pix_y(j, vp);
  const float vp22 = vp[10], vp23 = vp[11];

  const int32_t cnt = a.counts[gt];
  const int64_t start = a.offsets[gt];
  const bool overflow = start + cnt > f.capacity;
  // overflowing tile lists fall back to every triangle of the scene
  const int32_t n_items = overflow ? (sc.tri_end - sc.tri_begin) : cnt;

  uint32_t best_key = 0xFFFFFFFFu;
  int32_t best_id = 0x7FFFFFFF;
  for (int32_t base = 0; base < n_items; base += NT) {
    const int32_t n = min(NT, n_items - base);
    __syncthreads();
    if (tid < n) {
      const int32_t t = overflow ? (sc.tri_begin + base + tid) : a.bins[start + base + tid];
      const float4 r0 = a.rec[4 * (int64_t)t], r1 = a.rec[4 * (int64_t)t + 1],
                   r2 = a.rec[4 * (int64_t)t + 2];
      const uint32_t fl = __float_as_uint(a.rec[4 * (int64_t)t + 3].y);
      lds_rec[3 * tid + 0] = r0;
      lds_rec[3 * tid + 1] = r1;
      lds_rec[3 * tid + 2] = r2;
      // triangles that are not kept or not front-facing never win the chooser
      lds_id[tid] = ((fl & (kKeep | kFront)) == (kKeep | kFront)) ? t : -1;
    }
    __syncthreads();
    for (int q = 0; q < n; ++q) {
      const int32_t t = lds_id[q];
      const float4 r0 = lds_rec[3 * q + 0], r1 = lds_rec[3 * q + 1],
                   r2 = lds_rec[3 * q + 2];
      const float e0 = edge(r0, x, y), e1 = edge(r1, x, y), e2 = edge(r2, x, y);
      if (t >= 0 && e0 >= 0.0f && e1 >= 0.0f && e2 >= 0.0f) {
        const uint32_t key = depth_key(win_depth(r0, r1, r2, e0, e1, e2, vp22, vp23));
        if (key < best_key || (key == best_key && t < best_id)) {
          best_key = key;
          best_id = t;
        }
      }
    }
  }
  if (!inside) return;

  // argmin over all primitives returns index 0 of the scene when no candidate
  // has a finite (or NaN) depth (renderer/shader.py:207-217); the chosen
  // primitive's own keep & in_triangle then decides (renderer/pipeline.py:275).
  int32_t w;
  bool cover;
  if (sc.tri_end <= sc.tri_begin) {
    w = -1;  // a scene without primitives draws nothing
    cover = false;
  } else if (best_key >= kInfKey) {
    w = sc.tri_begin;
    const float4 r0 = a.rec[4 * (int64_t)w], r1 = a.rec[4 * (int64_t)w + 1],
                 r2 = a.rec[4 * (int64_t)w + 2];
    const uint32_t fl = __float_as_uint(a.rec[4 * (int64_t)w + 3].y);
    const float e0 = edge(r0, x, y), e1 = edge(r1, x, y), e2 = edge(r2, x, y);
    cover = (fl & kKeep) && e0 >= 0.0f && e1 >= 0.0f && e2 >= 0.0f;
  } else {
    w = best_id;
    cover = true;
  }

  const int64_t pix = (int64_t)s * f.W * f.H + (int64_t)i * f.H + j;
  bool keep = false;
  float z = 0.0f, b[3] = {0.0f, 0.0f, 0.0f}, col[3] = {0.0f, 0.0f, 0.0f};
  if (cover) {
    const float4 r0 = a.rec[4 * (int64_t)w], r1 = a.rec[4 * (int64_t)w + 1],
                 r2 = a.rec[4 * (int64_t)w + 2];
    const uint32_t fl = __float_as_uint(a.rec[4 * (int64_t)w + 3].y);
    const float e0 = edge(r0, x, y), e1 = edge(r1, x, y), e2 = edge(r2, x, y);
    float wr = e0 + e1;
    wr = wr + e2;
    b[0] = e0 / wr;
    b[1] = e1 / wr;
    b[2] = e2 / wr;
    z = win_depth(r0, r1, r2, e0, e1, e2, vp22, vp23);
    Corner c;
    c.slot[0] = a.in.faces[3 * (int64_t)w + 0];
    c.slot[1] = a.in.faces[3 * (int64_t)w + 1];
    c.slot[2] = a.in.faces[3 * (int64_t)w + 2];
    keep = Shade<SHADER>::run(a.in, sc, c, b, (fl & kFront) != 0, col, s);
  }

  // merge (renderer/pipeline.py:401-440) and the shadow-map offset
  // (renderer/shadow.py:116)
  if (keep) {
    a.out.zbuffer[pix] = a.apply_offset ? z + a.depth_offset : z;
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
  if (a.out.prim_id) a.out.prim_id[pix] = keep ? w : -1;
  if (a.out.bary) {
    a.out.bary[3 * pix + 0] = keep ? b[0] : 0.0f;
    a.out.bary[3 * pix + 1] = keep ? b[1] : 0.0f;
    a.out.bary[3 * pix + 2] = keep ? b[2] : 0.0f;
  }
}

template <int SHADER, int TW, int TH>
__global__ __launch_bounds__(TW* TH) void k_raster(RasterArgs a) {
  __shared__ float4 lds_rec[3 * TW * TH];
  __shared__ int32_t lds_id[TW * TH];
  raster_tile<SHADER, TW, TH>(a, blockIdx.x, lds_rec, lds_id);
}

// Persistent form: a fixed grid drains the (scene, tile) queue.  Every group
// exits once the shared counter passes the item count.
template <int SHADER, int TW, int TH>
__global__ __launch_bounds__(TW* TH) void k_raster_persistent(RasterArgs a,
                                                              int64_t n_items) {
  __shared__ float4 lds_rec[3 * TW * TH];
  __shared__ int32_t lds_id[TW * TH];
  __shared__ int32_t item;
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) item = atomicAdd(a.work_counter, 1);
    __syncthreads();
    const int32_t gt = item;
    if (gt >= n_items) break;
    raster_tile<SHADER, TW, TH>(a, gt, lds_rec, lds_id);
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
std::vector<std::array<hipEvent_t, 5>> g_timed;

struct StageTimer {
  hipStream_t st;
  std::array<hipEvent_t, 5> ev{};
  bool on = false;
  explicit StageTimer(hipStream_t s) : st(s) {
    std::lock_guard<std::mutex> g(g_timing_mu);
    on = g_timing;
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
  int64_t rec, rects, counts, cursor, offsets, block_sums, misc, bins, total;
  int64_t n_tiles, nblk;
};

WsLayout ws_layout(const jr_frame* fr, int32_t n_tris) {
  WsLayout L{};
  const int64_t tiles_x = (fr->width + fr->tile_w - 1) / fr->tile_w;
  const int64_t tiles_y = (fr->height + fr->tile_h - 1) / fr->tile_h;
  L.n_tiles = tiles_x * tiles_y * (int64_t)fr->n_scenes;
  L.nblk = (L.n_tiles + kScanBlock - 1) / kScanBlock;
  int64_t o = 0;
  L.rec = o;        o += align_up(64 * (int64_t)std::max(n_tris, 1));
  L.rects = o;      o += align_up(16 * (int64_t)std::max(n_tris, 1));
  L.counts = o;     o += align_up(4 * L.n_tiles);
  L.cursor = o;     o += align_up(4 * L.n_tiles);
  L.misc = o;       o += align_up(64);   // [0] total pairs, [1] work counter
  L.offsets = o;    o += align_up(4 * (L.n_tiles + 1));
  L.block_sums = o; o += align_up(4 * std::max<int64_t>(L.nblk, 1));
  L.bins = o;       o += align_up(4 * std::max<int64_t>(fr->bin_capacity, 1));
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
  if (fr->bin_capacity < 0 || fr->bin_capacity > 0x7FFFFFFFLL)
    return fail(JR_E_INVALID, "bin_capacity out of range");
  return JR_OK;
}

template <int SHADER, int TW, int TH>
void launch_raster_t(const RasterArgs& a, int64_t n_items, bool persistent,
                     hipStream_t st) {
  if (persistent) {
    int dev = 0, cus = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    const int per_cu = 2048 / (TW * TH);  // up to 32 waves per CU
    const int64_t grid = std::min<int64_t>(n_items, (int64_t)std::max(cus, 1) * per_cu);
    hipLaunchKernelGGL((k_raster_persistent<SHADER, TW, TH>), dim3((unsigned)grid),
                       dim3(TW * TH), 0, st, a, n_items);
  } else {
    hipLaunchKernelGGL((k_raster<SHADER, TW, TH>), dim3((unsigned)n_items),
                       dim3(TW * TH), 0, st, a);
  }
}

template <int SHADER>
void launch_raster_s(const RasterArgs& a, int64_t n_items, bool persistent,
                     hipStream_t st) {
  const int tw = a.f.TW, th = a.f.TH;
  if (tw == 8 && th == 8) launch_raster_t<SHADER, 8, 8>(a, n_items, persistent, st);
  else if (tw == 16 && th == 16) launch_raster_t<SHADER, 16, 16>(a, n_items, persistent, st);
  else if (tw == 8 && th == 32) launch_raster_t<SHADER, 8, 32>(a, n_items, persistent, st);
  else launch_raster_t<SHADER, 4, 16>(a, n_items, persistent, st);
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
    case JR_SHADER_PHONG_REFLECTION_SHADOW: launch_raster_s<JR_SHADER_PHONG_REFLECTION_SHADOW>(a, n_items, persistent, st); break;
    default: launch_raster_s<JR_SHADER_VISIBILITY>(a, n_items, persistent, st); break;
  }
}

int32_t check_inputs(const jr_frame* fr, const jr_inputs* in, const jr_outputs* out) {
  if (!in || !out) return fail(JR_E_INVALID, "inputs/outputs NULL");
  if (in->n_tris < 0 || in->n_slots < 0) return fail(JR_E_INVALID, "negative sizes");
  if (!out->zbuffer) return fail(JR_E_INVALID, "zbuffer is required");
  if (in->n_tris == 0) return JR_OK;
  if ((!in->positions && !in->clip) || !in->faces)
    return fail(JR_E_INVALID, "positions (or clip) and faces are required");
  if (in->clip && !(fr->shader == JR_SHADER_VISIBILITY || fr->shader == JR_SHADER_DEPTH))
    return fail(JR_E_INVALID, "clip input is only valid with the depth/visibility shaders");
  const int sh = fr->shader;
  const bool needs_canvas = !(sh == JR_SHADER_DEPTH || sh == JR_SHADER_VISIBILITY);
  if (needs_canvas && !out->canvas) return fail(JR_E_INVALID, "shader writes a canvas but canvas is NULL");
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
      (!in->shadow_maps || in->shadow_w <= 0 || in->shadow_h <= 0))
    return fail(JR_E_INVALID, "phong_reflection_shadow needs shadow maps");
  return JR_OK;
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
    default: return -1;
  }
}

int64_t jr_workspace_size(const jr_frame* frame, int32_t n_tris) {
  if (check_frame(frame) != JR_OK || n_tris < 0) return -1;
  return ws_layout(frame, n_tris).total;
}

int32_t jr_forward(const jr_frame* fr, const jr_inputs* in, const jr_scene* scenes,
                   jr_outputs* out, void* workspace, int64_t workspace_bytes,
                   jr_stream_t stream_) {
  int32_t rc = check_frame(fr);
  if (rc) return rc;
  rc = check_inputs(fr, in, out);
  if (rc) return rc;
  if (!scenes) return fail(JR_E_INVALID, "scenes is NULL");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream_);
  const WsLayout L = ws_layout(fr, in->n_tris);
  if (!workspace || workspace_bytes < L.total)
    return fail(JR_E_WORKSPACE, "workspace smaller than jr_workspace_size()");
  char* ws = static_cast<char*>(workspace);
  Frame f;
  f.W = fr->width;
  f.H = fr->height;
  f.TW = fr->tile_w;
  f.TH = fr->tile_h;
  f.tiles_x = (f.W + f.TW - 1) / f.TW;
  f.tiles_y = (f.H + f.TH - 1) / f.TH;
  f.T = f.tiles_x * f.tiles_y;
  f.n_scenes = fr->n_scenes;
  f.capacity = fr->bin_capacity;

  float4* rec = reinterpret_cast<float4*>(ws + L.rec);
  Rect* rects = reinterpret_cast<Rect*>(ws + L.rects);
  int32_t* counts = reinterpret_cast<int32_t*>(ws + L.counts);
  int32_t* cursor = reinterpret_cast<int32_t*>(ws + L.cursor);
  int32_t* misc = reinterpret_cast<int32_t*>(ws + L.misc);
  int32_t* offsets = reinterpret_cast<int32_t*>(ws + L.offsets);
  int32_t* block_sums = reinterpret_cast<int32_t*>(ws + L.block_sums);
  int32_t* bins = reinterpret_cast<int32_t*>(ws + L.bins);

  StageTimer timer(st);
  // counts, cursor and misc are contiguous: clear them with one fill kernel
  // (a kernel rather than hipMemsetAsync so the sequence can be graph-captured)
  const int64_t nz = (L.offsets - L.counts) / 4;
  hipLaunchKernelGGL(k_zero, dim3((unsigned)std::min<int64_t>((nz + 255) / 256, 4096)),
                     dim3(256), 0, st, counts, nz);
  const int n_tris = in->n_tris;
  if (n_tris > 0) {
    hipLaunchKernelGGL(k_setup, dim3((n_tris + 255) / 256), dim3(256), 0, st, *in,
                       scenes, f, rec, rects, counts);
  }
  timer.mark(1);
  hipLaunchKernelGGL(k_scan_local, dim3((unsigned)std::max<int64_t>(L.nblk, 1)),
                     dim3(kScanThreads), 0, st, counts, offsets, block_sums, L.n_tiles);
  hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kScanThreads), 0, st, block_sums,
                     L.nblk, misc);
  hipLaunchKernelGGL(k_scan_add, dim3((unsigned)std::max<int64_t>(L.nblk, 1)),
                     dim3(kScanThreads), 0, st, offsets, block_sums, L.n_tiles);
  timer.mark(2);
  if (n_tris > 0) {
    hipLaunchKernelGGL(k_bin, dim3((n_tris + 255) / 256), dim3(256), 0, st, *in,
                       scenes, f, rec, rects, offsets, cursor, bins);
  }
  timer.mark(3);
  RasterArgs a;
  a.in = *in;
  a.scenes = scenes;
  a.out = *out;
  a.f = f;
  a.rec = rec;
  a.counts = counts;
  a.offsets = offsets;
  a.bins = bins;
  a.work_counter = misc + 1;
  a.init_buffers = fr->init_buffers;
  a.apply_offset = fr->apply_offset;
  a.zdef = fr->zbuffer_default;
  a.cdef0 = fr->colour_default[0];
  a.cdef1 = fr->colour_default[1];
  a.cdef2 = fr->colour_default[2];
  a.depth_offset = fr->depth_offset;
  if (n_tris == 0) {
    // no primitive: every pixel keeps its old value (plus the offset)
    if (!fr->init_buffers && !fr->apply_offset) {
      timer.mark(4);
      return JR_OK;
    }
  }
  launch_raster(fr->shader, a, L.n_tiles, fr->persistent != 0, st);
  timer.mark(4);
  const hipError_t err = hipGetLastError();
  if (err != hipSuccess) return fail(JR_E_HIP, std::string("launch failed: ") + hipGetErrorString(err));
  return JR_OK;
}

int32_t jr_stage_timing(int32_t enable) {
  std::lock_guard<std::mutex> g(g_timing_mu);
  g_timing = enable != 0;
  return JR_OK;
}

int32_t jr_stage_times(float* ms, int32_t max_calls) {
  std::vector<std::array<hipEvent_t, 5>> calls;
  {
    std::lock_guard<std::mutex> g(g_timing_mu);
    calls.swap(g_timed);
  }
  int32_t n = 0;
  for (auto& ev : calls) {
    if (n < max_calls && ms) {
      (void)hipEventSynchronize(ev[4]);
      for (int k = 0; k < 4; ++k) {
        float t = -1.0f;
        if (ev[k] && ev[k + 1] && hipEventElapsedTime(&t, ev[k], ev[k + 1]) != hipSuccess)
          t = -1.0f;
        ms[4 * n + k] = t;
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
