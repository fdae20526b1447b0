// jr_pose.hip -- object poses of a batch on the device: the per-object world
// transform of merge_objects (renderer/model.py:447-555) for every object of
// every scene in three launches, so that a batch whose objects only move is
// updated without leaving HBM.
//
//   k_pose_matrices   one thread per object: inv(T)^T (renderer/model.py:526)
//   k_pose_positions  one thread per vertex: Camera.apply_pos
//                     (renderer/geometry.py:321-346) of the scaled vertex
//   k_pose_normals    one wavefront per object: Camera.apply_vec
//                     (renderer/geometry.py:352-386) of the object's unit
//                     normals, then their division by the 2-norm of all of
//                     them, whose float32 sum of squares is added in the
//                     order of numpy's np.sum: chunks of 8192 values added
//                     in turn, each chunk summed pairwise (blocks of at most
//                     128 values summed in eight interleaved partial sums,
//                     blocks combined along the halving tree)
//
// Compiled with -ffp-contract=off: every product and sum is rounded on its
// own, in the order written, as the host-side merge_objects computes them.
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/jrhip.h"
#include "jr_math.h"

namespace jr {

int32_t fail(int32_t code, const std::string& msg);  // jr_forward.hip: sets jr_last_error()

namespace {

constexpr int kPoseThreads = 256;
constexpr int kLeafLen = 128;  // numpy's pairwise block size
constexpr int kChunk = 8192;   // numpy's reduction buffer, in values

// Leaves of the pairwise trees of an object with at most m normals: the
// halving keeps every leaf of a chunk of more than 128 values at 64 or more,
// and a shorter chunk is one leaf.
int32_t pose_leaf_cap(int32_t m) {
  const int64_t n = 3 * (int64_t)m;
  return (int32_t)(n / 64 + (n + kChunk - 1) / kChunk);
}

__global__ __launch_bounds__(kPoseThreads) void k_pose_matrices(int32_t n_objects,
                                                                const float* __restrict__ T,
                                                                float* __restrict__ nm) {
  const int32_t o = blockIdx.x * kPoseThreads + threadIdx.x;
  if (o >= n_objects) return;
  float a[16], x[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) a[i] = T[16 * (int64_t)o + i];
  inv_lu<4>(a, x);
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) nm[16 * (int64_t)o + 4 * r + c] = x[4 * c + r];
}

__global__ __launch_bounds__(kPoseThreads) void k_pose_positions(
    int32_t n_verts, const int32_t* __restrict__ vert_object, const float* __restrict__ local,
    const float* __restrict__ scaling, const float* __restrict__ T, float* __restrict__ out) {
  const int32_t v = blockIdx.x * kPoseThreads + threadIdx.x;
  if (v >= n_verts) return;
  const int64_t o = vert_object[v];
  const float p0 = local[3 * (int64_t)v + 0] * scaling[3 * o + 0];
  const float p1 = local[3 * (int64_t)v + 1] * scaling[3 * o + 1];
  const float p2 = local[3 * (int64_t)v + 2] * scaling[3 * o + 2];
  float h[4];
  apply4(T + 16 * o, p0, p1, p2, 1.0f, h);
  // to_cartesian: the homogeneous division unless w is zero
  const bool w0 = h[3] == 0.0f;
  out[3 * (int64_t)v + 0] = w0 ? h[0] : h[0] / h[3];
  out[3 * (int64_t)v + 1] = w0 ? h[1] : h[1] / h[3];
  out[3 * (int64_t)v + 2] = w0 ? h[2] : h[2] / h[3];
}

// Component k of the object's flattened [n, 3] array of transformed normals.
JR_DEV float pose_h(const float* __restrict__ unit, const float* M, int64_t first, int32_t k) {
  const int32_t i = k / 3, c = k - 3 * (k / 3);
  const float* u = unit + 3 * (first + i);
  float s = u[0] * M[4 * c + 0];
  s = s + u[1] * M[4 * c + 1];
  s = s + u[2] * M[4 * c + 2];
  s = s + 0.0f * M[4 * c + 3];
  return s;
}

// The halving of numpy's pairwise sum: a range of more than 128 values splits
// at n2 = n / 2 rounded down to a multiple of 8.
JR_DEV int32_t pose_split(int32_t n) {
  const int32_t n2 = n / 2;
  return n2 - n2 % 8;
}

// Objects of at most this many normal components keep their transformed
// normals in LDS between the passes; larger ones recompute them.
constexpr int kStageMax = 8192;

// Shared memory: cap leaf starts, cap leaf lengths, cap leaf sums, then (when
// stage) 3 * max_norms transformed components.
//
// numpy's pairwise sum of a leaf (at most 128 squares): fewer than 8 values
// are added to 0 one by one; otherwise eight partial sums over strides of 8
// (one lane each here), combined as ((r0 + r1) + (r2 + r3)) + ((r4 + r5) +
// (r6 + r7)), then the values past the last multiple of 8 added one by one.
__global__ __launch_bounds__(64) void k_pose_normals(const int32_t* __restrict__ norm_begin,
                                                     const float* __restrict__ unit,
                                                     const float* __restrict__ nm,
                                                     float* __restrict__ out,
                                                     int32_t max_norms, int32_t cap, int32_t stage) {
  extern __shared__ int32_t pose_lds[];
  int32_t* leaf_s = pose_lds;
  int32_t* leaf_n = pose_lds + cap;
  float* leaf_sum = reinterpret_cast<float*>(pose_lds + 2 * cap);
  float* hs = reinterpret_cast<float*>(pose_lds + 3 * cap);
  __shared__ float part[64];
  __shared__ int32_t n_leaves;
  __shared__ float norm;
  // lane 0's walk stacks (in LDS: a private array indexed at run time would
  // live in scratch memory, and every step of the walks depends on the last)
  __shared__ int32_t st_s[32], st_n[32], st_has[32];
  __shared__ float st_left[32];
  const int32_t o = blockIdx.x;
  const int lane = threadIdx.x;
  const int64_t first = norm_begin[o];
  const int32_t count = norm_begin[o + 1] - norm_begin[o];
  // an object beyond the stated maximum is left unwritten (its leaves would
  // not fit)
  if (count < 0 || count > max_norms) return;
  const int32_t n = 3 * count;
  float M[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) M[i] = nm[16 * (int64_t)o + i];
  if (stage)
    for (int32_t k = lane; k < n; k += 64) hs[k] = pose_h(unit, M, first, k);

  // the leaves of every chunk's pairwise tree, left to right (a depth-first
  // walk that visits the left half of every split first)
  if (lane == 0) {
    int32_t k = 0;
    for (int32_t cs = 0; cs < n; cs += kChunk) {
      st_s[0] = cs;
      st_n[0] = min(kChunk, n - cs);
      int32_t d = 1;
      while (d > 0) {
        --d;
        const int32_t s = st_s[d], len = st_n[d];
        if (len <= kLeafLen) {
          leaf_s[k] = s;
          leaf_n[k] = len;
          ++k;
        } else {
          const int32_t n2 = pose_split(len);
          st_s[d] = s + n2;  // right half, taken after the left one
          st_n[d] = len - n2;
          st_s[d + 1] = s;
          st_n[d + 1] = n2;
          d += 2;
        }
      }
    }
    n_leaves = k;
  }
  __syncthreads();
  const int32_t L = n_leaves;
  // eight leaves at a time: partial sum j of leaf base + q on lane 8q + j,
  // then lane q combines leaf base + q's eight
  for (int32_t base = 0; base < L; base += 8) {
    {
      const int32_t k = base + (lane >> 3), j = lane & 7;
      float r = 0.0f;
      if (k < L && leaf_n[k] >= 8) {
        const int32_t s = leaf_s[k], len = leaf_n[k];
        float h = stage ? hs[s + j] : pose_h(unit, M, first, s + j);
        r = h * h;
        for (int32_t i = 8; i < len - (len % 8); i += 8) {
          h = stage ? hs[s + i + j] : pose_h(unit, M, first, s + i + j);
          r = r + h * h;
        }
      }
      part[lane] = r;
    }
    __syncthreads();
    const int32_t k = base + lane;
    if (lane < 8 && k < L) {
      const int32_t s = leaf_s[k], len = leaf_n[k];
      const float* r = part + 8 * lane;
      float res = 0.0f;
      int32_t i = 0;
      if (len >= 8) {
        res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
        i = len - (len % 8);
      }
      for (; i < len; ++i) {
        const float h = stage ? hs[s + i] : pose_h(unit, M, first, s + i);
        res = res + h * h;
      }
      leaf_sum[k] = res;
    }
    __syncthreads();
  }

  // combine each chunk's leaves along its tree (post-order: each split adds
  // its left half's sum to its right half's), and the chunks' sums in turn
  if (lane == 0) {
    float total = 0.0f;
    int32_t next = 0;
    for (int32_t cs = 0; cs < n; cs += kChunk) {
      int32_t d = 0, len = min(kChunk, n - cs);
      float v;
      for (;;) {
        while (len > kLeafLen) {
          const int32_t n2 = pose_split(len);
          st_n[d] = len - n2;  // the right half still to do
          st_has[d] = 0;
          ++d;
          len = n2;
        }
        v = leaf_sum[next++];
        while (d > 0 && st_has[d - 1]) {
          v = st_left[d - 1] + v;
          --d;
        }
        if (d == 0) break;
        st_left[d - 1] = v;  // a left half done: its right half is next
        st_has[d - 1] = 1;
        len = st_n[d - 1];
      }
      total = total + v;
    }
    norm = sqrtf(total);
  }
  __syncthreads();
  const float sn = norm;
  for (int32_t k = lane; k < n; k += 64)
    out[3 * first + k] = (stage ? hs[k] : pose_h(unit, M, first, k)) / sn;
}

}  // namespace
}  // namespace jr

using namespace jr;

extern "C" {

int32_t jr_pose(const jr_pose_args* a, jr_stream_t stream) {
  if (!a) return fail(JR_E_INVALID, "jr_pose: args is NULL");
  if (a->n_objects < 0 || a->n_verts < 0 || a->n_norms < 0)
    return fail(JR_E_INVALID, "jr_pose: negative count");
  if (a->n_objects == 0) return JR_OK;
  if (!a->transforms || !a->normal_matrices || !a->scaling)
    return fail(JR_E_INVALID, "jr_pose: scaling, transforms and normal_matrices are required");
  if (a->n_verts > 0 && (!a->vert_object || !a->local_verts || !a->positions))
    return fail(JR_E_INVALID, "jr_pose: vertices without vert_object, local_verts or positions");
  if (a->normals && (!a->norm_begin || (a->n_norms > 0 && !a->unit_norms)))
    return fail(JR_E_INVALID, "jr_pose: normals without norm_begin or unit_norms");
  if (a->normals && (a->max_object_norms < 0 || a->max_object_norms > JR_POSE_MAX_NORMALS))
    return fail(JR_E_INVALID, "jr_pose: max_object_norms must be in [0, " +
                                  std::to_string(JR_POSE_MAX_NORMALS) + "]");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const unsigned g_obj = (unsigned)((a->n_objects + kPoseThreads - 1) / kPoseThreads);
  hipLaunchKernelGGL(k_pose_matrices, dim3(g_obj), dim3(kPoseThreads), 0, st, a->n_objects,
                     a->transforms, a->normal_matrices);
  if (a->n_verts > 0) {
    const unsigned g_v = (unsigned)((a->n_verts + kPoseThreads - 1) / kPoseThreads);
    hipLaunchKernelGGL(k_pose_positions, dim3(g_v), dim3(kPoseThreads), 0, st, a->n_verts,
                       a->vert_object, a->local_verts, a->scaling, a->transforms, a->positions);
  }
  if (a->normals) {
    const int32_t cap = pose_leaf_cap(a->max_object_norms);
    const int32_t stage = 3 * a->max_object_norms <= kStageMax ? 1 : 0;
    const size_t lds = (size_t)cap * 12 + (stage ? (size_t)3 * a->max_object_norms * 4 : 0);
    hipLaunchKernelGGL(k_pose_normals, dim3((unsigned)a->n_objects), dim3(64), lds, st,
                       a->norm_begin, a->unit_norms, a->normal_matrices, a->normals,
                       a->max_object_norms, cap, stage);
  }
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(JR_E_HIP, std::string("jr_pose: ") + hipGetErrorString(e));
  return JR_OK;
}

}  // extern "C"
