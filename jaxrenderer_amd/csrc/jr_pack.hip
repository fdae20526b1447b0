// jr_pack.hip -- the per-corner flattening of a batch of scenes on the GPU.
//
// Renderer.render flattens each scene per corner (renderer/renderer.py:277-296:
// position = verts[faces.ravel()], normal = norms[faces_norm.ravel()],
// uv = uvs[faces_uv.ravel()], texture_index = texture_index[faces_uv.ravel()]).
// A batch keeps every scene's attribute rows concatenated and expresses that
// flattening as global index arrays (jr_inputs.pos_index, nrm_index,
// uv_index, tex_index): scene s's local indices plus the offsets of its rows.
// One thread per face row writes its three corners; the face's scene comes
// from a binary search over the scenes' face offsets.  A local index outside
// its scene's rows sets the error word (the caller reports it) and is written
// clamped, so nothing downstream reads another scene's rows or past an array.
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/jrhip.h"

namespace jr {
namespace {

thread_local std::string g_pack_err;

int32_t pfail(int32_t code, const char* msg) {
  g_pack_err = msg;
  return code;
}

// Last scene s with tri_off[s] <= t (tri_off ascending, tri_off[0] = 0).
__device__ __forceinline__ int32_t scene_of(const int64_t* __restrict__ tri_off,
                                            int32_t n_scenes, int64_t t) {
  int32_t lo = 0, hi = n_scenes - 1;
  while (lo < hi) {
    const int32_t mid = (lo + hi + 1) >> 1;
    if (tri_off[mid] <= t) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__global__ __launch_bounds__(256) void k_pack(jr_pack_args a) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.n_faces) return;
  const int32_t s = scene_of(a.tri_off, a.n_scenes, t);
  const int64_t v0 = a.vert_off[s], nv = a.vert_off[s + 1] - v0;
  const int64_t n0 = a.norm_off[s], nn = a.norm_off[s + 1] - n0;
  const int64_t u0 = a.uv_off[s], nu = a.uv_off[s + 1] - u0;
  const int64_t x0 = a.texidx_off[s], nx = a.texidx_off[s + 1] - x0;
  bool bad = false;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const int64_t c = 3 * t + k;
    int64_t p = a.faces[c], n = a.faces_norm[c], u = a.faces_uv[c];
    bad = bad || p < 0 || p >= nv || n < 0 || n >= nn || u < 0 || u >= nu;
    p = p < 0 ? 0 : (p >= nv ? nv - 1 : p);
    n = n < 0 ? 0 : (n >= nn ? nn - 1 : n);
    u = u < 0 ? 0 : (u >= nu ? nu - 1 : u);
    a.pos_index[c] = (int32_t)(v0 + p);
    a.nrm_index[c] = (int32_t)(n0 + n);
    a.uv_index[c] = (int32_t)(u0 + u);
    // the per-vertex texture index gathered with faces_uv (a JAX gather:
    // out-of-range rows clamp)
    if (a.tex_index)
      a.tex_index[c] = nx > 0 ? a.texture_index[x0 + (u < nx ? u : nx - 1)] : 0;
  }
  if (bad) atomicOr(a.error, 1);
}

}  // namespace
}  // namespace jr

using namespace jr;

extern "C" {

const char* jr_pack_last_error(void) { return g_pack_err.c_str(); }

int32_t jr_pack(const jr_pack_args* a, jr_stream_t stream_) {
  if (!a) return pfail(JR_E_INVALID, "args is NULL");
  if (a->n_scenes <= 0 || a->n_faces < 0 || a->n_faces > 0x7FFFFFFFLL / 3)
    return pfail(JR_E_INVALID, "n_scenes must be positive and 3 n_faces below 2^31");
  if (!a->tri_off || !a->vert_off || !a->norm_off || !a->uv_off || !a->error)
    return pfail(JR_E_INVALID, "offset arrays and the error word are required");
  if (a->tex_index && (!a->texidx_off || !a->texture_index))
    return pfail(JR_E_INVALID, "tex_index needs texidx_off and texture_index");
  if (a->n_faces == 0) return JR_OK;
  if (!a->faces || !a->faces_norm || !a->faces_uv || !a->pos_index || !a->nrm_index ||
      !a->uv_index)
    return pfail(JR_E_INVALID, "face arrays and index outputs are required");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream_);
  jr_pack_args args = *a;
  if (!args.texidx_off) args.texidx_off = args.uv_off;  // unread without tex_index
  hipLaunchKernelGGL(k_pack, dim3((unsigned)((a->n_faces + 255) / 256)), dim3(256), 0, st,
                     args);
  const hipError_t err = hipGetLastError();
  if (err != hipSuccess) return pfail(JR_E_HIP, hipGetErrorString(err));
  return JR_OK;
}

}  // extern "C"
