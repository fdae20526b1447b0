/*
 * jrhip.h -- C ABI of the MI355X rasterisation library (libjrhip.so).
 *
 * The library replaces the body of jaxrenderer's `pipeline.render`
 * (renderer/pipeline.py:470-537) and the two shadow passes
 * (renderer/shadow.py:40-125, renderer/renderer.py:346-382) with HIP kernels
 * for gfx950.  Every pointer in these structs is a device pointer owned by the
 * caller; the library never allocates device memory: scratch space comes in as
 * a workspace argument sized by jr_workspace_size().  Every entry point is
 * stream ordered on the hipStream_t it receives and returns 0 on success or a
 * negative JR_E_* code; jr_last_error() returns the thread's last message.
 *
 * Buffer layout follows the reference's types (renderer/types.py:63-64,88):
 * image buffers are x-major, [scene][width][height](,channel), float32, and a
 * pixel (i, j) samples NDC ((i - vp[0][3]) / vp[0][0], (j - vp[1][3]) / vp[1][1])
 * with integer i, j (renderer/pipeline.py:177).
 *
 * The same structs are consumed by the CPU oracle (oracle/jr_oracle.c), which
 * takes host pointers instead.  This header is plain C.
 */
#ifndef JRHIP_H
#define JRHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define JR_ABI_VERSION 2

/* Built-in shaders (renderer/shaders/ directory).  JR_SHADER_VISIBILITY runs only
 * the default primitive chooser (renderer/shader.py:155-251) and the default
 * fragment (renderer/shader.py:292-333): it writes depth, primitive id and
 * barycentrics; the Python layer runs user-defined hooks on top of it. */
enum {
  JR_SHADER_DEPTH = 0,                   /* shaders/depth.py:38-61 */
  JR_SHADER_GOURAUD = 1,                 /* shaders/gouraud.py:49-124 */
  JR_SHADER_GOURAUD_TEXTURE = 2,         /* shaders/gouraud_texture.py:55-167 */
  JR_SHADER_PHONG = 3,                   /* shaders/phong.py:65-187 */
  JR_SHADER_PHONG_DARBOUX = 4,           /* shaders/phong_darboux.py:106-316 */
  JR_SHADER_PHONG_REFLECTION = 5,        /* shaders/phong_reflection.py:91-258 */
  JR_SHADER_PHONG_REFLECTION_SHADOW = 6, /* shaders/phong_reflection_shadow.py:99-295 */
  JR_SHADER_VISIBILITY = 7,
  JR_SHADER_COUNT = 8
};

/* error codes */
enum {
  JR_OK = 0,
  JR_E_INVALID = -1,   /* bad argument (shape, null pointer, unknown shader) */
  JR_E_WORKSPACE = -2, /* workspace too small */
  JR_E_HIP = -3,       /* a HIP runtime call failed */
  JR_E_UNSUPPORTED = -4
};

/* Per-scene uniforms.  One struct per scene, in device memory for
 * jr_forward / jr_backward and host memory for the oracle.  Matrices are
 * row-major 4x4 float32, as the reference's `Camera` fields
 * (renderer/geometry.py:205-278). */
typedef struct jr_scene {
  int64_t tex_base;        /* float offset of the diffuse map [tex_w][tex_h][tex_c] */
  int64_t spec_base;       /* float offset of the specular map [tex_w][tex_h] */
  int64_t nmap_base;       /* float offset of the normal map [nmap_w][nmap_h][3] */
  float world_to_clip[16]; /* Camera.world_to_clip */
  float viewport[16];      /* Camera.viewport */
  float world_to_eye_norm[16]; /* Camera.world_to_eye_norm */
  float shadow_world_to_clip[16]; /* Shadow.camera.world_to_clip */
  float shadow_viewport[16];      /* Shadow.camera.viewport */
  float light_direction[3]; /* LightSource.direction as handed to the shader */
  float light_colour[3];    /* LightSource.colour */
  float light_dir_eye[3];   /* normalised light direction in eye space */
  float ambient[3];
  float diffuse[3];
  float specular[3];
  float shadow_strength[3]; /* Shadow.strength */
  float pad_f;
  int32_t tri_begin;       /* this scene owns faces rows [tri_begin, tri_end) */
  int32_t tri_end;
  int32_t tex_w, tex_h, tex_c;
  int32_t nmap_w, nmap_h;
  int32_t texshape_base;   /* int32-pair offset of MergedModel.texture_shape */
  int32_t texshape_count;  /* number of objects (rows of texture_shape) */
  int32_t tex_offset;      /* MergedModel.offset */
  int32_t spec_w, spec_h;  /* specular map dims (clamp bounds of its reads) */
} jr_scene;

/* Geometry and maps shared by all scenes of a call.  A vertex "slot" is one
 * gl_VertexID of the reference's vertex stage (renderer/pipeline.py:516-518);
 * the per-slot attribute is attr[attr_index[slot]] (identity when the index
 * pointer is NULL).  This expresses Renderer.render's per-corner flattening
 * (renderer/renderer.py:277-296) without materialising it.  All indices are
 * global (already offset per scene). */
typedef struct jr_inputs {
  const float* positions;       /* [*,3] world space */
  const int32_t* pos_index;     /* [n_slots] or NULL */
  const float* normals;         /* [*,3] world space */
  const int32_t* nrm_index;
  const float* uvs;             /* [*,2] */
  const int32_t* uv_index;
  const int32_t* tex_index;     /* [*] texture-map index (MergedModel.texture_index) */
  const int32_t* tex_index_index;
  const float* colours;         /* [*,3] per-vertex colour (gouraud) */
  const int32_t* col_index;
  const int32_t* faces;         /* [n_tris,3] slot ids: pipeline `face_indices`; NULL:
                                   the slots of row t are 3t, 3t+1, 3t+2 (a per-corner
                                   flattening; needs n_slots >= 3 n_tris) */
  const int32_t* id_to_face;    /* darboux: [n_slots] face row per slot */
  const int32_t* face_slots;    /* darboux: [*,3] slot ids per face row */
  const float* textures;        /* flat diffuse maps */
  const float* specular_maps;   /* flat specular maps */
  const float* normal_maps;     /* flat normal maps (darboux) */
  const int32_t* texture_shapes;/* flat [*,2] (phong_reflection*) */
  const float* shadow_maps;     /* [n_scenes][shadow_w][shadow_h] */
  const float* clip;            /* optional [n_slots,4] gl_Position computed by a
                                   caller-side vertex stage; replaces
                                   world_to_clip @ positions when set */
  int32_t n_slots;
  int32_t n_tris;
  int32_t shadow_w;
  int32_t shadow_h;
  /* optional HOST pointer (phong_reflection_shadow only): look the shadow
   * depths up from the shadow pass's binning instead of reading shadow_maps
   * (see jr_shadow_source below); NULL: read shadow_maps */
  const struct jr_shadow_source* shadow_source;
} jr_inputs;

/* Output buffers, [n_scenes][width][height](,3).  Pixels whose final `keep`
 * is false are left untouched unless jr_frame.init_buffers is set
 * (renderer/pipeline.py:401-440: keep ? new : old). */
typedef struct jr_outputs {
  float* zbuffer;   /* required */
  float* canvas;    /* NULL for DEPTH / VISIBILITY */
  int32_t* prim_id; /* optional: chosen face row where written, else -1 */
  float* bary;      /* optional: perspective-correct barycentrics (3) */
  /* optional: [n_scenes][height][width][3] bytes, the display image of the
   * merged canvas: transpose_for_display((clamp(0, canvas, 1) * 255)
   * .astype(uint8)) (renderer/utils.py:72-96, the output step of
   * notebooks/84x84/A100.ipynb cell 7): axes swapped, rows flipped, values
   * truncated toward zero, NaN -> 0.  Written for every pixel. */
  uint8_t* display;
} jr_outputs;

typedef struct jr_frame {
  int64_t bin_capacity;   /* triangle-tile pairs the workspace can hold */
  int32_t width;          /* buffer axis 0 */
  int32_t height;         /* buffer axis 1 */
  int32_t n_scenes;
  int32_t shader;         /* JR_SHADER_* */
  int32_t init_buffers;   /* 1: unwritten pixels take the defaults below */
  int32_t apply_offset;   /* 1: zbuffer += depth_offset everywhere (shadow map) */
  float zbuffer_default;
  float colour_default[3];
  float depth_offset;
  int32_t tile_w;         /* pixels per tile along width (x) */
  int32_t tile_h;         /* pixels per tile along height (y) */
  int32_t persistent;     /* 1: one persistent raster launch draining a work queue */
  int32_t tiles_per_group; /* otherwise: consecutive tiles each workgroup rasterises
                              (0 or 1: one); with one wavefront per tile and at most 4,
                              side by side (a wavefront each), else one after another */
  int32_t pixels_per_lane; /* pixels per raster thread: 0/1, or 4 with 16x16 or 8x32
                              tiles (one wavefront per tile) */
  int32_t part;            /* JR_PART_ALL: every stage; JR_PART_BIN: stages 1-5 (clear,
                              setup, wide, scan, bin) only; JR_PART_RASTER: stage 6 only,
                              on a workspace holding this frame's JR_PART_BIN results */
} jr_frame;

/* jr_frame.part.  Splitting a pass lets a caller overlap one pass's binning
 * with another pass's raster (the main pass's binning does not read the
 * shadow map, only its raster does). */
enum {
  JR_PART_ALL = 0,
  JR_PART_BIN = 1,
  JR_PART_RASTER = 2
};

/* Shadow depths on demand.  The shadow-map pass of renderer/shadow.py:40-125
 * fills every texel of the map, but Shadow.get (renderer/shadow.py:127-153)
 * reads only the texels that shaded pixels project to.  With
 * jr_inputs.shadow_source set, a phong_reflection_shadow pass computes the
 * depth of each texel it reads from the shadow pass's binning -- its
 * small-triangle keys and tile lists, with the same per-texel arithmetic,
 * chooser, fallback and offset as that pass's raster stage -- so the shadow
 * pass runs only its binning stages (jr_frame.part = JR_PART_BIN).  Each depth
 * found is also written to texels[scene][x][y] (unless texels is NULL), so
 * the map is right at every texel a pixel read: jr_backward, given
 * shadow_maps = texels, reads the same values.  The shadow pass must be a
 * depth-only pass in the narrow form (its small-triangle keys kept as 32-bit
 * depth keys, tagged per binning, in its workspace): JR_SHADER_DEPTH,
 * init_buffers = 1, apply_offset = 1, a depth_offset other than -0, no
 * prim_id / bary outputs; and its JR_PART_BIN call must have run on
 * `workspace` (stream ordered before this call).  All pointers but `frame`
 * are device pointers. */
typedef struct jr_shadow_source {
  const jr_frame* frame;    /* the shadow pass's frame (host) */
  const jr_scene* scenes;   /* its n_scenes scene records */
  const void* workspace;    /* its workspace */
  int64_t workspace_bytes;
  const float* keys;        /* not read (may be NULL): the keys stay in the workspace */
  float* texels;            /* [n_scenes][frame->width][frame->height], written;
                               NULL: not kept (a forward-only render) */
  int32_t n_tris;           /* its triangle count */
  int32_t pad;
} jr_shadow_source;

/* Gradient buffers for jr_backward (all optional, accumulated with +=). */
typedef struct jr_grads {
  float* d_positions;   /* [*,3] scatter-added through pos_index */
  float* d_colours;     /* [*,3] through col_index (gouraud) */
  float* d_normals;     /* [*,3] through nrm_index */
  float* d_textures;    /* flat, like jr_inputs.textures */
  float* d_specular_maps;
  float* d_scene;       /* [n_scenes][JR_DSCENE_FLOATS]: see JR_DSCENE_* */
  float* d_uvs;         /* [*,2] through uv_index (phong_darboux's tangent frame) */
  float* d_normal_maps; /* flat, like jr_inputs.normal_maps (phong_darboux) */
} jr_grads;

/* Layout of one row of jr_grads.d_scene. */
enum {
  JR_DSCENE_LIGHT_DIRECTION = 0,  /* 3 */
  JR_DSCENE_LIGHT_COLOUR = 3,     /* 3 */
  JR_DSCENE_LIGHT_DIR_EYE = 6,    /* 3 */
  JR_DSCENE_AMBIENT = 9,          /* 3 */
  JR_DSCENE_DIFFUSE = 12,         /* 3 */
  JR_DSCENE_SPECULAR = 15,        /* 3 */
  JR_DSCENE_WORLD_TO_CLIP = 18,   /* 16 */
  JR_DSCENE_VIEWPORT = 34,        /* 16 */
  JR_DSCENE_WORLD_TO_EYE_NORM = 50, /* 16 */
  JR_DSCENE_SHADOW_STRENGTH = 66, /* 3 */
  JR_DSCENE_FLOATS = 69
};

typedef struct hipStream_t_opaque* jr_stream_t; /* a hipStream_t */

/* ABI version of the loaded library (== JR_ABI_VERSION). */
int32_t jr_version(void);

/* Message of the calling thread's last failed call. */
const char* jr_last_error(void);

/* sizeof of the ABI structs, for binding-side layout checks:
 * 0 jr_scene, 1 jr_inputs, 2 jr_outputs, 3 jr_frame, 4 jr_grads,
 * 5 jr_pose_args, 6 jr_shadow_source, 7 jr_pack_args. */
int64_t jr_sizeof(int32_t which);

/* Bytes of workspace jr_forward needs for this frame (bin_capacity included).
 * Replaces nothing in the reference: JAX sizes its temporaries itself.  A
 * depth pass in the narrow form (JR_SHADER_DEPTH, init_buffers, apply_offset,
 * an offset other than -0, no prim_id / bary) keeps its small-triangle keys
 * in the workspace from call to call, each binning's tagged apart from the
 * older ones, so that it never clears them: zero a newly allocated workspace
 * (memory that held a workspace of the same size before could pass for one
 * whose keys are in order), keep one workspace per stream and do not write
 * to it between calls.  A workspace used by another kind of pass in between,
 * or sized for another pixel count, is detected and its keys are refilled. */
int64_t jr_workspace_size(const jr_frame* frame, int32_t n_tris);

/* Render n_scenes scenes in one call: vertex transform, triangle setup,
 * screen-tile binning, per-tile visibility (default primitive chooser), the
 * shader's interpolate/fragment/mix and the buffer merge.
 * Replaces renderer/pipeline.py:470-537 `render(camera, shader, buffers,
 * face_indices, extra, loop_unroll)` for the built-in shaders, and with
 * frame->shader == JR_SHADER_DEPTH, apply_offset == 1 the shadow-map pass of
 * renderer/shadow.py:40-125.  scenes is a device array of n_scenes structs. */
int32_t jr_forward(const jr_frame* frame, const jr_inputs* inputs,
                   const jr_scene* scenes, jr_outputs* outputs, void* workspace,
                   int64_t workspace_bytes, jr_stream_t stream);

/* jr_forward without a workspace argument and without any allocation: no
 * counterpart in the reference (JAX allocates its own temporaries).  It is
 * for callers that cannot allocate device memory -- code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc.  The
 * workspace is the library's own per-GPU scratch array of JR_SCRATCH_BYTES,
 * resident from the library's load to the end of the process; a frame whose
 * jr_workspace_size exceeds it is rejected with JR_E_WORKSPACE before any
 * launch.  Whole passes only (frame->part must be JR_PART_ALL); stage timing
 * is never recorded.  Every call on a GPU shares that GPU's scratch: calls
 * that overlap in time on one GPU (two streams, two threads) race and must be
 * serialised by the caller.  Each call clears what it reads before writing
 * it, except the narrow form's tagged small-triangle keys (see
 * jr_workspace_size), whose older tags lose to the call's own: no result
 * depends on an earlier call. */
#define JR_SCRATCH_BYTES (64LL << 20)
int32_t jr_forward_noalloc(const jr_frame* frame, const jr_inputs* inputs,
                           const jr_scene* scenes, jr_outputs* outputs,
                           jr_stream_t stream);

/* Number of triangle-tile pairs the last jr_forward on `workspace` binned
 * (device value, read asynchronously into *pairs on `stream`; the caller
 * synchronises).  Used to size bin_capacity. */
int32_t jr_binned_pairs(const void* workspace, const jr_frame* frame,
                        int32_t n_tris, int64_t* pairs_host,
                        jr_stream_t stream);

/* Instrumentation (no counterpart in the reference, whose notebooks time whole
 * jitted calls): when enabled, each jr_forward records HIP events on its own
 * stream around its JR_STAGES stages: 0 clear (k_clear), 1 setup (k_setup),
 * 2 wide (k_wide counting), 3 scan (k_scan_*), 4 bin (k_bin and k_wide
 * binning), 5 raster (k_raster).
 * jr_stage_times waits for them and writes JR_STAGES floats (ms) per call,
 * oldest first, for up to max_calls calls; returns the number written and
 * forgets all recorded calls. */
#define JR_STAGES 6
int32_t jr_stage_timing(int32_t enable);
int32_t jr_stage_times(float* ms, int32_t max_calls);

/* Reverse-mode pass of jr_forward with visibility held fixed (the primitive
 * chosen per pixel, comparisons, floor/round lookups carry no gradient, as
 * under jax.grad).  Takes the upstream gradients of the canvas (and
 * optionally the zbuffer) and the prim_id buffer jr_forward wrote (entries
 * outside [0, n_tris) count as pixels no triangle shows); for every written
 * pixel it recomputes the chosen primitive's edge values,
 * barycentrics and fragment attributes exactly as the forward did and
 * accumulates (+=) gradients into `grads`.  Replaces jax.grad through
 * renderer/pipeline.py:163-328 (edge functions, depth, interpolation), the
 * shaders' vertex and fragment stages and jnp.linalg.inv
 * (renderer/pipeline.py:105).  Supported shaders: all seven built-ins
 * (DEPTH, GOURAUD, GOURAUD_TEXTURE, PHONG, PHONG_DARBOUX, PHONG_REFLECTION,
 * PHONG_REFLECTION_SHADOW); PHONG_DARBOUX also differentiates its tangent
 * frame (renderer/shaders/phong_darboux.py:212-293: the face's NDC corners,
 * their uvs, jnp.linalg.inv of the frame matrix) and its normal-map texel.
 * frame->width/height/n_scenes/shader/tile_w/tile_h must match the forward
 * call; the workspace is sized by jr_backward_workspace_size.  The workspace
 * carries state from call to call (its per-triangle accumulators are left
 * zero, so a call clears only what it used): keep one per stream, do not
 * write to it between calls; a new, resized or foreign workspace is detected
 * by its header and cleared in full. */
int64_t jr_backward_workspace_size(const jr_frame* frame, int32_t n_tris);
/* Message of the calling thread's last failed jr_backward. */
const char* jr_backward_last_error(void);
int32_t jr_backward(const jr_frame* frame, const jr_inputs* inputs,
                    const jr_scene* scenes, const float* d_canvas,
                    const float* d_zbuffer, const int32_t* prim_id,
                    jr_grads* grads, void* workspace, int64_t workspace_bytes,
                    jr_stream_t stream);

/* jr_backward of the image loss (scale / 2) sum((canvas - target)^2): the
 * canvas cotangent of every pixel is scale * (canvas - target) (float32,
 * subtraction first), computed where jr_backward would read d_canvas, so a
 * training step needs no separate loss-gradient kernels or d_canvas buffer.
 * canvas is the forward's output, target [n_scenes][width][height][3].  No
 * counterpart in the reference, where jax.grad differentiates the loss and
 * the renderer together; scale = 2 gives the gradient of sum((canvas -
 * target)^2). */
int32_t jr_backward_l2(const jr_frame* frame, const jr_inputs* inputs,
                       const jr_scene* scenes, const float* canvas, const float* target,
                       float scale, const float* d_zbuffer, const int32_t* prim_id,
                       jr_grads* grads, void* workspace, int64_t workspace_bytes,
                       jr_stream_t stream);

/* Object poses of a batch: the per-object world transform of merge_objects
 * (renderer/model.py:447-555), for every object of every scene in one call,
 * so that a batch whose objects only move is updated on the device.
 * Object o owns model vertices [vert_begin[o], vert_begin[o + 1]) and normals
 * [norm_begin[o], norm_begin[o + 1]); transforms[o] is its row-major 4x4 pose
 * T and scaling[o] its local scaling.  Writes
 *   positions[v] = to_cartesian(T @ [local_verts[v] * scaling[o], 1])
 *     (Camera.apply_pos, renderer/geometry.py:321-346, sums left to right);
 *   normal_matrices[o] = inv(T)^T (renderer/model.py:526, jnp.linalg.inv in
 *     the canonical LU order of the triangle setup);
 *   h = normal_matrices[o] @ [unit_norms[i], 0], normals[i] = h[:3] / s with s
 *     the 2-norm of all the object's h[:3] together (Camera.apply_vec,
 *     renderer/geometry.py:352-386; geometry.normalise, :39-50), its sum of
 *     squares added in the order of numpy's float32 np.sum (8192-value
 *     chunks added in turn, each summed pairwise).
 * unit_norms are the model's normals already divided by their own
 * whole-array norm (the first normalise of apply_vec, which no pose changes).
 * vert_object[v] is the object of vertex v.  normals may be NULL (positions
 * only).  max_object_norms is the largest normal count of any object (at most
 * JR_POSE_MAX_NORMALS, whose pairwise-sum tables take 24 KiB of shared
 * memory; it sizes the normal kernel's shared memory): an object with more
 * normals is left unwritten. */
#define JR_POSE_MAX_NORMALS 43690
typedef struct jr_pose_args {
  int32_t n_objects;
  int32_t n_verts;
  int32_t n_norms;
  int32_t max_object_norms;
  const int32_t* vert_begin;  /* [n_objects + 1]; optional, not read: vert_object
                                 carries the same ownership per vertex */
  const int32_t* norm_begin;  /* [n_objects + 1] */
  const int32_t* vert_object; /* [n_verts] */
  const float* local_verts;   /* [n_verts, 3] */
  const float* unit_norms;    /* [n_norms, 3] */
  const float* scaling;       /* [n_objects, 3] */
  const float* transforms;    /* [n_objects, 16] */
  float* normal_matrices;     /* [n_objects, 16], written */
  float* positions;           /* [n_verts, 3], written */
  float* normals;             /* [n_norms, 3], written; NULL: skipped */
} jr_pose_args;
int32_t jr_pose(const jr_pose_args* args, jr_stream_t stream);

/* The per-corner flattening of a batch (renderer/renderer.py:277-296,
 * Renderer.render's position = verts[faces.ravel()] and likewise normals, uvs
 * and texture_index[faces_uv.ravel()]) for scenes whose rows are concatenated:
 * writes the global index arrays jr_inputs.pos_index / nrm_index / uv_index /
 * tex_index from each scene's local face tables and the offsets of its rows.
 * Replaces the index arithmetic a batch planner does per scene on the host.
 * The *_off arrays hold n_scenes + 1 ascending int64 offsets (first 0): scene
 * s owns face rows [tri_off[s], tri_off[s + 1]) (3 corners each, in faces,
 * faces_norm, faces_uv), vertex rows [vert_off[s], vert_off[s + 1]), and so on.
 * A local index outside its scene's rows sets *error to 1 (the caller
 * synchronises and reports it) and is written clamped into the scene's rows.
 * tex_index may be NULL (then texidx_off and texture_index are not read). */
typedef struct jr_pack_args {
  int32_t n_scenes;
  int32_t pad;
  int64_t n_faces;            /* total face rows */
  const int64_t* tri_off;     /* [n_scenes + 1] */
  const int64_t* vert_off;    /* [n_scenes + 1] */
  const int64_t* norm_off;    /* [n_scenes + 1] */
  const int64_t* uv_off;      /* [n_scenes + 1] */
  const int64_t* texidx_off;  /* [n_scenes + 1] rows of texture_index */
  const int32_t* faces;       /* [3 n_faces] local vertex indices */
  const int32_t* faces_norm;  /* [3 n_faces] local normal indices */
  const int32_t* faces_uv;    /* [3 n_faces] local uv indices */
  const int32_t* texture_index; /* per-vertex texture map index, concatenated */
  int32_t* pos_index;         /* [3 n_faces], written */
  int32_t* nrm_index;         /* [3 n_faces], written */
  int32_t* uv_index;          /* [3 n_faces], written */
  int32_t* tex_index;         /* [3 n_faces], written; optional */
  int32_t* error;             /* one word, OR-ed with 1 on a bad index */
} jr_pack_args;
int32_t jr_pack(const jr_pack_args* args, jr_stream_t stream);
/* Message of the calling thread's last failed jr_pack. */
const char* jr_pack_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* JRHIP_H */
