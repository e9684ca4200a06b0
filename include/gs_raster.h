/* gs_raster.h -- C ABI of libgs_hip.so, the MI355X (gfx950) Gaussian rasteriser.
 *
 * Drop-in native layer beneath the Python packages `diff_gaussian_rasterization`
 * and `simple_knn`, the two un-vendored submodules CF-3DGS imports
 * (reference .gitmodules:1-6; imported at scene/gaussian_model_cf.py:20,28-31).
 * Each entry point names the reference interface it replaces.
 *
 * Conventions
 *  - Every pointer is a DEVICE pointer owned by the caller (e.g. torch
 *    tensor data_ptr()); the library allocates nothing on these paths.
 *  - Work is enqueued on `stream` (a hipStream_t passed as void*); calls return
 *    after enqueueing.  No host synchronisation happens inside the library.
 *  - Return value 0 = GS_OK, otherwise a GS_E_* code (argument checks are
 *    done on the host before anything is enqueued).  Conditions only the
 *    device can see (binning capacity exceeded, a culled point under
 *    `prefiltered`) are reported in the device status word counters[1].
 *  - Matrices are the 16 floats of the reference's torch tensors in memory
 *    order: viewmatrix = world_view_transform, projmatrix =
 *    full_proj_transform (scene/cameras.py:58-73), i.e. column-major.
 */
#ifndef GS_RASTER_H
#define GS_RASTER_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GS_OK 0
#define GS_E_ARG 1       /* bad argument (shape, null pointer, size) */
#define GS_E_WORKSPACE 2 /* workspace smaller than the layout requires */
#define GS_E_LAUNCH 3    /* the runtime refused a launch */
#define GS_E_STATIC_LIMIT 4 /* more items than an entry point's static scratch holds */
#define GS_E_SCRATCH 5      /* the device address of a static scratch could not be obtained */

/* counters: uint32[4] on the device.
 *   [0] num_rendered of the last preprocess
 *   [1] status of the last frame (cleared by every preprocess)
 *   [2] scratch (element count of the depth sort)
 *   [3] sticky status: capacity overflows OR-ed in and never cleared by the
 *       library, so a caller replaying many frames (e.g. a captured graph)
 *       clears it once and checks it once at the end. */
#define GS_STATUS_CAPACITY 1u    /* num_rendered exceeded the binning capacity */
#define GS_STATUS_PREFILTERED 2u /* a point was near-culled while prefiltered=1 */

#define GS_TILE 16

/* One camera / frame.  Mirrors GaussianRasterizationSettings
 * (constructed at scene/gaussian_model_cf.py:811-824), plus an optional rigid
 * pose applied to the means, which fuses the reference's
 * `P[i].retr().act(xyz)` (scene/gaussian_model_cf.py:123-133) into preprocess. */
typedef struct gs_frame {
  int width, height;
  float tanfovx, tanfovy;
  float scale_modifier;
  int sh_degree;          /* active SH degree D, (D+1)^2 <= M */
  int prefiltered;
  const float* viewmatrix; /* [16] */
  const float* projmatrix; /* [16] */
  const float* campos;     /* [3] */
  const float* bg;         /* [3] */
  const float* pose;       /* [7] = tx,ty,tz,qx,qy,qz,qw (lietorch SE3 layout) or NULL */
  int binning;             /* GS_BINNING_DEPTH_FIRST (default) or GS_BINNING_SORT64 */
  int deterministic;       /* backward only: 0 = per-Gaussian gradients summed with float
                              atomics (order varies run to run, as the reference's);
                              1 = every reduction in a fixed order, so gradients are
                              bitwise reproducible (needs the larger scratch below) */
} gs_frame_t;

/* Two binning schedules with identical results (same sorted Gaussian ids and
 * tile ranges):
 *  SORT64       the reference's structure: duplicate (tile << 32 | depth bits)
 *               keys in Gaussian order, stable radix sort on 32 + bits(T) bits;
 *  DEPTH_FIRST  stable sort of the visible Gaussians by depth bits, duplicate
 *               tile ids in that order, stable sort on bits(T) bits -- equal
 *               (tile, depth) keys keep Gaussian order in both. */
#define GS_BINNING_DEPTH_FIRST 0
#define GS_BINNING_SORT64 1

/* The Gaussians: exactly one of shs / colors_precomp and exactly one of
 * (scales, rotations) / cov3D_precomp, as GaussianRasterizer.forward demands. */
typedef struct gs_gaussians {
  int P;
  int M;                     /* SH coefficients per Gaussian in shs */
  const float* means3D;      /* [P,3] */
  const float* shs;          /* [P,M,3] or NULL */
  const float* colors_precomp; /* [P,3] or NULL */
  const float* opacities;    /* [P] */
  const float* scales;       /* [P,3] or NULL */
  const float* rotations;    /* [P,4] (w,x,y,z) or NULL */
  const float* cov3D_precomp; /* [P,6] or NULL */
} gs_gaussians_t;

/* Gradients written by gs_backward (every entry is written, no accumulation).
 * Any per-Gaussian output may be NULL when it is not wanted (a pose-only
 * solve passes only dL_dpose). */
typedef struct gs_grads {
  float* dL_dmeans2D;   /* [P,3] NDC-space screen mean gradient (z = 0) */
  float* dL_dcolors;    /* [P,3] gradient w.r.t. the blended per-Gaussian colour */
  float* dL_dopacity;   /* [P] */
  float* dL_dmeans3D;   /* [P,3] (w.r.t. the means before the pose, if any) */
  float* dL_dcov3D;     /* [P,6] */
  float* dL_dsh;        /* [P,M,3] or NULL */
  float* dL_dscales;    /* [P,3] or NULL */
  float* dL_drotations; /* [P,4] or NULL */
  float* dL_dpose;      /* [7] or NULL (requires frame.pose) */
} gs_grads_t;

/* Byte offsets of the sub-buffers of the three workspaces. */
typedef struct gs_geom_layout {
  size_t splat;    /* float4[3P]: (x,y,conic_a,conic_b) (conic_c,opacity,depth,clamp bits) (r,g,b,0) */
  size_t cov3D;    /* float[6P] */
  size_t tiles;    /* uint32[P] tiles touched */
  size_t offsets;  /* uint32[P] inclusive prefix sum of tiles within each 256-Gaussian block */
  size_t scan_tmp; /* uint32[ceil(P/256)] exclusive prefix of the block totals */
  size_t total;
} gs_geom_layout_t;

typedef struct gs_binning_layout {
  size_t keys_dup, vals_dup;   /* [capacity] duplicated keys (uint64 tile|depth, or uint32 tile), Gaussian ids */
  size_t keys_alt, vals_alt;   /* ping-pong buffers of the radix sort */
  size_t keys_sorted, vals_sorted; /* whichever of the above holds the sorted result */
  size_t hist;                 /* sort digit histograms */
  int sort_passes;             /* 8-bit LSD passes of the list sort */
  int key_bytes;               /* 8 (SORT64) or 4 (DEPTH_FIRST) */
  int id_bits;                 /* DEPTH_FIRST with bits(P) + bits(n_tiles) <= 32: the list entries are
                                  single words tile << id_bits | Gaussian id, sorted in place of the keys
                                  (vals_sorted == keys_sorted); 0: ids in their own array */
  size_t dkeys_a, dvals_a, dkeys_b, dvals_b; /* uint32[P] depth sort (DEPTH_FIRST) */
  size_t dscan, dblocks;       /* uint32[P] tiles prefix in depth order, per-256 block sums */
  size_t drect;                /* uint32[2P] tile rect per Gaussian: x0 | y0 << 16, x1 | y1 << 16 (DEPTH_FIRST) */
  size_t ckpt;                 /* float[capacity/256 + 1][5][256] blend state at list batch boundaries */
  size_t bweight;              /* uint32[capacity/256 + 2 n_tiles + 2] per (tile, list batch): the blend's
                                  longest per-wave list, the backward's estimate of that chunk's cost */
  size_t items;                /* uint32[1 + 2 (capacity/256 + n_tiles + 1)] backward work items:
                                  count, then (tile, chunk) pairs, longest estimated first */
  size_t item_log;             /* same size: the blend's record of the work items, in completion
                                  order: count, then (tile, chunk << 6 | cost class) */
  size_t total;
} gs_binning_layout_t;

typedef struct gs_image_layout {
  size_t ranges;    /* uint32[2*n_tiles] [start, end) into the sorted list */
  size_t final_T;   /* float[H*W] */
  size_t n_contrib; /* uint32[H*W] index (1-based) of the last contributor */
  size_t accum;     /* float[4][H*W] colour sums before the background term, depth */
  size_t ckpt_base; /* uint32[n_tiles+1] first checkpoint of each tile */
  size_t tile_maxc; /* uint32[n_tiles] largest contributor index in the tile */
  size_t tile_order; /* uint32[n_tiles] the blend's tile order, longest list first */
  size_t total;
} gs_image_layout_t;

int gs_geom_layout(int P, gs_geom_layout_t* out);
int gs_binning_layout(long long capacity, int P, int width, int height, int binning, gs_binning_layout_t* out);
int gs_image_layout(int width, int height, gs_image_layout_t* out);
/* Backward scratch for P Gaussians; a deterministic backward also keeps one
 * 10-float partial per (Gaussian, tile) pair, so it needs the binning capacity. */
size_t gs_backward_scratch_bytes(int P, long long capacity, int deterministic);

/* Forward, phase 1 (replaces the preprocess + InclusiveSum part of
 * diff_gaussian_rasterization._C.rasterize_gaussians, called from
 * _RasterizeGaussians.forward; reached at scene/gaussian_model_cf.py:873-882).
 * Writes radii[P] and the geometry workspace; counters[0] = num_rendered
 * (sum of tiles touched), counters[1] = status (cleared, then flagged). */
int gs_preprocess(const gs_frame_t* frame, const gs_gaussians_t* g, int* radii, void* geom,
                  size_t geom_bytes, uint32_t* counters, void* stream);

/* Forward, phase 2 (replaces duplicateWithKeys + SortPairs + identifyTileRanges
 * + render of the same call).  Reads num_rendered from counters[0]; if it
 * exceeds `capacity` nothing is binned and counters[1] |= GS_STATUS_CAPACITY.
 * out_color [3,H,W], out_depth [H,W], out_alpha [H,W] (out_alpha may be NULL). */
int gs_render(const gs_frame_t* frame, int P, const void* geom, uint32_t* counters, long long capacity,
              void* binning, size_t binning_bytes, void* image, size_t image_bytes, float* out_color,
              float* out_depth, float* out_alpha, void* stream);

/* gs_render split at the num_rendered readback (no reference counterpart; the
 * reference waits for the readback before it bins anything).
 * gs_render_prepare queues everything that does not read num_rendered: the
 * tile ranges are cleared and, in the depth-first schedule, the Gaussians are
 * put in depth order and their tile counts prefix-summed in that order.  A
 * caller that guesses `capacity` (e.g. from the previous frame) queues it
 * before it waits for counters[0], so the GPU has work while the host reads
 * the count.  gs_render_resume, with the same arguments, then does the rest
 * of gs_render; if counters[0] turned out larger than the guess, the caller
 * calls gs_render with a larger capacity and a workspace of that size instead.
 * prepare + resume == gs_render. */
int gs_render_prepare(const gs_frame_t* frame, int P, const void* geom, uint32_t* counters, long long capacity,
                      void* binning, size_t binning_bytes, void* image, size_t image_bytes, void* stream);
int gs_render_resume(const gs_frame_t* frame, int P, const void* geom, uint32_t* counters, long long capacity,
                     void* binning, size_t binning_bytes, void* image, size_t image_bytes, float* out_color,
                     float* out_depth, float* out_alpha, void* stream);

/* Backward (replaces diff_gaussian_rasterization._C.rasterize_gaussians_backward,
 * triggered by loss.backward() at trainer/cf3dgs_trainer.py:133).  dL_ddepth and
 * dL_dalpha may be NULL (treated as zero).  scratch: gs_backward_scratch_bytes(P,
 * capacity, frame->deterministic).
 * The image workspace is the forward's and is only read; in the binning
 * workspace the backward writes its work-item table (items) and reads the
 * rest. */
int gs_backward(const gs_frame_t* frame, const gs_gaussians_t* g, const int* radii, const void* geom,
                void* binning, long long capacity, const void* image, const uint32_t* counters,
                const float* dL_dcolor, const float* dL_ddepth, const float* dL_dalpha, const gs_grads_t* out,
                void* scratch, size_t scratch_bytes, void* stream);

/* Adam for the means and the SH coefficients applied inside the backward
 * (no reference counterpart: the reference runs loss.backward() and then
 * optimizer.step(), scene/gaussian_model_cf.py:256-268, trainer/cf3dgs_trainer.py:81-173).
 * gs_backward_adam is gs_backward followed by one Adam step of the means and
 * the SH coefficients (torch.optim.Adam, the update of gs_adam_step_rows, bit
 * for bit), except that those two gradients are never written: each
 * Gaussian's gradient goes straight into its Adam step.  g->means3D and g->shs
 * are updated in place; out->dL_dmeans3D and out->dL_dsh must be NULL.  The
 * frame must have no pose, and M <= 16.  Step count `step` >= 1 (bias
 * corrections 1 - beta^step, computed as gs_adam_step_rows does).  SH element
 * e of the [P, M, 3] tensor steps with shs_lead_lr when e % (3 M) < shs_lead,
 * else with shs_lr.  All pointers 16-byte aligned. */
typedef struct gs_adam_fused {
  float* means_m;
  float* means_v;
  float means_lr;
  float* shs_m;
  float* shs_v;
  float shs_lr;
  float shs_lead_lr;
  int shs_lead;
  float beta1, beta2, eps;
  int step;
} gs_adam_fused_t;
int gs_backward_adam(const gs_frame_t* frame, const gs_gaussians_t* g, const int* radii, const void* geom,
                     void* binning, long long capacity, const void* image, const uint32_t* counters,
                     const float* dL_dcolor, const float* dL_ddepth, const float* dL_dalpha, const gs_grads_t* out,
                     const gs_adam_fused_t* adam, void* scratch, size_t scratch_bytes, void* stream);

/* ---- Many frames per launch (the batched pose solver, the serving loop) ----
 * No single reference counterpart: these split gs_render / gs_backward so that
 * the per-pixel kernels of many independent frames (each with its own
 * workspaces, as above) run as ONE launch that fills the device.  A "set" is
 * a device buffer describing every frame's workspaces; it is built on the host
 * once (the *_build calls write a host buffer of *_bytes(n) bytes, which the
 * caller copies to device memory) and stays valid while those workspaces do.
 *
 *   gs_render  = gs_render_binning (per frame)  + blend
 *   gs_backward = gs_backward_prepare (per frame) + render backward
 *                 + gs_backward_finish (per frame)
 * with the middle step of every frame replaced by one gs_multi_blend /
 * gs_multi_render_bwd launch over the set.  total_tiles / total_blocks are
 * returned by the builds.  All frames of a render-backward set share one
 * `deterministic` setting.  gs_multi_render_bwd's `flags`: bit 0 (1) the
 * fixed-order reductions; GS_RB_COLOR_ONLY when no frame of the set has a
 * depth or alpha cotangent (their terms are then skipped; a frame that has
 * one must not be launched with it); GS_RB_POSE_ONLY (with 1 and
 * GS_RB_COLOR_ONLY) when gs_backward_is_pose_only holds for the frames: the
 * pose-only backward, whose prepare / finish steps gs_backward_prepare and
 * gs_backward_finish then take by themselves.
 * gs_backward_is_pose_only: 1 when gs_backward (and its prepare / finish
 * steps) would run the pose-only backward for these arguments -- fixed-order
 * reductions, only out->dL_dpose set, no depth or alpha cotangent, active SH
 * degree 0 or precomputed colours.  It sums five terms per (Gaussian, tile)
 * pair instead of ten; the pose gradient has the same bits as the full
 * backward's. */
#define GS_RB_COLOR_ONLY 2
#define GS_RB_POSE_ONLY 4
int gs_render_binning(const gs_frame_t* frame, int P, const void* geom, uint32_t* counters, long long capacity,
                      void* binning, size_t binning_bytes, void* image, size_t image_bytes, void* stream);
size_t gs_multi_blend_bytes(int n);
int gs_multi_blend_build(int n, const gs_frame_t* frames, const int* P, const void* const* geom,
                         const long long* capacity, void* const* binning, void* const* image,
                         float* const* out_color, float* const* out_depth, float* const* out_alpha,
                         void* host_set, size_t set_bytes, unsigned* total_tiles);
int gs_multi_blend(int n, const void* dev_set, unsigned total_tiles, void* stream);
/* The depth-first binning of many frames around ONE depth sort: per frame
 * gs_render_depth_keys (the depth keys), then gs_multi_depth_sort over the
 * set (each frame's 32-bit depth sort on its own buffers, with the same
 * passes and result as gs_render_binning's, but each of the 4 x 3 pass
 * launches covers every frame), then per frame
 * gs_render_binning_after_depth_sort (tile-count prefix, duplicate, list
 * sort, checkpoint scan) or gs_render_after_depth_sort (the same and the
 * blend).  Together they leave exactly what gs_render_binning / gs_render
 * leave.  Frames must use GS_BINNING_DEPTH_FIRST with P in (0, 2^21] and a
 * positive capacity; binning_bytes are the frames' binning workspace sizes.
 * Replaces, for many frames at once, the key sort of forward phase 2 above
 * (the upstream rasteriser's SortPairs of the point list, reached from
 * scene/gaussian_model_cf.py:873-882 once per frame). */
int gs_render_depth_keys(const gs_frame_t* frame, int P, const void* geom, uint32_t* counters, long long capacity,
                         void* binning, size_t binning_bytes, void* image, size_t image_bytes, void* stream);
size_t gs_multi_depth_sort_bytes(int n);
int gs_multi_depth_sort_build(int n, const gs_frame_t* frames, const int* P, const long long* capacity,
                              void* const* binning, const size_t* binning_bytes, uint32_t* const* counters,
                              void* host_set, size_t set_bytes, unsigned* blocks);
int gs_multi_depth_sort(int n, const void* dev_set, unsigned blocks, void* stream);
/* gs_preprocess followed by gs_render_depth_keys for many frames, in three
 * launches (every frame's preprocess; every frame's tile-count scan, one
 * workgroup each; every frame's depth keys), with the results of the
 * per-frame calls, bit for bit.  The build takes the per-frame arguments of
 * the two calls as arrays, checks them as those do (frames with P > 0 on
 * GS_BINNING_DEPTH_FIRST and a positive capacity, as gs_render_depth_keys
 * requires) and returns launch[3]: the two per-Gaussian launches' workgroup
 * counts and the preprocess launch's LDS bytes, which gs_multi_preprocess_keys
 * takes back.  Replaces, for many frames at once, the per-frame preprocess of
 * forward phase 1 and the key step of phase 2 above. */
size_t gs_multi_preprocess_keys_bytes(int n);
int gs_multi_preprocess_keys_build(int n, const gs_frame_t* frames, const gs_gaussians_t* g, int* const* radii,
                                   void* const* geom, const size_t* geom_bytes, uint32_t* const* counters,
                                   const long long* capacity, void* const* binning, const size_t* binning_bytes,
                                   void* const* image, const size_t* image_bytes, void* host_set, size_t set_bytes,
                                   unsigned* launch);
int gs_multi_preprocess_keys(int n, const void* dev_set, const unsigned* launch, void* stream);
int gs_render_binning_after_depth_sort(const gs_frame_t* frame, int P, const void* geom, uint32_t* counters,
                                       long long capacity, void* binning, size_t binning_bytes, void* image,
                                       size_t image_bytes, void* stream);
int gs_render_after_depth_sort(const gs_frame_t* frame, int P, const void* geom, uint32_t* counters,
                               long long capacity, void* binning, size_t binning_bytes, void* image,
                               size_t image_bytes, float* out_color, float* out_depth, float* out_alpha,
                               void* stream);
/* The tile-list sort shared the same way: after the shared depth sort, per
 * frame gs_render_binning_duplicate (tile-count prefix and duplicate), then
 * gs_multi_list_sort over the set (each frame's list sort on its tile bits,
 * with its own pass count, leaving the list and the tile ranges where
 * gs_render_binning does; frames with one-word list entries only, capacity
 * up to 2^21), then per frame gs_render_binning_after_list_sort (checkpoint
 * scan) or gs_render_after_list_sort (the same and the blend).  blocks[4]
 * holds the grid of each pass (0 past the last). */
int gs_render_binning_duplicate(const gs_frame_t* frame, int P, const void* geom, uint32_t* counters,
                                long long capacity, void* binning, size_t binning_bytes, void* image,
                                size_t image_bytes, void* stream);
size_t gs_multi_list_sort_bytes(int n);
int gs_multi_list_sort_build(int n, const gs_frame_t* frames, const int* P, const long long* capacity,
                             void* const* binning, const size_t* binning_bytes, uint32_t* const* counters,
                             void* const* image, void* host_set, size_t set_bytes, unsigned* blocks);
int gs_multi_list_sort(int n, const void* dev_set, const unsigned* blocks, void* stream);
/* The depth-first binning's steps on either side of the list sort for many
 * frames: gs_multi_binning_duplicate is gs_render_binning_duplicate of every
 * frame (three launches: the workgroup-local scans of the tile counts in
 * depth order, one workgroup per frame for the block prefix, the
 * duplicates); gs_multi_binning_after_list_sort is
 * gs_render_binning_after_list_sort of every frame (one launch, a workgroup
 * per frame).  The results are those of the per-frame calls, bit for bit.
 * The build takes those calls' per-frame arguments as arrays and checks them
 * as they do; total_blocks is returned by the build.  Replaces, for many
 * frames at once, the duplicate and the tile-range step of forward phase 2
 * above. */
size_t gs_multi_binning_steps_bytes(int n);
int gs_multi_binning_steps_build(int n, const gs_frame_t* frames, const int* P, const void* const* geom,
                                 uint32_t* const* counters, const long long* capacity, void* const* binning,
                                 const size_t* binning_bytes, void* const* image, const size_t* image_bytes,
                                 void* host_set, size_t set_bytes, unsigned* total_blocks);
int gs_multi_binning_duplicate(int n, const void* dev_set, unsigned total_blocks, void* stream);
int gs_multi_binning_after_list_sort(int n, const void* dev_set, void* stream);
int gs_render_binning_after_list_sort(const gs_frame_t* frame, int P, const void* geom, uint32_t* counters,
                                      long long capacity, void* binning, size_t binning_bytes, void* image,
                                      size_t image_bytes, void* stream);
int gs_render_after_list_sort(const gs_frame_t* frame, int P, const void* geom, uint32_t* counters,
                              long long capacity, void* binning, size_t binning_bytes, void* image,
                              size_t image_bytes, float* out_color, float* out_depth, float* out_alpha,
                              void* stream);
int gs_backward_prepare(const gs_frame_t* frame, const gs_gaussians_t* g, const int* radii, const void* geom,
                        void* binning, long long capacity, const void* image, const uint32_t* counters,
                        const float* dL_dcolor, const float* dL_ddepth, const float* dL_dalpha,
                        const gs_grads_t* out, void* scratch, size_t scratch_bytes, void* stream);
int gs_backward_finish(const gs_frame_t* frame, const gs_gaussians_t* g, const int* radii, const void* geom,
                       void* binning, long long capacity, const void* image, const uint32_t* counters,
                       const float* dL_dcolor, const float* dL_ddepth, const float* dL_dalpha,
                       const gs_grads_t* out, void* scratch, size_t scratch_bytes, void* stream);
size_t gs_multi_render_bwd_bytes(int n);
int gs_multi_render_bwd_build(int n, const gs_frame_t* frames, const int* P, const void* const* geom,
                              const long long* capacity, void* const* binning, const void* const* image,
                              const float* const* dL_dcolor, const float* const* dL_ddepth,
                              const float* const* dL_dalpha, void* const* scratch, void* host_set,
                              size_t set_bytes, unsigned* total_blocks);
int gs_multi_render_bwd(int n, const void* dev_set, unsigned total_blocks, int flags, void* stream);
int gs_backward_is_pose_only(const gs_frame_t* f, const gs_gaussians_t* g, const gs_grads_t* out,
                             int has_depth_or_alpha);
/* The set-up and finish steps of many pose-only frames, one launch for the
 * set-up and two for the finish: gs_multi_pose_bwd_prepare is
 * gs_backward_prepare of every frame; gs_multi_pose_bwd_finish computes every
 * frame's per-Gaussian pose partials (one launch over all frames' Gaussians),
 * then each frame's pose reduction into out[i].dL_dpose (one workgroup per
 * frame).  The results are those of the per-frame calls, bit for bit.  The
 * build takes the per-frame arguments of gs_backward_prepare / _finish as
 * arrays (no depth or alpha cotangent), checks them as those do, and refuses
 * (GS_E_ARG) a frame for which gs_backward_is_pose_only(frame, g, out, 0) does
 * not hold; it returns blocks[2] (set-up and per-Gaussian workgroups), which
 * the two launches take back.  The backward of the pose-fitting loop,
 * trainer/cf3dgs_trainer.py:261-285, for many frames at once. */
size_t gs_multi_pose_bwd_bytes(int n);
int gs_multi_pose_bwd_build(int n, const gs_frame_t* frames, const gs_gaussians_t* g, const int* const* radii,
                            const void* const* geom, void* const* binning, const long long* capacity,
                            const void* const* image, const uint32_t* const* counters,
                            const float* const* dL_dcolor, const gs_grads_t* out, void* const* scratch,
                            const size_t* scratch_bytes, void* host_set, size_t set_bytes, unsigned* blocks);
int gs_multi_pose_bwd_prepare(int n, const void* dev_set, const unsigned* blocks, void* stream);
int gs_multi_pose_bwd_finish(int n, const void* dev_set, const unsigned* blocks, void* stream);

/* Near-plane visibility (replaces diff_gaussian_rasterization._C.mark_visible,
 * GaussianRasterizer.markVisible). visible[P] = 1 if the view-space z > 0.2. */
int gs_mark_visible(const float* means3D, int P, const float* viewmatrix, const float* projmatrix,
                    uint8_t* visible, void* stream);

/* Stable LSD radix sort of (uint64 key, uint32 value) pairs on bits
 * [0, end_bit), n read from *n_dev (<= capacity).  Input in (keys_a, vals_a);
 * the result lands in a or b: *result_in_b is set on the host (depends only on
 * end_bit).  tmp: gs_sort_tmp_bytes(capacity). */
size_t gs_sort_tmp_bytes(long long capacity);
int gs_sort_pairs_u64(uint64_t* keys_a, uint32_t* vals_a, uint64_t* keys_b, uint32_t* vals_b,
                      const uint32_t* n_dev, long long capacity, int end_bit, void* tmp, size_t tmp_bytes,
                      int* result_in_b, void* stream);
int gs_sort_pairs_u32(uint32_t* keys_a, uint32_t* vals_a, uint32_t* keys_b, uint32_t* vals_b,
                      const uint32_t* n_dev, long long capacity, int end_bit, void* tmp, size_t tmp_bytes,
                      int* result_in_b, void* stream);

/* Mean squared distance to the 3 nearest other points (replaces
 * simple_knn._C.distCUDA2, called at scene/gaussian_model_cf.py:194-196).
 * Exact (same value as a brute-force search).  ws: gs_knn_workspace_bytes(P). */
size_t gs_knn_workspace_bytes(int P);
int gs_knn3_meandist(const float* points, int P, float* out, void* ws, size_t ws_bytes, void* stream);

/* gs_knn3_meandist without a workspace argument: no counterpart in the
 * reference.  It is for callers that cannot allocate device memory (code
 * under graph capture, serving loops).  Its workspace is a static device
 * array of the library (about 24 MiB, one per GPU, held until the process
 * ends), sized for at most GS_KNN_STATIC_MAX_POINTS points; more returns
 * GS_E_STATIC_LIMIT before anything is enqueued.  It allocates nothing.
 * Every call on one GPU shares that GPU's array: calls that overlap on one
 * GPU (two streams, two threads) race and must be serialised by the caller.
 * The stream must belong to the current device.  Same result as
 * gs_knn3_meandist, bit for bit. */
#define GS_KNN_STATIC_MAX_POINTS (1 << 20)
int gs_knn3_meandist_static(const float* points, int P, float* out, void* stream);

/* Frame ingest: a depth map and its image become a frame's initial Gaussians
 * on the device.  The reference does this on the host: kornia depth_to_3d,
 * numpy, open3d voxel_down_sample(0.01), then create_from_pcd
 * (trainer/trainer.py:389-423,476-512; scene/gaussian_model_cf.py:178-212).
 *
 * Colour sources are per point (pixel i = point i of a lifted frame).  Value
 * (point i, channel c) sits at element i*color_pix_stride + c*color_chan_stride
 * of a uint8 array (colors_u8 = 1; v becomes v / 255.0 in double, as
 * np.asarray(image) / 255.0) or of a float32 array (colors_u8 = 0).
 *
 * gs_depth_to_points:   kornia depth_to_3d(normalize_points=False) over an
 *                       [H, W] depth.  points[i = v*W + u] =
 *                       (((u - cx) / fx) * d, ((v - cy) / fy) * d, d) in
 *                       float32.  When `colors` is non-NULL it also gets the
 *                       source colours as float32 [H*W, 3].
 * gs_points_bounds:     bounds6 (device) = min xyz, max xyz of [N, 3] points;
 *                       all NaN when a coordinate is not finite.
 *                       ws: gs_bounds_workspace_bytes().
 * gs_voxel_down_sample: open3d VoxelDownSample.  voxel_min3 (HOST, double) is
 *                       min - voxel_size / 2.  The index per axis is
 *                       floor((p - voxel_min) / voxel_size) in double, and
 *                       must fit bits_x / bits_y / bits_z bits (total <= 64).
 *                       Each occupied voxel's mean point and mean colour
 *                       (double sums in point order, then float32) go to
 *                       out_points / out_colors ([N, 3] capacity), in
 *                       ascending (ix, iy, iz) order; *n_out (device) = the
 *                       number of voxels.  ws: gs_voxel_workspace_bytes(N).
 * gs_create_from_pcd:   create_from_pcd's raw parameters from points, colours
 *                       and gs_knn3_meandist's dist2: xyz; features [P, n_coef,
 *                       3] (row 0 = RGB2SH(colour), the rest 0); scaling =
 *                       log(sqrt(max(dist2, 1e-7))) x3; rotation (1, 0, 0, 0);
 *                       opacity = inverse_sigmoid(0.1). */
int gs_depth_to_points(int W, int H, const float* depth, float fx, float fy, float cx, float cy, float* points,
                       const void* colors_src, int colors_u8, long long color_pix_stride,
                       long long color_chan_stride, float* colors, void* stream);
size_t gs_bounds_workspace_bytes(void);
int gs_points_bounds(const float* points, long long N, float* bounds6, void* ws, size_t ws_bytes, void* stream);
size_t gs_voxel_workspace_bytes(long long N);
int gs_voxel_down_sample(const float* points, long long N, const void* colors_src, int colors_u8,
                         long long color_pix_stride, long long color_chan_stride, const double* voxel_min3,
                         double voxel_size, int bits_x, int bits_y, int bits_z, float* out_points,
                         float* out_colors, uint32_t* n_out, void* ws, size_t ws_bytes, void* stream);
int gs_create_from_pcd(int P, int n_coef, const float* points, const float* colors, const float* dist2, float* xyz,
                       float* features, float* scaling, float* rotation, float* opacity, void* stream);

/* Fused photometric loss of the training step (replaces Loss.forward's
 * (1-l)*l1_loss + l*(1 - SSIM_V2) in torch, trainer/losses.py:98-136 and
 * :164-251; l = lambda_dssim = 0.2 at arguments/__init__.py:113).  pred and
 * target are [C,H,W]; window11 is a HOST array of the 11 separable window
 * taps (the normalised 1-D Gaussian, sigma 1.5, of trainer/losses.py:143-159).
 * out3 (device) = {loss, mean SSIM, mean |pred - target|}.  The workspace keeps
 * the per-pixel SSIM partials for the backward, which writes
 * dpred = grad_out[0] * d loss / d pred. */
size_t gs_loss_workspace_bytes(int C, int H, int W);
int gs_ssim_l1_forward(const float* pred, const float* target, int C, int H, int W, const float* window11,
                       float lambda_dssim, int clamp01, float* out3, void* ws, size_t ws_bytes, void* stream);
int gs_ssim_l1_backward(const float* pred, const float* target, int C, int H, int W, const float* window11,
                        float lambda_dssim, int clamp01, const float* grad_out, const void* ws, size_t ws_bytes,
                        float* dpred, void* stream);
/* (clamp01 != 0: the loss of clamp(pred, 0, 1), the caller's image.clamp(0, 1)
 * at scene/gaussian_model_cf.py:885, with clamp's gradient mask.) */
/* gs_ssim_l1_forward then gs_ssim_l1_backward in two launches instead of
 * three, for callers that do not look at the loss before the backward (the
 * batched pose solve; the reference's loop checks torch.isnan(loss) in
 * between, trainer/losses.py:126): out3 and dpred are the same bits. */
int gs_ssim_l1_forward_backward(const float* pred, const float* target, int C, int H, int W, const float* window11,
                                float lambda_dssim, int clamp01, float* out3, const float* grad_out, void* ws,
                                size_t ws_bytes, float* dpred, void* stream);
/* gs_ssim_l1_forward_backward of n images (each with its own size, buffers,
 * workspace and out3; one window, lambda, clamp and grad_out for all) in one
 * launch of each of its two kernels.  Every image's results have the bits of
 * its own call.  No reference counterpart (the reference computes one
 * frame's loss at a time); for batched solves. */
size_t gs_multi_loss_bytes(int n);
int gs_multi_loss_build(int n, const float* const* pred, const float* const* target, const int* C, const int* H,
                        const int* W, const float* window11, float lambda_dssim, int clamp01, float* const* out3,
                        const float* grad_out, void* const* ws, const size_t* ws_bytes, float* const* dpred,
                        void* host_set, size_t set_bytes, unsigned* blocks);
int gs_multi_loss(int n, const void* dev_set, unsigned blocks, void* stream);

/* SE(3) pose parameters as lietorch's LieGroupParameter holds them
 * (scene/gaussian_model_cf.py:328-340 init_RT, :123-133 P[0].retr().act(xyz)):
 * the optimised tensor is a tangent xi = [tau(3), phi(3)] (zero at start) and
 * the pose is X = Exp(xi) * X0, with X = [t(3), q(4) x,y,z,w] and
 * Exp(tau, phi) = (V(phi) tau, quat(phi)),
 * V = I + (1 - cos th)/th^2 [phi]x + (th - sin th)/th^3 [phi]x^2, th = |phi|;
 * (t1,q1) * (t2,q2) = (t1 + R(q1) t2, q1 q2).  Arrays are [n][6] / [n][7].
 *
 * gs_se3_retract: pose_out = Exp(xi) * X0.
 * gs_se3_grad:    grad_xi = (d pose / d xi)^T grad_pose, the exact derivative
 *                 of the map above (forward-mode, double precision), i.e. the
 *                 gradient lietorch's Exp/retr backward delivers to P[i].
 * gs_se3_adam:    the per-frame optimiser of training_setup_fix_position
 *                 (torch.optim.Adam(lr=rotation_lr, eps=1e-15) on P[0],
 *                 scene/gaussian_model_cf.py:303-322): forms grad_xi as above,
 *                 one Adam step on xi (step count read from *step as t-1 and
 *                 incremented on the device, so a captured graph can be
 *                 replayed), then pose_out = Exp(xi) * X0. */
int gs_se3_retract(int n, const float* xi, const float* X0, float* pose_out, void* stream);
int gs_se3_grad(int n, const float* xi, const float* X0, const float* grad_pose, float* grad_xi, void* stream);
int gs_se3_adam(int n, float* xi, const float* X0, const float* grad_pose, float* exp_avg, float* exp_avg_sq,
                int* step, float lr, float beta1, float beta2, float eps, float* pose_out, void* stream);

/* Adam over up to 16 float32 tensors in one launch (replaces
 * torch.optim.Adam(..., eps=1e-15).step() of training_setup,
 * scene/gaussian_model_cf.py:256-268; the step loop calls it at
 * trainer/cf3dgs_trainer.py:163-165).  Arrays are HOST arrays of device
 * pointers (16-byte aligned), element counts and per-tensor learning rates;
 * `step` is the 1-based step count used for the bias corrections. */
int gs_adam_step(int ntensors, float* const* params, const float* const* grads, float* const* exp_avg,
                 float* const* exp_avg_sq, const long long* numel, const float* lr, float beta1, float beta2,
                 float eps, int step, void* stream);

/* gs_adam_step with a second learning rate per tensor: element i of tensor k
 * steps with lead_lr[k] when (i mod row[k]) < lead[k], with lr[k] otherwise
 * (row[k] = 0: lr[k] throughout; row[k] % 4 == 0; numel[k] < 2^31 when
 * row[k] > 0).  One [P,16,3] SH tensor then carries both of the reference's
 * SH groups (scene/gaussian_model_cf.py:257-260: f_dc at feature_lr, the
 * leading 3 floats of every 48, and f_rest at feature_lr / 20). */
int gs_adam_step_rows(int ntensors, float* const* params, const float* const* grads, float* const* exp_avg,
                      float* const* exp_avg_sq, const long long* numel, const float* lr, const float* lead_lr,
                      const int* lead, const int* row, float beta1, float beta2, float eps, int step, void* stream);

/* The Gaussian model's parameter activations (scene/gaussian_model_cf.py:48-56,
 * :116-121, :168-169), one launch each way:
 *   scales = exp(raw_scale) [P,3]; rotations = F.normalize(raw_rot) [P,4]
 *   (x / max(||x||, 1e-12)); opacities = sigmoid(raw_opacity) [P,1].
 * The backward is the chain rule of those torch ops (exp: g * scale; sigmoid:
 * g * (1 - o) * o; normalize through norm, clamp_min and division).  An output
 * pointer may be NULL (not computed); a NULL incoming gradient counts as zero.
 * Rotation arrays are 16-byte aligned. */
int gs_activate_forward(int P, const float* raw_scale, const float* raw_rot, const float* raw_opacity,
                        float* scales, float* rotations, float* opacities, void* stream);
int gs_activate_backward(int P, const float* raw_rot, const float* scales, const float* opacities,
                         const float* dL_dscales, const float* dL_drotations, const float* dL_dopacities,
                         float* dL_draw_scale, float* dL_draw_rot, float* dL_draw_opacity, void* stream);

/* The activation backward followed by one Adam step of the raw parameters
 * (no reference counterpart; the caller side of gs_backward_adam): the raw
 * gradients gs_activate_backward would write go straight into Adam, and
 * raw_scale / raw_rot / raw_opacity are updated in place.  All three incoming
 * gradients are required.  Same update, bit for bit, as gs_activate_backward
 * followed by gs_adam_step_rows (bias corrections 1 - beta^step).  raw_rot,
 * dL_drotations and the rotation moments 16-byte aligned. */
typedef struct gs_adam_act {
  float* scale_m;
  float* scale_v;
  float scale_lr;
  float* rot_m;
  float* rot_v;
  float rot_lr;
  float* opacity_m;
  float* opacity_v;
  float opacity_lr;
  float beta1, beta2, eps;
  int step;
} gs_adam_act_t;
int gs_activate_backward_adam(int P, float* raw_scale, float* raw_rot, float* raw_opacity, const float* scales,
                              const float* opacities, const float* dL_dscales, const float* dL_drotations,
                              const float* dL_dopacities, const gs_adam_act_t* adam, void* stream);

/* Opt-in per-stage timing with HIP events recorded on the launch stream (no
 * reference counterpart; the reference's debug mode only synchronises).
 * Stages: preprocess, scan, duplicate, sort, ranges, blend, render_bwd,
 * preprocess_bwd, knn.  gs_timing_read waits for the recorded events. */
/* Adaptive density control (replaces the torch-op bookkeeping of
 * CFGaussianModel: add_densification_stats scene/gaussian_model_cf.py:716-719
 * with the trainer's max_radii2D update trainer/cf3dgs_trainer.py:144-152;
 * densify_and_clone / densify_and_split / prune_points / densify_and_prune /
 * prune / densify :523-714, including _prune_optimizer and
 * cat_tensors_to_optimizer's moment handling).
 *
 * gs_densify_stats: for every i with radii[i] > 0: grad_accum[i] +=
 *   |grad_means2D[i].xy|, denom[i] += 1, max_radii2D[i] = max(., radii[i]).
 * gs_densify_mark: per Gaussian (raw log-scales [P,3], raw opacities [P]):
 *   g = grad_accum / denom (NaN -> 0); with `densify`: clone if
 *   |g| >= grad_threshold and max exp(scale) <= dense_extent, split if
 *   g >= grad_threshold and max exp(scale) > dense_extent; with `prune`:
 *   sigmoid(opacity) < min_opacity, or (max_screen > 0) max_radii2D >
 *   max_screen, or (big_extent > 0) max exp(scale) > big_extent -- tested for
 *   the original, its clone, and its split copies (scale log(exp(s) * 0.625),
 *   statistics zero).  Writes flags[P] (4-byte aligned) and, on the device,
 *   totals[4] = (originals kept, clones kept, split copies kept per copy,
 *   split-selected); ws: gs_densify_workspace_bytes(P).
 * gs_densify_emit: src/dst[15] = xyz [P,3], features [P,M,3], opacity [P,1],
 *   scaling [P,3], rotation [P,4], then their exp_avg, then their exp_avg_sq
 *   (the ten moment pointers all NULL when there are none); dst arrays hold
 *   totals[0] + totals[1] + 2 totals[2] rows, in the order
 *   [originals kept][clones kept][split copy 1 kept][split copy 2 kept].
 *   split_samples [2 totals[3], 3]: the N(0, exp(scale)) draws of the split
 *   copies (rows of copy 1, then of copy 2, selected Gaussians in index
 *   order); new points get zero moments. */
int gs_densify_stats(int P, const float* grad_means2D, const int* radii, float* grad_accum, float* denom,
                     float* max_radii2D, void* stream);
size_t gs_densify_workspace_bytes(int P);
int gs_densify_mark(int P, const float* grad_accum, const float* denom, const float* scaling, const float* opacity,
                    const float* max_radii2D, float grad_threshold, float dense_extent, float min_opacity,
                    float big_extent, float max_screen, int densify, int prune, uint8_t* flags, int* totals,
                    void* ws, size_t ws_bytes, void* stream);
int gs_densify_emit(int P, int M, const float* const* src, float* const* dst, const float* split_samples,
                    const uint8_t* flags, const int* totals, const void* ws, size_t ws_bytes, void* stream);

int gs_timing_enable(int on);
int gs_timing_reset(void);
int gs_timing_read(const char* stage, double* total_ms, int* launches);

const char* gs_error_string(int code);
/* 2: binning layout with the backward work items, one-word list entries and
 * packed tile rects; split forward (gs_render_prepare / gs_render_resume).
 * 3: image layout with the blend's tile order (tile_order).
 * 4: the depth-first binning split around a depth sort and a list sort
 * shared by many frames (gs_render_depth_keys, gs_multi_depth_sort*,
 * gs_render_binning_duplicate, gs_multi_list_sort*, gs_render_*after_*_sort)
 * and the loss of many images in one launch per kernel (gs_multi_loss*). */
int gs_abi_version(void);

#ifdef __cplusplus
}
#endif

#endif /* GS_RASTER_H */
