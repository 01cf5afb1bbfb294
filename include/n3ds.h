/*
 * n3ds.h - C-ABI of libn3ds.so, the gfx950 (MI355X) implementation of the
 * Neural 3D Strokes stroke-field hot path.
 *
 * Conventions (all entry points):
 *   - every pointer is a device pointer (hipMalloc / torch CUDA tensor storage)
 *     unless stated otherwise; tensors are dense, row-major, float32 unless a
 *     dtype argument says otherwise;
 *   - an absent optional tensor is NULL (the reference passes an empty tensor);
 *   - the caller allocates every output; outputs documented as "accumulated"
 *     must be zero-initialised by the caller (the reference's torch.zeros);
 *   - work is enqueued asynchronously on `stream` (a hipStream_t; NULL = the
 *     legacy default stream); no entry point synchronises or allocates except
 *     where noted;
 *   - the return value is 0 on success, 1 for an invalid argument and 2 for a
 *     HIP runtime error; n3ds_last_error() then returns the message (the
 *     reference raises the TORCH_CHECK message as a Python RuntimeError).
 *
 * Reference files are cited relative to the reference repository root.
 */
#ifndef N3DS_H
#define N3DS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Last error message of the calling thread ("" if none). */
const char* n3ds_last_error(void);

/* ABI version of this library (1). */
int n3ds_version(void);

/* ------------------------------------------------------------------------
 * Per-(point, stroke) stroke field.
 * Replaces the pybind `stroke_forward` of
 * source/strokelib/src/strokes.h:14-27 (kernel strokes_forward.cu:11-109,
 * host strokes_forward.cu:170-225), called from strokes.py:146-183.
 *
 * alpha_out [N,S], color_out [N,S,3], sdf_out [N,S] or NULL,
 * texcoord_out [N,S,2] or NULL; x [N,3], radius [N], viewdir [R,3] with
 * N = R * samples (viewdir row = point / (N/R)); shape_params [S,P],
 * color_params [S,Q].  sdf_id = base<<4 | T | R<<1 | SS<<2 | MS<<3,
 * color_id 0..3 (constant_rgb, gradient_rgb, constant_sh2, constant_sh3).
 * ------------------------------------------------------------------------ */
int n3ds_stroke_forward(float* alpha_out, float* color_out, float* sdf_out, float* texcoord_out,
                        const float* x, const float* radius, const float* viewdir,
                        const float* shape_params, const float* color_params, int64_t n_points,
                        int64_t n_viewdirs, int64_t n_strokes, int64_t n_shape_params,
                        int64_t n_color_params, uint32_t sdf_id, uint32_t color_id, float sdf_delta,
                        int use_laplace_transform, int inv_scale_radius, void* stream);

/* Replaces the pybind `stroke_backward` of source/strokelib/src/strokes.h:38-54
 * (kernel strokes_backward.cu:11-157, host :224-287), called from
 * strokes.py:185-224.  grad_shape_params [S,P], grad_color_params [S,Q] and
 * grad_x [N,3] (or NULL) are accumulated.  grad_sdf may be NULL.
 * adjoint_exact = 1: the true derivative of the forward; 0: the reference's
 * hand-written adjoint term for term. */
int n3ds_stroke_backward(float* grad_shape_params, float* grad_color_params, float* grad_x,
                         const float* grad_alpha, const float* grad_color, const float* grad_sdf,
                         const float* x, const float* radius, const float* viewdir, const float* alpha,
                         const float* shape_params, const float* color_params, int64_t n_points,
                         int64_t n_viewdirs, int64_t n_strokes, int64_t n_shape_params,
                         int64_t n_color_params, uint32_t sdf_id, uint32_t color_id, float sdf_delta,
                         int use_laplace_transform, int inv_scale_radius, int adjoint_exact, void* stream);

/* ------------------------------------------------------------------------
 * Ordered "over" compositing of per-stroke alpha / colour / density.
 * Replaces the pybind `compose_forward` of
 * source/strokelib/src/compositing.h:11-20 (kernel compositing.cu:9-59,
 * host :159-204), called from strokes.py:316-346.
 * density_out [N], color_out [N,D]; alphas [N,S], colors [N,S,D] with
 * D in {1,3}, density_params [S].  Stroke S-1 is in front.
 * ------------------------------------------------------------------------ */
int n3ds_compose_forward(float* density_out, float* color_out, const float* alphas, const float* colors,
                         const float* density_params, int64_t n_points, int64_t n_strokes, int64_t color_dim,
                         void* stream);

/* Replaces the pybind `compose_backward` of compositing.h:22-32 (kernel
 * compositing.cu:61-157, host :206-263), called from strokes.py:348-373.
 * grad_alphas [N,S] and grad_colors [N,S,D] are written; grad_density_params
 * [S] is accumulated.  grad_density_out / grad_color_out may be NULL (zeros).
 * n_strokes <= 2^20 (beyond 8192 strokes the per-point transmittance is
 * checkpointed every few chunks and recomputed, and grad_density_params is
 * accumulated with global atomics). */
int n3ds_compose_backward(float* grad_alphas, float* grad_colors, float* grad_density_params,
                          const float* grad_density_out, const float* grad_color_out, const float* alphas,
                          const float* colors, const float* density_params, int64_t n_points,
                          int64_t n_strokes, int64_t color_dim, int adjoint_exact, void* stream);

/* ------------------------------------------------------------------------
 * Multiresolution hash-grid encoder.
 * Replaces the pybind `grid_encode_forward` of
 * source/gridencoder/src/gridencoder.h:12 (kernel_grid gridencoder.cu:87-245,
 * host :372-400, :448-471), called from grid.py:24-63.
 * inputs [B,D] float32 in [0,1]; embeddings [n_rows,C] of `dtype`
 * (0 = float32, 1 = float16); offsets [L+1] int32; outputs [B, L*C] of
 * `dtype` (already in the [B, L*C] order that grid.py obtains by permuting
 * [L,B,C]); S = log2(per_level_scale), H = base_resolution; dy_dx
 * [B, L*D*C] of `dtype` or NULL.  gridtype 0 = hash, 1 = tiled; interp
 * 0 = linear, 1 = smoothstep.  D in {2,3}, C in {1,2,4,8}; float16 tables
 * are interpolated in float32.
 * ------------------------------------------------------------------------ */
int n3ds_grid_encode_forward(const float* inputs, const void* embeddings, const int32_t* offsets,
                             void* outputs, int64_t B, int D, int C, int L, float S, int H, void* dy_dx,
                             int gridtype, int align_corners, int interp, int dtype, void* stream);

/* Replaces the pybind `grid_encode_backward` of gridencoder.h:13
 * (kernel_grid_backward gridencoder.cu:248-340, kernel_input_backward
 * :343-369, host :402-444, :473-503), called from grid.py:65-89.
 * grad [B, L*C] of `dtype`; grad_embeddings_f32 [n_rows,C] float32 is
 * accumulated (zero-initialised by the caller); grad_embeddings_out
 * [n_rows,C] of `dtype` receives the rounded result when dtype = 1 (may be
 * NULL); grad_inputs [B,D] (written) needs dy_dx, both may be NULL.
 * workspace (device, at least n3ds_grid_encode_backward_workspace_bytes):
 * the table gradient is scattered by bucket of consecutive rows and summed in
 * LDS, with no float atomics to global memory on the hot path; NULL: one
 * float atomic per (point, level, corner, feature), as the reference does. */
int n3ds_grid_encode_backward(const void* grad, const float* inputs, const int32_t* offsets,
                              float* grad_embeddings_f32, void* grad_embeddings_out, int64_t n_rows, int64_t B,
                              int D, int C, int L, float S, int H, const void* dy_dx, void* grad_inputs,
                              int gridtype, int align_corners, int interp, int dtype, void* workspace,
                              int64_t workspace_bytes, void* stream);

/* Bytes of device workspace the bucketed backward needs (host-side size
 * query; -1 for invalid sizes).  No counterpart in the reference (its
 * backward scatters with atomics only). */
int64_t n3ds_grid_encode_backward_workspace_bytes(int64_t B, int D, int C, int L, int64_t n_rows);

/* Fused proposal density field: the forward of the reference's PropMLP with
 * disable_rgb (source/models.py:286-523: _warp_coords :19-33, GridEncoder with
 * bound 1, density_layer Linear(F, 64) -> ReLU -> Linear(64, 1), softplus of
 * raw + density_bias), F = num_levels * level_dim <= 32, in one kernel.
 * coords [N,3]; warp 0: none (coords already in [-1, 1]), 1: coord.contract
 * then / 2 (the 'contract' warp), 2: divide by box_size / 2; the table as for
 * n3ds_grid_encode_forward (S = log2(per_level_scale), H = base resolution,
 * dtype 0 float32 / 1 float16); w1 [64, F], b1 [64], w2 [1, 64], b2 [1] float32
 * (nn.Linear's layout).  density_out [N]; coord_out [N,3] (the warped
 * coordinates the reference returns as `coord`, or NULL); features_out [N, F]
 * and grid_inputs_out [N,3] (the encoder's outputs and inputs, kept for the
 * adjoint; both or neither, NULL when no gradient is needed). */
int n3ds_prop_field_forward(const float* coords, int64_t n_points, int warp, float box_size, const void* embeddings,
                            const int32_t* offsets, int num_levels, int level_dim, float S, int H, int gridtype,
                            int align_corners, int interp, int dtype, const float* w1, const float* b1,
                            const float* w2, const float* b2, float density_bias, float* density_out,
                            float* coord_out, float* features_out, float* grid_inputs_out, void* stream);

/* Adjoint of n3ds_prop_field_forward for upstream grad_density [N]: reads the
 * kept features [N, F] (or, when NULL, gathers them again from the table and
 * writes grid_inputs [N,3]), recomputes the hidden units per point (the [N, 64]
 * activations are never stored), writes grad_features [N, F] (dL/d of the
 * encoder outputs) for n3ds_grid_encode_backward, and adds the weight gradients
 * into grad_w1 [64,F], grad_b1 [64], grad_w2 [64], grad_b2 [1]
 * (zero-initialised by the caller). */
int n3ds_prop_field_backward(const float* coords, int64_t n_points, int warp, float box_size, const void* embeddings,
                             const int32_t* offsets, int num_levels, int level_dim, float S, int H, int gridtype,
                             int align_corners, int interp, int dtype, const float* w1, const float* b1,
                             const float* w2, const float* b2, float density_bias, const float* grad_density,
                             const float* features, float* grad_features, float* grid_inputs, float* grad_w1,
                             float* grad_b1, float* grad_w2, float* grad_b2, void* stream);

/* Replaces the pybind `grad_total_variation` of gridencoder.h:15
 * (kernel_grad_tv gridencoder.cu:506-645), called from grid.py:172-194.
 * Adds weight * TV gradient into grad_f32 [n_rows,C] float32. */
int n3ds_grid_grad_tv(const float* inputs, const void* embeddings, float* grad_f32, const int32_t* offsets,
                      float weight, int64_t B, int D, int C, int L, float S, int H, int gridtype,
                      int align_corners, int dtype, void* stream);

/* ------------------------------------------------------------------------
 * Fused stroke-field renderer (one pass per ray: per-sample stroke SDF ->
 * alpha / colour for a mixed stroke palette, ordered "over" compositing and
 * the volume-render ray march).  It replaces, as one call, the chain of
 * source/models.py:660-711 (StrokeField.predict_density: one stroke_forward
 * per stroke type, columns by stroke index, compose_strokes) followed by
 * source/utils/render.py:29-48 (compute_alpha_weights) and :51-116
 * (volumetric_rendering), with its adjoint.
 * ------------------------------------------------------------------------ */

/* Floats of the stroke table for n_strokes strokes (host-side size query). */
int64_t n3ds_render_table_floats(int n_strokes);

/* Builds the per-stroke table (rotation, scale, culling bounds) from
 * stroke_ids [S] int32 (one sdf_id per stroke) and shape_params [S,shape_width]
 * (each row's leading num_shape_params(sdf_id) entries are used).  *status
 * (device int, zero-initialised) becomes non-zero for an invalid id. */
int n3ds_render_prepare(float* table, const int32_t* stroke_ids, const float* shape_params, int n_strokes,
                        int shape_width, int* status, void* stream);

/* Forward.  origins/directions/viewdirs [R,3], radii [R], tdist [R,n+1]
 * (sample k at the midpoint of tdist[k], tdist[k+1], as render.py:5-26),
 * color_params [S,color_width], density_params [S]; n <= 256.
 * Outputs rgb [R,3], acc [R], depth [R]; state_out [5, R*n] is the per-sample
 * composite state the backward reads.  counters [25] uint64 (or NULL) receive
 * per-base-type counts of evaluated and alpha>0 (sample, stroke) pairs.
 * cone != 0: sample radius = radii * t; coord_div divides sample coordinates
 * (models.py:663 warp); bg_color is a host pointer to 3 floats, the background
 * of every ray unless bg_rgbs (device [R,3], per-ray backgrounds as drawn by
 * models.py:213-222, or NULL) is given; opaque_background makes the last
 * interval infinitely wide (render.py:35-40: its alpha is 1); cull_eps is the
 * alpha below which the Laplace transfer's tail is culled (linear transfer:
 * culling is exact).
 * Optional activations kept for the backward (hit_dir == NULL: none, the
 * backward recomputes everything): hit_dir int32 [R, ceil(S/64), 3],
 * hit_buf int32 [R, cap_hits, 4], pair_buf float [R, npf, cap_pairs] with
 * npf = 4 for color_id 1 (gradient_rgb) else 1.  Per ray, the culled hits
 * and the alpha (and colour) of every evaluated (sample, stroke) pair are
 * kept chunk by chunk (64 strokes) until the ray's capacity is used up; the
 * backward recomputes the chunks that did not fit.
 * workspace (device, n3ds_render_forward_workspace_bytes; or NULL): when
 * given, culling runs first as its own pass and the forward reads its hits;
 * for up to 64 chunks of 64 strokes the pass first bins the strokes in a
 * uniform grid (per cell and chunk, the strokes whose culling bound overlaps
 * the cell) and tests each ray only against the strokes of the cells its
 * segment crosses (n3ds_render_tuning parameter 5); when NULL the forward
 * culls as it goes (no extra memory).  Results are identical. */
int n3ds_render_forward(const float* origins, const float* directions, const float* viewdirs, const float* radii,
                        const float* tdist, const float* table, const float* color_params,
                        const float* density_params, int64_t n_rays, int n_samples, int n_strokes,
                        int shape_width, int color_width, int color_id, float sdf_delta,
                        int use_laplace_transform, int inv_scale_radius, int cone, float coord_div,
                        const float* bg_color, const float* bg_rgbs, int opaque_background,
                        float cull_eps, float* rgb_out, float* acc_out, float* depth_out,
                        float* state_out, unsigned long long* counters, int* hit_dir, int* hit_buf,
                        float* pair_buf, int cap_hits, int cap_pairs, void* workspace, int64_t workspace_bytes,
                        void* stream);

/* Eval-time extras of the ray march: replaces the compute_extras branch of
 * volumetric_rendering (source/utils/render.py:88-116, called from
 * models.py:224-233 when render_image asks for extras) for the fused
 * renderer.  From the per-sample composite density in state_out of
 * n3ds_render_forward (row 0 of [5, R*n]), tdist [R, n+1], directions [R,3]
 * and the far distance t_far [R] (batch['far']): extras_out [R, 4] =
 * distance_mean (exp of the weighted mean log depth, clipped to the ray's
 * fenceposts), and the 5th, 50th and 95th weighted percentiles of the
 * fenceposts extended by t_far with the background weight
 * (stepfun.weighted_percentile, stepfun.py:351-361).  opaque_background as
 * passed to the forward. */
int n3ds_render_extras(const float* state, const float* tdist, const float* directions, const float* t_far,
                       int64_t n_rays, int n_samples, int opaque_background, float* extras_out, void* stream);

/* Launch-shape parameters of the render kernels (no counterpart in the
 * reference), set by the autotuner for the current process: 0 = pending
 * (sample, stroke) pairs that trigger an evaluation window (16..256, default
 * 128), 1 = backward persistent blocks per resident block (1..8, default 1),
 * 2 = batch backward blocks per resident block (1..64, default 8; read by
 * n3ds_render_batch_plan), 3 = 64-stroke chunks per wavefront of the culling
 * pass (1..16, default 1; a batch's plan depends on it), 4 = the adjoint's
 * gradient groups queued per SDF family across rays (0..1, default 1), 5 = the
 * culling pass of one scene with the stroke grid (0..1, default 1), 6 = the
 * split adjoint when its workspace is given (0..1, default 1), 7 = records per
 * ray and SDF family the split adjoint's lists hold (1..1024, default 32; a
 * fuller list sends every ray to the one-kernel adjoint).  value < 0 only
 * queries.  Returns the previous
 * value, or -1 for an unknown parameter or a value out of range.  Results do
 * not depend on these parameters beyond float summation order. */
int n3ds_render_tuning(int param, int value);

/* Bytes of the forward's culling-pass workspace: 260 B per (ray, chunk of 64
 * strokes), plus, for at most 64 chunks, the stroke grid (at most 4 MiB of cell
 * masks, 12 B per ray and 8 B per (ray, chunk)); max_strokes = the stroke count
 * (a batch: its largest scene's). */
int64_t n3ds_render_forward_workspace_bytes(int64_t n_rays, int max_strokes);

/* Floats of device workspace n3ds_render_backward needs (queries the current
 * device; -1 on a HIP error). */
int64_t n3ds_render_backward_workspace_floats(int64_t n_rays, int n_strokes, int shape_width, int color_width);

/* Floats of device workspace the split adjoint needs, or 0 when it does not apply
 * (no counterpart in the reference: the reference's adjoint is the per-operator chain).
 * With the forward's kept hits (hit_dir != NULL), n3ds_render_backward runs the
 * adjoint as a per-ray kernel that lists the live (sample, stroke) pairs' adjoint
 * records per SDF family and a kernel over those records, when its workspace holds
 * max(this, n3ds_render_backward_workspace_floats(...)) floats; same results up to
 * the float summation order of the gradients. */
int64_t n3ds_render_backward_split_floats(int64_t n_rays, int n_strokes, int shape_width, int color_width);

/* Backward: accumulates dL/d(shape_params) [S,shape_width], dL/d(color_params)
 * [S,color_width] and dL/d(density_params) [S] for upstream grad_rgb [R,3],
 * grad_acc [R] (or NULL) and grad_depth [R] (or NULL), from the forward's
 * `state` and, when given, the hits / pair values it kept (same buffers and
 * capacities as passed to n3ds_render_forward); the per-pair SDF gradients
 * are always recomputed.  adjoint_exact as for n3ds_stroke_backward. */
int n3ds_render_backward(const float* origins, const float* directions, const float* viewdirs, const float* radii,
                         const float* tdist, const float* table, const float* color_params,
                         const float* density_params, int64_t n_rays, int n_samples, int n_strokes,
                         int shape_width, int color_width, int color_id, float sdf_delta,
                         int use_laplace_transform, int inv_scale_radius, int cone, float coord_div,
                         const float* bg_color, const float* bg_rgbs, int opaque_background,
                        float cull_eps, int adjoint_exact, const float* grad_rgb,
                         const float* grad_acc, const float* grad_depth, const float* state, const int* hit_dir,
                         const int* hit_buf, const float* pair_buf, int cap_hits, int cap_pairs, float* workspace,
                         int64_t workspace_floats, float* grad_shape, float* grad_color, float* grad_density,
                         void* stream);

/* Allocation-free variant of n3ds_render_backward.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments and results as n3ds_render_backward without the workspace:
 * the per-block partial gradient rows live in a fixed device array of the
 * library (n3ds_render_backward_noalloc_scratch_floats() floats = 32 MiB, one
 * copy per GPU, held until the process ends).  It allocates nothing and
 * rejects n_strokes * (shape_width + color_width + 1) beyond that size
 * (N3DS_EINVAL); the persistent block count is lowered to fit the scratch,
 * which changes only the float summation order of the gradients.  Every call
 * on one GPU shares that GPU's array: calls that overlap on one GPU (other
 * streams, other host threads) race and must be serialised by the caller. */
int64_t n3ds_render_backward_noalloc_scratch_floats(void);
int n3ds_render_backward_noalloc(const float* origins, const float* directions, const float* viewdirs,
                                 const float* radii, const float* tdist, const float* table,
                                 const float* color_params, const float* density_params, int64_t n_rays,
                                 int n_samples, int n_strokes, int shape_width, int color_width, int color_id,
                                 float sdf_delta, int use_laplace_transform, int inv_scale_radius, int cone,
                                 float coord_div, const float* bg_color, const float* bg_rgbs, int opaque_background,
                        float cull_eps, int adjoint_exact,
                                 const float* grad_rgb, const float* grad_acc, const float* grad_depth,
                                 const float* state, const int* hit_dir, const int* hit_buf, const float* pair_buf,
                                 int cap_hits, int cap_pairs, float* grad_shape, float* grad_color,
                                 float* grad_density, void* stream);

/* ------------------------------------------------------------------------
 * Batches of independent scenes.  No counterpart in the reference, which
 * trains one scene per process (train.py:127-144): many scenes with uneven
 * stroke and ray counts render in one launch per stage.  Rays and strokes are
 * packed scene after scene in the same per-ray / per-stroke arrays as above
 * (the stroke table is n3ds_render_prepare over all packed strokes).  The
 * scene table is int32: [n_scenes, 10] records {first ray, rays, first stroke,
 * strokes, first backward block, backward blocks, offset of the scene's
 * partial gradient rows, backward wavefronts per block (0: accumulators in
 * global memory), first culling task (the scene's rays x its own 64-stroke
 * chunks, scene after scene), 0} followed by the scene of every backward block.  It is passed
 * twice, in host memory (validated and used to size launches) and as a device
 * copy (read by the kernels).  Results per scene equal n3ds_render_forward /
 * _backward on that scene alone.
 * ------------------------------------------------------------------------ */

/* Fills record fields 4-9 (the backward plan: blocks per scene in proportion
 * to each scene's rays x strokes, oversubscribing the device so the hardware
 * dispatcher balances uneven scenes) from fields 0-3, and the block map when
 * block_map holds block_map_cap >= the block count; returns the number of
 * backward blocks, or the negated status on error. */
int n3ds_render_batch_plan(int32_t* scenes_host, int n_scenes, int n_samples, int shape_width, int color_width,
                           int color_id, int32_t* block_map, int block_map_cap);

/* Forward of every scene of the batch: as n3ds_render_forward with
 * n_strokes = all packed strokes; the kept-hit layout is sized by the
 * largest scene's strokes. */
int n3ds_render_batch_forward(const float* origins, const float* directions, const float* viewdirs,
                              const float* radii, const float* tdist, const float* table, const float* color_params,
                              const float* density_params, const int32_t* scenes_host, const int32_t* scenes,
                              int n_scenes, int64_t n_rays, int n_samples, int n_strokes, int shape_width,
                              int color_width, int color_id, float sdf_delta, int use_laplace_transform,
                              int inv_scale_radius, int cone, float coord_div, const float* bg_color, const float* bg_rgbs, int opaque_background,
                        float cull_eps,
                              float* rgb_out, float* acc_out, float* depth_out, float* state_out,
                              unsigned long long* counters, int* hit_dir, int* hit_buf, float* pair_buf, int cap_hits,
                              int cap_pairs, void* workspace, int64_t workspace_bytes, void* stream);

/* Floats of device workspace n3ds_render_batch_backward needs for a planned
 * scene table (-1 on a HIP error). */
int64_t n3ds_render_batch_backward_workspace_floats(const int32_t* scenes_host, int n_scenes, int n_samples,
                                                    int shape_width, int color_width, int color_id);

/* Backward of every scene of the batch: one launch per class of scenes (scenes
 * whose LDS accumulators allow the same wavefronts per block; usually one or
 * two classes), each block working on one scene, plus one reduction;
 * accumulates into the packed gradient tensors.  The scene table must be
 * planned (n3ds_render_batch_plan, block map included in both copies). */
int n3ds_render_batch_backward(const float* origins, const float* directions, const float* viewdirs,
                               const float* radii, const float* tdist, const float* table, const float* color_params,
                               const float* density_params, const int32_t* scenes_host, const int32_t* scenes,
                               int n_scenes, int64_t n_rays, int n_samples, int n_strokes, int shape_width,
                               int color_width, int color_id, float sdf_delta, int use_laplace_transform,
                               int inv_scale_radius, int cone, float coord_div, const float* bg_color, const float* bg_rgbs, int opaque_background,
                        float cull_eps,
                               int adjoint_exact, const float* grad_rgb, const float* grad_acc,
                               const float* grad_depth, const float* state, const int* hit_dir, const int* hit_buf,
                               const float* pair_buf, int cap_hits, int cap_pairs, float* workspace,
                               int64_t workspace_floats, float* grad_shape, float* grad_color, float* grad_density,
                               void* stream);

/* ------------------------------------------------------------------------
 * Stroke-parameter update (one training step's optimizer stage).
 * Together these replace, for fp32 CUDA parameters, the reference's
 * clip_gradients (source/utils/training.py:67-77), torch.optim.Adam as built
 * by create_optimizer (training.py:119-136; betas / eps from
 * source/configs.py:82-84) and StrokeField.clip_params
 * (source/models.py:576-584).
 * ------------------------------------------------------------------------ */

/* *acc (device double, zero before the first tensor of a step) += sum of
 * grads[i]^2 over n floats (the global norm of clip_grad_norm_ is accumulated
 * tensor by tensor). */
int n3ds_grad_sqnorm_accumulate(const float* grads, int64_t n, double* acc, void* stream);

/* *coef = min(max_norm / (sqrt(*acc) + 1e-6), 1), NaN propagating
 * (torch.nn.utils.clip_grad_norm_), then *acc = 0 for the next step. */
int n3ds_grad_clip_coef(double* acc, float max_norm, float* coef, void* stream);

/* In place, for one parameter tensor [n_rows, n_cols]:
 *   g *= *grad_coef (if grad_coef != NULL); g = clamp(g, -v, v) if
 *   grad_max_val = v > 0; g = nan_to_num(g) (written back to grads);
 *   Adam (torch.optim.Adam semantics, step = t after increment, t >= 1).
 *   When step_dev / lr_dev (device floats) are given, t / lr are read from
 *   them, so a captured step replays with the current count and rate
 *   (the per-step rate of learning_rate_decay, training.py:89-116);
 *   p = clamp(p, col_lo[c], col_hi[c]) per column when col_lo / col_hi
 *   (device [n_cols], +-inf for an open side) are given, then
 *   p = clamp(p, lo_all, hi_all).  n_cols <= 1024. */
int n3ds_adam_step(float* params, float* grads, float* exp_avg, float* exp_avg_sq, int64_t n_rows, int n_cols,
                   const float* col_lo, const float* col_hi, double lr, const float* lr_dev, double beta1,
                   double beta2, double eps, int64_t step, const float* step_dev, float grad_max_val,
                   const float* grad_coef, float lo_all, float hi_all, void* stream);

/* ------------------------------------------------------------------------
 * Image texture on stroke colours.
 *
 * Replaces the torch ops of ImageTexture.forward / sample_texture
 * (source/textures.py:51-82; applied between the stroke and compositing
 * operators at models.py:704-706):
 *   tint = grid_sample(image, texcoords * 2 - 1, bilinear, zeros padding,
 *                      align_corners=True);
 *   colors_out = (colors * tint_ratio + (1 - tint_ratio)) * tint;
 *   alphas_out = alphas * mean(tint)   (only when modulate_alpha).
 * quads: the image packed by n3ds_texture_pack (16-byte aligned),
 * texcoords [N, 2] (u along W, v along H), colors / colors_out [N, 3],
 * alphas / alphas_out [N] (read / written only when modulate_alpha).
 * ------------------------------------------------------------------------ */
int n3ds_texture_forward(float* colors_out, float* alphas_out, const float* texcoords, const float* colors,
                         const float* alphas, const float* quads, int64_t height, int64_t width, int64_t channels,
                         int64_t n, double tint_ratio, int modulate_alpha, void* stream);

/* Floats of the packed form of an H x W image with `channels` in {1, 3}
 * ((H+1)(W+1) entries of 16 or 4 floats; H*W <= 2^26), or -1. */
int64_t n3ds_texture_pack_floats(int64_t height, int64_t width, int64_t channels);

/* Packs image [C, H, W] float (the reference's image_texture[0],
 * textures.py:45-49) into quads: entry (j, i) holds the corners
 * (x, y) = (i-1, j-1), (i, j-1), (i-1, j), (i, j) channel by channel, 0 off
 * the image, so one bilinear tap reads one aligned entry. */
int n3ds_texture_pack(float* quads, const float* image, int64_t height, int64_t width, int64_t channels,
                      void* stream);

/* ------------------------------------------------------------------------
 * Proposal-level interval resampling.
 *
 * Replaces the torch ops of source/utils/stepfun.py:sample_intervals
 * (:280-317, with sample :216-254, invert_cdf :195-202, integrate_weights
 * :149-170, sorted_interp :5-24), called at models.py:175-181.
 * t [R, n_bins+1] sorted bin edges, w_logits [R, n_bins]; u_base [n_samples]
 * the sample positions (the reference's torch.linspace); jitter [R,
 * jitter_width] (torch.rand, jitter_width 1 = single_jitter, or n_samples)
 * or NULL, scaled by max_jitter; t_out [R, n_samples+1].  Centres are the
 * inverse CDF of softmax(w_logits) at u; fenceposts are their midpoints and
 * the end reflections clamped to [minval, maxval].  n_samples > 1,
 * 2 (n_bins + 1) + n_samples <= 16384.
 * ------------------------------------------------------------------------ */
int n3ds_sample_intervals(float* t_out, const float* t, const float* w_logits, const float* u_base,
                          const float* jitter, int jitter_width, float max_jitter, int64_t n_rays, int n_bins,
                          int n_samples, float minval, float maxval, void* stream);

/* Replaces the torch ops of stepfun.py:max_dilate_weights (:136-146, over
 * max_dilate :120-133, weight_to_pdf :109-112, pdf_to_weight :115-117),
 * called at models.py:159-166.  t [R, n_bins+1] sorted, w [R, n_bins];
 * t_out [R, 3 n_bins + 1] (t, t[:-1] - dilation, t[1:] + dilation in sorted
 * order, clipped to [minval, maxval]); w_out [R, 3 n_bins] the max-dilated
 * PDF back as weights, divided by max(sum, eps) when renormalize.
 * n_bins <= 3276. */
int n3ds_max_dilate_weights(float* t_out, float* w_out, const float* t, const float* w, int64_t n_rays, int n_bins,
                            float dilation, float minval, float maxval, int renormalize, void* stream);

/* Adjoint of n3ds_texture_forward: what torch's autograd gives the ops of
 * textures.py:51-82 (the image is a non-persistent buffer without gradient).
 * grad_colors [N,3] and, when modulate_alpha, grad_alphas [N] are written;
 * grad_texcoords [N,2] is written when not NULL (the reference's stroke
 * backward discards it, strokes.py:188-224).  grad_colors_out /
 * grad_alphas_out may be NULL (zeros). */
int n3ds_texture_backward(float* grad_colors, float* grad_alphas, float* grad_texcoords,
                          const float* grad_colors_out, const float* grad_alphas_out, const float* texcoords,
                          const float* colors, const float* alphas, const float* quads, int64_t height,
                          int64_t width, int64_t channels, int64_t n, double tint_ratio, int modulate_alpha,
                          void* stream);

#ifdef __cplusplus
}
#endif

#endif /* N3DS_H */
