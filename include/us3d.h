/*
 * us3d.h -- C ABI of the MI355X (gfx950) differentiable Gaussian slice renderer.
 *
 * Two renderers share this library:
 *   - UltraGauss slice splat ("ug"): plane/Gaussian intersection preprocess, (tile, depth) keys with a
 *     per-tile LDS sort and tile ranges, per-tile splatting of echo intensity, and its adjoint.
 *   - UltraG-Ray B-mode ray renderer ("ur"): per-Gaussian scanline range cull, (ray, depth) keys with a
 *     per-ray LDS sort, transmittance x echo compositing along each scanline, and its adjoint.
 *
 * Conventions
 *   - Every pointer inside a plan, and every array argument, is a DEVICE pointer unless stated.
 *   - Nothing is allocated inside the library. All scratch lives in a caller-provided workspace whose
 *     size comes from the matching *_workspace_bytes() query (the one exception, us3d_ug_render_static, uses a
 *     fixed static per-GPU array instead). Every entry point is stream-ordered on
 *     `stream` (a hipStream_t passed as void*; NULL = legacy default stream), performs no host
 *     synchronisation and issues no hipMemset, so a call sequence can be captured into a hipGraph.
 *   - Results that depend on data (instance and visible-pair counts, overflow) are written to the us3d_stats
 *     block at offset 0 of the workspace. If the instance count exceeds the plan's k_capacity, or the visible
 *     (frame, Gaussian) pairs exceed its v_capacity, the kernels that would write past them return early,
 *     stats.overflow is set and the caller must retry with larger capacities (the reference instead blocks on
 *     a device->host copy to learn the count).
 *   - Return value: US3D_OK or an error code; us3d_last_error() describes the last failure on the
 *     calling thread.
 *   - Floating point: FP32 storage and arithmetic. Tile rectangles, keys, and the q <= chi2 test use
 *     IEEE division/sqrt and explicit FMAs only (library built with -ffp-contract=off), so they are
 *     reproducible bit-for-bit by a CPU restatement that performs the same operations.
 */
#ifndef US3D_H
#define US3D_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define US3D_OK 0
#define US3D_ERR_ARGUMENT 1
#define US3D_ERR_HIP 2
#define US3D_ERR_CAPACITY 3

/* One probe frame. 100 bytes, all fields 4 bytes, no padding. Axis vectors are used as given
 * (the training loop normalises them).
 * geometry 0, linear probe: pixel (x,y) at origin + ((x - col_offset) delta_lat) a_lat + (near_mm + y delta_ax) a_ax.
 * geometry 1, fan (sector) probe, UltraGauss only: origin is the apex, delta_lat the angular step (radians per
 * column), col_offset the column at angle 0, near_mm the radius of row 0 and delta_ax the radial step; pixel
 * (x,y) at origin + r sin(t) a_lat + r cos(t) a_ax with t = (x - col_offset) delta_lat, r = near_mm + y delta_ax.
 * Explicit pixel positions of a fan frame must lie on that grid to within half a pixel. */
typedef struct us3d_frame {
    float origin[3];        /* probe origin, world mm */
    float axis_lat[3];      /* lateral unit axis */
    float axis_ax[3];       /* axial (depth) unit axis */
    float axis_ele[3];      /* elevational unit axis */
    float delta_lat;        /* mm per pixel, lateral (UltraGauss); width_mm / W (UltraG-Ray) */
    float delta_ax;         /* mm per pixel, axial */
    float near_mm;          /* depth of row 0 */
    float col_offset;       /* lateral_col_offset: column of lateral coordinate 0 */
    int32_t height;         /* H: rows (UltraGauss) or depth samples per scanline (UltraG-Ray) */
    int32_t width;          /* W: columns / scanlines */
    int32_t tiles_x;        /* tiles per row (UltraGauss); 1 (UltraG-Ray) */
    int32_t tiles_y;        /* tile rows (UltraGauss); W scanlines (UltraG-Ray) */
    int32_t tile_base;      /* first global tile (UltraG-Ray: scanline) index of this frame */
    int32_t pixel_base;     /* first pixel of this frame in the concatenated pixel arrays */
    int32_t sweep;          /* index into the sweep table */
    int32_t ray_base;       /* first ray of this frame in concatenated per-ray arrays (UltraG-Ray) */
    int32_t geometry;       /* 0: linear probe grid; 1: fan (see above) */
} us3d_frame;

/* A sweep: a set of frames rendered from one Gaussian set. Gaussians [g_begin, g_begin+g_count) of the
 * concatenated parameter arrays; frames [f_begin, f_begin+f_count). (frame, Gaussian) pairs of a sweep are
 * numbered Gaussian-major from pair_base: pair = pair_base + (g - g_begin) * f_count + (f - f_begin).
 * The sweep table is ordered by g_begin and the Gaussian ranges are disjoint; they need not cover all n_gauss
 * Gaussians: one outside every range is rendered in no frame of the batch and its gradients are zero. */
typedef struct us3d_sweep {
    int32_t g_begin;
    int32_t g_count;
    int32_t f_begin;
    int32_t f_count;
    int64_t pair_base;
} us3d_sweep;

/* Data-dependent results, at offset 0 of every workspace. Every map call resets the block except
 * overflow_any, which only the caller clears (write 0 to it after allocating the workspace, e.g. with
 * us3d_fill_i32): a run of replayed steps can then be checked for any overflowed step with one read. */
typedef struct us3d_stats {
    int64_t n_instances;    /* K: (tile, Gaussian) or (ray, Gaussian) instances of the batch (0 when the visible
                             * pairs overflowed: the instances are then not counted) */
    int64_t max_segment;    /* longest per-tile / per-ray list */
    int32_t overflow;       /* 1 if n_visible > v_capacity or n_instances > k_capacity (outputs are then invalid) */
    int32_t n_sorted_big;   /* segments that took the multi-chunk merge path */
    int32_t overflow_any;   /* sticky: set to 1 by every map call that overflows, never cleared by the library */
    int32_t reserved_i32;
    int64_t n_visible;      /* visible (frame, Gaussian) pairs: those with at least one instance */
    int64_t reserved[3];    /* mapper scratch (reserved[0]: tile total, reserved[1]: frame-major total); zeroed
                             * by every map call */
} us3d_stats;

/* A batch: frames (possibly from several sweeps) that are mapped, rendered and differentiated together.
 * Host struct holding device pointers. n_tiles = sum of frames' tiles_x*tiles_y (UltraGauss) or of their
 * widths (UltraG-Ray); n_pairs = sum over sweeps of g_count*f_count. n_pairs and n_pixels must be < 2^31.
 * Per-pair records are kept only for visible pairs (at least one instance), at most v_capacity of them; the
 * workspace holds 0.625 bytes per (frame, Gaussian) pair besides (two bit masks and the scans of their words). */
typedef struct us3d_plan {
    int32_t n_gauss;                /* rows of the concatenated Gaussian arrays */
    int32_t n_frames;
    int32_t n_sweeps;
    int32_t n_tiles;
    int64_t n_pairs;
    int64_t n_pixels;               /* sum of H*W */
    int64_t k_capacity;             /* capacity of instance arrays in the workspace (< 2^31) */
    int64_t v_capacity;             /* capacity of the visible-pair arrays in the workspace (< 2^31) */
    const us3d_frame* frames;       /* [n_frames] */
    const int32_t* frame_tile_base; /* [n_frames+1] prefix of per-frame tile counts */
    const int32_t* frame_pixel_base;/* [n_frames+1] prefix of per-frame pixel counts */
    const us3d_sweep* sweeps;       /* [n_sweeps] */
    void* workspace;
    int64_t workspace_bytes;
} us3d_plan;

/* ---------------------------------------------------------------- UltraGauss slice splat */

typedef struct us3d_ug_settings {
    float chi2_threshold;   /* q = d^T Lambda d cut (7.815); 120 reproduces the fallback's exp clamp at -60 */
    float diag_beta;        /* l_ii = raw_ii^2 + beta */
    float bg_alpha;         /* fixed background alpha */
    float bg_intensity;     /* fixed background intensity */
    float alpha_min;        /* pairs with alpha_hat <= alpha_min are skipped (1e-8); negative disables */
    int32_t sort_by_depth;  /* 1: key = (tile, depth key); 0: key = (tile) */
    int32_t tile_size;      /* 8, 16 or 32 pixels */
    int32_t alpha_excludes_bg; /* alpha output = sum alpha_hat (1) or sum alpha_hat + bg_alpha (0) */
    int32_t raw_activations;   /* 1: the opacity and intensity arrays hold logits. us3d_ug_map applies the
                                * sigmoid 1/(1+exp(-x)) of the training model's activations
                                * (pipeline_god_file.py:369-373) and us3d_ug_backward returns the gradients
                                * w.r.t. the logits (chain s(1-s)); 0: they hold opacity and intensity. */
    int32_t fwd_lds_batch;  /* Gaussians the forward splat stages in LDS per batch: 64, 128 or 256, or 0 for its
                             * thread count (64/128/256 at tiles 8/16/32). No counterpart in the reference (its
                             * splat reads the Gaussians one by one); the outputs do not depend on it. */
    int32_t bwd_lds_batch;  /* the same for the adjoint (0: 64/256/256 at tiles 8/16/32); the gradients depend on
                             * it only through the FP32 rounding of the per-Gaussian sums. */
    int32_t work_queue;     /* 0: the forward and the adjoint launch one workgroup per (frame, tile) in an XCD-aware
                             * order; 1: persistent workgroups (as many as the device holds) take the tiles from
                             * per-XCD queues in the same order. No counterpart in the reference (one launch per
                             * frame); the outputs do not depend on it. */
} us3d_ug_settings;

/* Byte offsets of the arrays inside a workspace (for inspection/debug outputs). Shared by the renderers.
 * (frame, Gaussian) pairs are numbered densely (us3d_sweep); visible pairs (at least one instance) get a compact
 * index v, their rank in that order, so a Gaussian's visible pairs are consecutive in frame order.
 * Arrays: queue (the heads of the per-XCD tile queues of the work_queue launches), pair_bits u64[ceil(n_pairs/64)]
 * (bit p: pair p is visible), word_cnt i32[words] and word_off
 * i32[words+1] (popcount of each word and its exclusive scan: v = word_off[p/64] + popcount of the bits below p),
 * fm_bits u32[fm_words] (frame-major visibility: per frame ceil(g_count/32) words, bit gl for Gaussian
 * g_begin+gl), fm_cnt / fm_off (their popcounts and scan: the slot of a visible pair in vis_list), vis_pair
 * i32[v_cap] (dense pair index of v), pair_count i32[v_cap] (instances of v), pair_off i32[v_cap+1] (first emit
 * index), pair_rect i32x4[v_cap] (tile rectangle [x0,x1)x[y0,y1); UltraG-Ray: (0, min_ray, 1, max_ray+1)),
 * pair_dk u32[v_cap] (depth key), pair_g i32[v_cap] (Gaussian of v), vis_list i32[v_cap] (frame-major list of
 * v: a frame's visible pairs are consecutive, in Gaussian order), tile_count i32[n_tiles], tile_off
 * i32[n_tiles+1], tile_cursor i32[n_tiles] (emit cursors), seg_key u64[K] (the instances bucket by bucket, each
 * as the key depth key << 32 | v of its visible pair v, sorted: inside a tile v ascends with the Gaussian, so this
 * is the reference's (depth key, Gaussian) order; an instance's emit index is pair_off[v] plus its row-major
 * position in the pair's rectangle), seg_key_tmp u64[K] (sort scratch), records (per Gaussian), partial (per
 * instance, at the emit index), inst_mask u16[K] (UltraGauss: the sub-block mask of each sorted instance, written
 * by the forward, read by the backward; empty for UltraG-Ray), pair_sect (per visible pair: UltraGauss f32[16],
 * its part of the sub-block test, the plane form of q on the frame's grid and its emit base; GaussianPile
 * f32[12], its focus footprint and emit base; written by the map, read by the forward and the backward). */
typedef struct us3d_layout {
    int64_t stats, queue, pair_bits, word_cnt, word_off, fm_bits, fm_cnt, fm_off, vis_pair;
    int64_t pair_count, pair_off, pair_rect, pair_dk, pair_g, vis_list, tile_count, tile_off;
    int64_t tile_cursor, seg_key, seg_key_tmp, records, partial, scan_tmp;
    int64_t inst_mask, pair_sect, total;
} us3d_layout;

int64_t us3d_ug_workspace_bytes(int64_t n_gauss, int64_t n_pairs, int64_t n_tiles, int64_t k_capacity,
                                int64_t v_capacity);
int us3d_ug_workspace_layout(int64_t n_gauss, int64_t n_pairs, int64_t n_tiles, int64_t k_capacity,
                             int64_t v_capacity, us3d_layout* out);

/* Mapper: replaces MapperForwardCUDA (ops/ultragauss_renderer/src/mapper/wrapper_mapper.cu:35-206,
 * declared at ops/ultragauss_renderer/include/mapper/wrapper_mapper.cuh:4-22). Packs the Gaussians,
 * computes per-(frame, Gaussian) tile rectangles and depth keys (core.cu:9-116), scans, emits
 * (tile, depth, Gaussian) instances (core.cu:118-143), sorts each tile's list in LDS by
 * (depth key, Gaussian) -- the order CUB's stable radix sort of (tile<<32|depth) produces --
 * and leaves per-tile ranges as tile_off (core.cu:145-164). means [G,3], tril [G,6], opacity [G],
 * intensity [G]. */
int us3d_ug_map(const us3d_plan* plan, const float* means, const float* tril, const float* opacity,
                const float* intensity, const us3d_ug_settings* settings, void* stream);

/* Forward splat: replaces ShaderForwardCUDA (ops/ultragauss_renderer/src/shader/wrapper_shader.cu:17-66,
 * forward.cu:17-76). pixels [n_pixels,3] world mm, concatenated per frame (frame.pixel_base); NULL means the
 * linear-probe grid origin + (x - col_offset)*delta_lat*a_lat + (near + y*delta_ax)*a_ax computed in-kernel
 * with the FP32 operations of pipeline_god_file.py:299-306 (bit-identical to that tensor).
 * Outputs [n_pixels]: image, alpha, depth; aux [n_pixels,4] = (sum alpha_hat, unclamped image,
 * sum alpha_hat*elev, 0) kept for the adjoint (sums accumulated in FP64, stored as FP32). Requires a prior us3d_ug_map on the same workspace;
 * leaves each instance's sub-block mask in the workspace (layout inst_mask) for the adjoint.
 * depth may be NULL: the depth numerator is then not accumulated, aux holds 0 in its place, and a backward on
 * this aux must pass d_depth = NULL. The backward with d_alpha = d_depth = NULL (image gradients only) skips
 * their terms as well. */
int us3d_ug_forward(const us3d_plan* plan, const float* pixels, const us3d_ug_settings* settings,
                    float* image, float* alpha, float* depth, float* aux, void* stream);

/* Adjoint: replaces ShaderBackwardCUDA (ops/ultragauss_renderer/src/shader/wrapper_shader.cu:68-122,
 * backward.cu:17-131). Requires the map and forward of the same call sequence on the same workspace, with
 * nothing else run on it in between (it reads their sorted lists and sub-block masks).
 * Recomputes per-pair weights from the stored per-pixel sums, accumulates per
 * (tile, Gaussian) in registers (no atomics) and reduces per Gaussian in a fixed order, so the result
 * is bitwise reproducible. d_alpha / d_depth may be NULL (zero). Outputs (overwritten): d_means [G,3],
 * d_tril [G,6] (w.r.t. the raw tril parameters), d_opacity [G], d_intensity [G]. */
int us3d_ug_backward(const us3d_plan* plan, const float* pixels, const float* tril,
                     const us3d_ug_settings* settings, const float* aux, const float* d_image,
                     const float* d_alpha, const float* d_depth, float* d_means, float* d_tril,
                     float* d_opacity, float* d_intensity, void* stream);

/* Training step with the image loss fused into the splat: the pair us3d_ug_forward + us3d_image_loss +
 * us3d_ug_backward(d_image = the loss gradient, d_alpha = d_depth = NULL) of the training loop
 * (pipeline_god_file.py:739-764 around image_loss, :604-637), without the two per-pixel loss passes.
 * us3d_ug_forward_loss renders as us3d_ug_forward and, in the same kernel, sums the frame statistics of the loss
 * against gt (indexed like the image) per (tile, wavefront); a small kernel then writes loss[n_frames] and the
 * frames' gradient coefficients into tmp (us3d_ug_loss_tmp_bytes(plan) bytes; scale, w_l1 and w_ssim as in
 * us3d_image_loss). us3d_ug_backward_loss is us3d_ug_backward with d_image formed from the clamped image, gt and
 * those coefficients while the adjoint stages its pixels; it needs the same plan, gt and tmp. The sums run in a
 * different (fixed) order than us3d_image_loss's (a thread's at most four pixels in FP32, then FP64), so loss and
 * gradients agree with the unfused sequence to rounding (1e-7 relative), and each is bitwise reproducible. No counterpart in the reference (its loss is torch ops). */
int64_t us3d_ug_loss_tmp_bytes(const us3d_plan* plan);
int us3d_ug_forward_loss(const us3d_plan* plan, const float* pixels, const us3d_ug_settings* settings,
                         float* image, float* alpha, float* depth, float* aux, const float* gt, float w_l1,
                         float w_ssim, float scale, float* loss, double* tmp, void* stream);
int us3d_ug_backward_loss(const us3d_plan* plan, const float* pixels, const float* tril,
                          const us3d_ug_settings* settings, const float* aux, const float* gt, const double* tmp,
                          float* d_means, float* d_tril, float* d_opacity, float* d_intensity, void* stream);

/* Per-pixel candidate lists (the reference's pixel_gauss_counts/offsets/indices extras,
 * core.cu:166-287): candidates of each pixel in sorted tile order filtered by the lateral/axial window.
 * counts [n_pixels] is written by pass 0; offsets [n_pixels] must be the exclusive scan of counts
 * (us3d_scan_i32) before pass 1 fills indices (Gaussian indices relative to the sweep's g_begin). */
int us3d_ug_pixel_lists(const us3d_plan* plan, const float* pixels, const us3d_ug_settings* settings,
                        int pass, int32_t* counts, const int32_t* offsets, int32_t* indices,
                        void* stream);

/* No-allocation render. No counterpart in the reference: for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). Runs us3d_ug_map then
 * us3d_ug_forward on the same arguments, with the workspace taken from a static per-GPU array of
 * US3D_STATIC_SCRATCH_BYTES instead of the plan's: plan->workspace, workspace_bytes, k_capacity and v_capacity
 * are ignored (may be 0); the instance capacity is the largest that fits with a visible-pair capacity of
 * min(n_pairs, instance capacity), which cannot overflow before the instances do. A batch whose workspace cannot fit is rejected with
 * US3D_ERR_CAPACITY before any device work. overflow_dev (device int, may be NULL) receives 1 if the instances
 * exceeded that capacity (the outputs are then invalid), else 0. Allocates nothing (no hipMalloc/hipFree); the first
 * call on a device looks the array up with hipGetSymbolAddress. Each GPU's array keeps the last call's contents, so
 * calls that overlap on one GPU (several streams or threads) race; serialise them. */
#define US3D_STATIC_SCRATCH_BYTES (64ll << 20)
int us3d_ug_render_static(const us3d_plan* plan, const float* means, const float* tril, const float* opacity,
                          const float* intensity, const float* pixels, const us3d_ug_settings* settings, float* image,
                          float* alpha, float* depth, float* aux, int32_t* overflow_dev, void* stream);

/* ---------------------------------------------------------------- UltraG-Ray B-mode ray renderer */

typedef struct us3d_ur_settings {
    float k_sigma;                      /* safe radius multiple of max scale */
    float quadrature_window_radius_mm;  /* R: GL3 window around closest approach (0 = whole ray) */
    float elev_cutoff_margin_mm;
    float local_transmission_max_window_mm;
    float grad_threshold;               /* pixels with |dL/dB| < grad_threshold are skipped (1e-6) */
    int32_t local_transmission_window_only;
    int32_t scales_are_log;             /* 1: the scales array holds log-scales. us3d_ur_map applies
                                         * max(exp(x), 1e-5) (pipeline_god_file.py:360-361) and us3d_ur_backward
                                         * returns the gradients w.r.t. the log-scales (chain exp(x) where
                                         * exp(x) >= 1e-5, else 0); 0: it holds the scales. */
    int32_t reserved1;
} us3d_ur_settings;

int64_t us3d_ur_workspace_bytes(int64_t n_gauss, int64_t n_pairs, int64_t n_rays, int64_t k_capacity,
                                int64_t v_capacity);
int us3d_ur_workspace_layout(int64_t n_gauss, int64_t n_pairs, int64_t n_rays, int64_t k_capacity,
                             int64_t v_capacity, us3d_layout* out);

/* Ray mapper: replaces MapperForwardCUDA of ops/ultragray_renderer (src/mapper/wrapper_mapper.cu:7-128,
 * core.cu:3-129). means [G,3], scales [G,3], quats [G,4] (w,x,y,z), attenuations [G] (raw logits),
 * sh [G,4]. Emits (ray, quantised axial depth, Gaussian) instances and sorts each ray's list in LDS.
 * Frames: delta_lat = width_mm / W (FP32), tiles_x = 1, tiles_y = W; ray_base indexes the per-ray arrays. */
int us3d_ur_map(const us3d_plan* plan, const float* means, const float* scales, const float* quats,
                const float* attenuations, const float* sh, const us3d_ur_settings* settings,
                void* stream);

/* B-mode forward: replaces ShaderForwardCUDA of ops/ultragray_renderer (src/shader/wrapper_shader.cu:7-58,
 * forward.cu:22-119). pixels: per frame (W,H,3) world mm (pixel (x,y) at ((x*H)+y)*3), concatenated;
 * ray_origins/ray_dirs [n_rays,3]; ray_offsets [n_rays] elevational OPS offsets (mm) or NULL for none.
 * Each of pixels, ray_origins and ray_dirs may be NULL: the kernels then use the linear probe's geometry
 * described by the frame -- sample (x,y) at origin + ((x - col_offset) delta_lat) a_lat +
 * (near_mm + y delta_ax) a_ax, ray origin origin + ((x - col_offset) delta_lat) a_lat, direction a_ax --
 * evaluated with the FP32 operations of the pipeline's grid (pipeline_god_file.py:292-333), so the results
 * equal those of the explicit arrays that grid produces, without the (P,3) array in HBM.
 * image [n_pixels] in (H,W) row-major per frame; aux [n_pixels,4] = (T, E = num/(den+1e-6), den, 0) per
 * sample in the pixels' ray-major order (sample (x,y) at x*H+y of its frame), accumulated in FP64 and stored
 * as FP32; it is the backward's input, not an image. */
int us3d_ur_forward(const us3d_plan* plan, const float* pixels, const float* ray_origins,
                    const float* ray_dirs, const float* ray_offsets, const us3d_ur_settings* settings,
                    float* image, float* aux, void* stream);

/* B-mode adjoint: replaces ShaderBackwardCUDA of ops/ultragray_renderer (wrapper_shader.cu:60-123,
 * backward.cu:26-325). Geometry arguments as for the forward (NULL = the frame's linear probe). Outputs (overwritten): d_means [G,3], d_scales [G,3], d_quats [G,4],
 * d_attenuations [G], d_sh [G,4]. */
int us3d_ur_backward(const us3d_plan* plan, const float* pixels, const float* ray_origins,
                     const float* ray_dirs, const float* ray_offsets, const us3d_ur_settings* settings,
                     const float* aux, const float* d_image, float* d_means, float* d_scales,
                     float* d_quats, float* d_attenuations, float* d_sh, void* stream);

/* ---------------------------------------------------------------- GaussianPile us_focus slice splat */

/* Settings of the GaussianPile renderer's "ours_focus" backend as its facade configures the reference's
 * ProjectedSliceForwardCUDA (ops/gaussianplie_renderer/gaussianplie_renderer/autograd.py:17-55): shader weight
 * US_FOCUS, tile mapper, additive compositing, constant elevational focus width, det_gamma 0.5. */
typedef struct us3d_gp_settings {
    float k_sigma;      /* radius_sigma: box half-widths k * sqrt(Sigma2_ii + |Sigma2_01|) */
    float sigma_ele;    /* sigma_z_mm: elevational focus width */
    float det_eps;      /* det_eps */
    float weight_min;   /* max(opacity_cull_threshold, 1e-12): smaller weights are dropped; must be > 0 */
    int32_t tile_size;  /* 8, 16 or 32 (tile_size_px) */
    int32_t raw_activations; /* 1: opacity is a logit (the map applies the sigmoid) and scales are log-scales
                                (the map applies max(exp(x), 1e-5)), as GaussianParameters activates them
                                (pipeline_god_file.py:360-370); the backward then returns the gradients w.r.t.
                                those raw values. 0: both are used as given. Quaternions are always normalised
                                by the kernels. Pass the same value to the map, forward and backward. */
    int32_t work_queue; /* 0: one workgroup per (frame, tile); 1: persistent workgroups over per-XCD tile queues (as
                           us3d_ug_settings.work_queue; the outputs do not depend on it) */
    int32_t reserved;
} us3d_gp_settings;

/* Workspace for a batch (same plan structure as UltraGauss: tiles of tile_size px per frame). */
int64_t us3d_gp_workspace_bytes(int64_t n_gauss, int64_t n_pairs, int64_t n_tiles, int64_t k_capacity,
                                int64_t v_capacity);
int us3d_gp_workspace_layout(int64_t n_gauss, int64_t n_pairs, int64_t n_tiles, int64_t k_capacity,
                             int64_t v_capacity, us3d_layout* out);

/* Mapper: replaces the tile mapper of ProjectedSliceForwardCUDA (src/shader/wrapper_shader.cu:107-129 ->
 * OursTileGaussMapperCUDA, src/mapper/core_tile.cu:25-379) for shader weight US_FOCUS. means/scales [G,3],
 * quats [G,4] (w, x, y, z, normalised inside), opacity [G], intensity [G] (NULL: all ones, as the facade
 * passes). Frames use the linear-probe grid of us3d_frame (axes unit length). Per (frame, Gaussian) pair the
 * focus footprint (us_focus_device.cuh:176-300); pairs whose peak weight is below weight_min are not listed
 * (they contribute nothing); each tile lists its Gaussians in index order. */
int us3d_gp_map(const us3d_plan* plan, const float* means, const float* scales, const float* quats,
                const float* opacity, const float* intensity, const us3d_gp_settings* settings, void* stream);

/* Forward: replaces kernel_us_focus_tile2d_additive_forward (src/shader/forward.cu:114-210). image [n_pixels]
 * = per pixel the FP32 sum, in list order, of intensity * o * slice_response * det2^-1/2 * exp(-quad/2) over
 * the listed Gaussians inside their k-sigma box with weight >= weight_min. Requires us3d_gp_map on the same
 * workspace. */
int us3d_gp_forward(const us3d_plan* plan, const us3d_gp_settings* settings, float* image, void* stream);

/* Adjoint: replaces the US_FOCUS branch of ProjectedSliceBackwardCUDA (src/shader/wrapper_shader.cu; kernels in
 * src/shader/us_focus.cu). The exact derivative of the forward for every parameter (the reference takes the
 * scale and quaternion gradients by finite differences of its adjoint, us_focus.cu:695-743). Each (frame, Gaussian)
 * pair's gradients are sanitised before the frames are summed, as the reference sanitises every render call's
 * (us_focus.cu:129-137): non-finite -> 0, magnitudes clamped to 1e3 (kGradClamp, plane_cond2d_device.cuh:12);
 * with raw_activations the chain rules apply after. Outputs (overwritten): d_means [G,3], d_scales [G,3],
 * d_quats [G,4], d_opacity [G], d_intensity [G] (may be NULL). Requires the map of the same call sequence on the
 * same workspace; deterministic (no atomics). */
int us3d_gp_backward(const us3d_plan* plan, const us3d_gp_settings* settings, const float* d_image,
                     float* d_means, float* d_scales, float* d_quats, float* d_opacity, float* d_intensity,
                     void* stream);

/* Training step with the image loss fused into the GaussianPile splat, as us3d_ug_forward_loss /
 * us3d_ug_backward_loss do for the slice splat: us3d_gp_forward_loss renders as us3d_gp_forward and sums the
 * frame statistics of the loss against gt in the same kernel, then writes loss[n_frames] and the frames' gradient
 * coefficients into tmp (us3d_gp_loss_tmp_bytes(plan) bytes); us3d_gp_backward_loss is us3d_gp_backward with
 * d_image formed from that image, gt and the coefficients while the adjoint stages its pixels (same plan, image,
 * gt and tmp). Agrees with us3d_gp_forward + us3d_image_loss + us3d_gp_backward to rounding. No counterpart in
 * the reference (its loss is torch ops). */
int64_t us3d_gp_loss_tmp_bytes(const us3d_plan* plan);
int us3d_gp_forward_loss(const us3d_plan* plan, const us3d_gp_settings* settings, float* image, const float* gt,
                         float w_l1, float w_ssim, float scale, float* loss, double* tmp, void* stream);
int us3d_gp_backward_loss(const us3d_plan* plan, const us3d_gp_settings* settings, const float* image,
                          const float* gt, const double* tmp, float* d_means, float* d_scales, float* d_quats,
                          float* d_opacity, float* d_intensity, void* stream);

/* ---------------------------------------------------------------- training-step helpers */

/* Exclusive scan of n int32 values into out[0..n) and the int64 total into *total_dev (device).
 * tmp must hold us3d_scan_tmp_bytes(n) bytes. Replaces the CUB DeviceScan::ExclusiveSum call sites
 * (ops/ultragauss_renderer/src/mapper/wrapper_mapper.cu:97-102,174-182). */
int64_t us3d_scan_tmp_bytes(int64_t n);
int us3d_scan_i32(const int32_t* in, int32_t* out, int64_t n, int64_t* total_dev, void* tmp,
                  void* stream);

/* Fused loss of pipeline_god_file.py:606-637 over a batch of frames: per frame
 * loss_f = w_l1 * mean|pred - gt| + w_ssim * (1 - global SSIM(pred, gt)) (single-window SSIM, :628-637); writes
 * loss[n_frames] and d_pred = d(scale * sum_f loss_f)/d pred. Frame statistics are reduced in FP64 with a fixed
 * order (bitwise reproducible). frame_pixel_base [n_frames+1]; tmp holds us3d_image_loss_tmp_bytes bytes;
 * 1 <= n_frames <= 65535. */
int64_t us3d_image_loss_tmp_bytes(int32_t n_frames);
int us3d_image_loss(const float* pred, const float* gt, const int32_t* frame_pixel_base, int32_t n_frames,
                    float w_l1, float w_ssim, float scale, float* loss, float* d_pred, double* tmp, void* stream);

/* Adam update with torch.optim.Adam semantics (pipeline_god_file.py:740,764; no weight decay) over n
 * parameters, bias corrections from the device step counter (t = *step_dev + 1). skip_dev (may be NULL):
 * when *skip_dev != 0 the update is skipped -- pass &stats.overflow so a step whose mapping overflowed never
 * touches the parameters. us3d_adam_advance increments *step_dev (unless skipped) once all parameter groups
 * are updated; both are kernels, so a captured graph advances the counter on every replay. */
int us3d_adam_step(float* param, const float* grad, float* exp_avg, float* exp_avg_sq, int64_t n,
                   int32_t* step_dev, const int32_t* skip_dev, float lr, float beta1, float beta2, float eps,
                   void* stream);
int us3d_adam_advance(int32_t* step_dev, const int32_t* skip_dev, void* stream);

/* Fill n int32 values with `value` (a kernel, usable inside captured graphs instead of hipMemset). */
int us3d_fill_i32(int32_t* dst, int64_t n, int32_t value, void* stream);

/* Opt-in per-kernel timing (no counterpart in the reference's CUDA path; the analogue of its
 * ULTRAGRAY_PROFILE stage timers, ops/ultragray_renderer/ultragray_renderer/autograd.py:13-70): while
 * enabled, every kernel launch of the library is bracketed by a hipEvent pair on its stream.
 * us3d_stage_timing_read synchronises on the recorded events, writes one "kernel launches total_ms" line per
 * kernel into buf (NUL-terminated) and clears the records; with buf == NULL it only returns a buffer size
 * that is large enough. */
int us3d_stage_timing(int enable);
int us3d_stage_timing_read(char* buf, int64_t bytes);

const char* us3d_last_error(void);
int us3d_version(void);

#ifdef __cplusplus
}
#endif

#endif /* US3D_H */
