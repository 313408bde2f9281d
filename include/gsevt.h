/*
 * gsevt.h - C ABI of libgsevt_hip.so, the gfx950 (MI355X) rasterizer and
 * event-tracking kernels of GS-EVT.
 *
 * The reference reaches its rasterizer through the Python module
 * diff_gaussian_rasterization, whose pybind11 extension `_C` exposes
 * rasterize_gaussians / rasterize_gaussians_backward / mark_visible.  The
 * module is imported at gaussian_splatting/gaussian_renderer/__init__.py:15-18
 * and called at __init__.py:338 (GaussianRasterizer(raster_settings=...)) and
 * __init__.py:347-375 (rasterizer(means3D=..., ..., theta=, rho=, w=, v=)).
 * Its source (submodule diff-gaussian-rasterization-vwRT, .gitmodules:1-3) is
 * not in the reference tree; the stages below split those two calls into the
 * steps of the published 3D Gaussian Splatting rasterizer.
 *
 * Conventions
 *  - every pointer argument is a device pointer unless stated "host";
 *  - 4x4 matrices use the transposed storage of the rasterizer settings:
 *    element (r, c) of the matrix is m[4 * c + r];
 *  - `stream` is a hipStream_t (NULL: the default stream);
 *  - functions return 0 on success and a nonzero code otherwise, with the
 *    message in gsevt_last_error() (thread-local); no function allocates or
 *    frees device memory: workspaces are passed in, their sizes come from the
 *    *_scratch_elems / *_count queries;
 *  - splat records of the single-render path are `channels == 3 ? 3 : 2`
 *    float4 per Gaussian: {x, y, conic_a, conic_b}, {conic_c, opacity, depth,
 *    c0}[, {c1, c2, 0, 0}] (the batched path's records: gsevt_tr_preprocess).
 */
#ifndef GSEVT_H_
#define GSEVT_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- errors ------------------------------------------------------------ */
const char* gsevt_last_error(void);
void gsevt_clear_error(void);
int gsevt_abi_version(void);

/* ---- single render (the diff_gaussian_rasterization drop-in) ----------- */

/* Preprocess: cull (view z <= 0.2), EWA 2D covariance (+0.3 px^2), conic,
 * radius ceil(3 sqrt(lambda_max)), tile rect, SH -> RGB (+0.5, clamped at 0)
 * or grey intensity (channels == 1, weights default 0.2989/0.5870/0.1140),
 * depth.  Exactly one of shs / colors_precomp and one of (scales, rotations)
 * / cov3D_precomp must be non-null (the reference raises otherwise).
 * Replaces the preprocess stage of _C.rasterize_gaussians, called from
 * gaussian_renderer/__init__.py:362-375. */
int gsevt_preprocess(int P, int D, int M, const float* means3D, const float* scales, float scale_modifier,
                     const float* rotations, const float* opacities, const float* shs, const float* colors_precomp,
                     const float* cov3D_precomp, const float* viewmatrix, const float* projmatrix,
                     const float* campos, int W, int H, float tanfovx, float tanfovy, int tile_w, int tile_h,
                     int channels, const float* intensity_weights /* host, 3 or NULL */, int* radii,
                     uint32_t* tiles_touched, float* splats, void* stream);

/* Inclusive scan of n uint32 (tiles_touched -> offsets); total_out receives
 * the last value (num_rendered).  Replaces the InclusiveSum step of
 * _C.rasterize_gaussians. */
long long gsevt_scan_scratch_elems(long long n);
int gsevt_scan_inclusive_u32(const uint32_t* in, uint32_t* out, long long n, uint32_t* scratch,
                             uint32_t* total_out, void* stream);

/* keys = tile << 32 | float_bits(depth), values = Gaussian index, for every
 * tile of every visible Gaussian's rect (row-major), at the scanned offsets.
 * Replaces duplicateWithKeys of _C.rasterize_gaussians. */
int gsevt_duplicate_with_keys(int P, const uint32_t* offsets_inclusive, const int* radii, const float* splats,
                              int channels, int W, int H, int tile_w, int tile_h, uint64_t* keys, uint32_t* values,
                              long long capacity, void* stream);

/* Stable LSD radix sort of (u64 key, u32 value) on bits [0, key_bits); the
 * count is read from the device (d_count), the grid is sized for capacity.
 * Result in keys/values (even number of passes).  Replaces the
 * DeviceRadixSort::SortPairs step of _C.rasterize_gaussians. */
long long gsevt_sort_scratch_elems(long long capacity);
int gsevt_sort_passes(int key_bits);
int gsevt_sort_pairs(uint64_t* keys, uint32_t* values, uint64_t* keys_alt, uint32_t* values_alt,
                     const uint32_t* d_count, long long capacity, int key_bits, uint32_t* scratch, void* stream);

/* ranges[2 t], ranges[2 t + 1] = first / one-past-last sorted fragment of
 * tile t (zero for empty tiles).  Replaces identifyTileRanges. */
int gsevt_tile_ranges(const uint32_t* d_num_rendered, long long capacity, const uint64_t* sorted_keys, int n_tiles,
                      uint32_t* ranges, void* stream);

/* Front-to-back blend per tile (alpha = min(0.99, o exp(power)), skip
 * alpha < 1/255 or power > 0, stop before T < 1e-4): colour (channels x H x
 * W, + T bg), depth = sum z alpha T, opacity = sum alpha T, final T, the
 * number of fragments up to the last blended one, and (optional) per
 * Gaussian the number of pixels it was blended into.  Replaces the render
 * stage of _C.rasterize_gaussians; outputs are the 5-tuple of
 * gaussian_renderer/__init__.py:362 (render, radii, depth, opacity,
 * n_touched). */
int gsevt_blend_forward(int W, int H, int tile_w, int tile_h, const uint32_t* ranges, const uint32_t* values,
                        const float* splats, int channels, const float* bg /* channels floats or NULL */, float* out_color,
                        float* out_depth, float* out_opacity, float* final_T, int* n_contrib, int* n_touched,
                        void* stream);

/* Reverse replay of the blend (T recovered by division); per Gaussian the
 * gradient of its 2D attributes, 10 floats: d mean (x, y; pixel units),
 * d conic (a, b, c), d opacity, d colour (3, or 1 in slot 6), d depth.
 * grad10 must be zero on entry.  Replaces the render stage of
 * _C.rasterize_gaussians_backward. */
int gsevt_blend_backward(int W, int H, int tile_w, int tile_h, const uint32_t* ranges, const uint32_t* values,
                         const float* splats, int channels, const float* bg /* channels floats or NULL */, const float* final_T,
                         const int* n_contrib, const float* dL_dcolor, const float* dL_ddepth,
                         const float* dL_dopacity, float* grad10, void* stream);

/* Chain the 2D gradients to view space and reduce the camera-pose partial
 * sums on device, 18 floats per workgroup (sum dL/dp_view, sum p x dL/dp,
 * dL/dW 3x3, sum dL/dcampos); optional per-Gaussian map gradients (NULL to
 * skip; dL_dmeans2D in NDC units).  Replaces the preprocess stage of
 * _C.rasterize_gaussians_backward. */
int gsevt_pose_partials_count(int P);
int gsevt_preprocess_backward(int P, int D, int M, const float* means3D, const float* scales, float scale_modifier,
                              const float* rotations, const float* shs, const float* colors_precomp,
                              const float* cov3D_precomp, const float* viewmatrix, const float* projmatrix,
                              const float* projmatrix_raw, const float* campos, int W, int H, float tanfovx,
                              float tanfovy, int tile_w, int tile_h, const int* radii, const float* grad10,
                              int channels, const float* intensity_weights /* host or NULL */, float* pose_partials,
                              float* dL_dmeans3D, float* dL_dscales, float* dL_drotations, float* dL_dshs,
                              float* dL_dopacity, float* dL_dmeans2D, float* dL_dcolors, float* dL_dcov3D,
                              void* stream);

/* Deterministic FP64 reduction of the partials and the exact adjoint of the
 * render pose T_r = Exp(dt [v; w]) Exp([rho; theta]) T_c:
 *   out12[0..2] dL/d theta, [3..5] dL/d rho, [6..8] dL/d w, [9..11] dL/d v
 * (the grads of cam_rot_delta, cam_trans_delta, cam_w_delta, cam_v_delta
 * that gaussian_renderer/__init__.py:301-308 route through the rasterizer).
 * vel_transform is the settings' vel_transofrm (NULL: identity), dt its
 * delta_time.  accumulate != 0 adds to out12. */
int gsevt_pose_finalize(const float* pose_partials, int n_partials, const float* viewmatrix,
                        const float* vel_transform, double delta_time, const float* linear_vel,
                        const float* angular_vel, float* out12, int accumulate, void* stream);

/* present[i] = 1 unless the view-space z of means3D[i] is <= 0.2, else 0
 * (uint8 per Gaussian; P == 0 launches nothing).  Replaces _C.mark_visible
 * behind GaussianRasterizer.markVisible of the rasterizer module imported at
 * gaussian_renderer/__init__.py:15-18; projmatrix is accepted for the same
 * signature and does not change the result. */
int gsevt_mark_visible(int P, const float* means3D, const float* viewmatrix, const float* projmatrix,
                       uint8_t* present, void* stream);

int gsevt_fill_u32(uint32_t* p, long long n, uint32_t value, void* stream);

/* ---- event frames (utils/event_camera/event.py:116-128, tracker.py:78-91) */

/* count[y * W + x] += polarity ? +1 : -1 (int32 atomics; exact).  Events
 * outside the image are skipped and counted in *rejected (optional).
 * Replaces the per-event Python loop at event.py:119-120. */
int gsevt_event_scatter(int n, const int* x, const int* y, const int* polarity, int W, int H, int* count,
                        int* rejected, void* stream);

/* Undistortion map (Brown-Conrady k1, k2, p1, p2, k3; new camera matrix = K)
 * in 1/32-pixel fixed point, int2 per destination pixel.  K, dist5 are
 * device arrays of 9 and 5 doubles.  With gsevt_event_frame it replaces
 * cv2.undistort at event.py:121. */
int gsevt_undistort_map(int W, int H, const double* K, const double* dist5, int* map_xy, void* stream);

/* count -> remap -> 9x9 Gaussian blur (sigma 0 -> 1.7, replicate border) ->
 * unit L2 norm -> signed frame and |frame| (event.py:121-126). */
long long gsevt_event_frame_scratch_bytes(int W, int H);
int gsevt_event_frame(int W, int H, const int* count, const int* map_xy, int ksize, double sigma, void* scratch,
                      float* out_signed, float* out_abs, void* stream);

/* gsevt_event_frame without a workspace argument and without any allocation:
 * the scratch is a static device array of the library, one per GPU, held
 * until the process ends (GSEVT_EVENT_FRAME_NOALLOC_MAX_PIXELS * 8 + 2048
 * bytes, about 10.5 MB).  No counterpart in the reference: it is for callers
 * that cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  W * H above the limit is rejected with an
 * error.  Calls on one GPU share its array: calls that overlap on one GPU
 * (concurrent streams or threads) race and must be serialised by the caller.
 * Output equals gsevt_event_frame's bit for bit. */
#define GSEVT_EVENT_FRAME_NOALLOC_MAX_PIXELS (1280 * 1024)
int gsevt_event_frame_noalloc(int W, int H, const int* count, const int* map_xy, int ksize, double sigma,
                              float* out_signed, float* out_abs, void* stream);

/* Nearest-neighbour resize, src = floor(dst * sw / dw) (cv2.INTER_NEAREST,
 * tracker.py:87-88). */
int gsevt_resize_nearest(int sw, int sh, const float* src, int dw, int dh, float* dst, void* stream);

/* ---- batched tracking iteration (items = hypothesis x window) ----------- *
 * The stages below run one tracking iteration (utils/tracker.py:204-222:
 * RenderFrame -> tracking_loss -> backward -> Adam step -> update_vwRT) for
 * B items at once, R = 2B renders per launch, with no host synchronisation;
 * the sequence can be captured in a HIP graph.  Item state lives in device
 * arrays described by gsevt_item_state. */
typedef struct gsevt_item_state {
  double* pose;        /* [B][12] world->camera R (row-major 9) and t (3) */
  float* lin;          /* [B][3] linear velocity */
  float* ang;          /* [B][3] angular velocity */
  float* adam_m;       /* [B][12] first moments (theta, rho, w, v) */
  float* adam_v;       /* [B][12] second moments */
  int* adam_t;         /* [B][4] Adam step per parameter group */
  int* iter;           /* [B] iteration inside the stage */
  float* dtau;         /* [B] window duration (s) */
  const int* window;   /* [B] event frame of the item */
  float* loss;         /* [B] last loss */
  float* loss_hist;    /* [B][hist_len] */
  int hist_len;
  float* grad12;       /* [B][12] last gradient */
  /* Stage of each item for gsevt_tr_update_staged and unsigned_mode = -1
   * (NULL otherwise): 0 coarse (pose only, unsigned images), 1 fine (pose
   * and velocity, signed images), 2 done (no further updates). */
  int* stage;          /* [B] */
  int* start_vel;      /* [B] iteration at which the fine stage started */
} gsevt_item_state;

int gsevt_tr_camera_record_bytes(void);
/* 3D covariance of the frozen map, once (gaussian_model.py:52-60). */
int gsevt_tr_map_cov(int P, const float* scales, float scale_modifier, const float* rotations, float* cov6,
                     void* stream);
/* Camera records of the last / next render of each item:
 * T_r = Exp(-/+ dtau/2 [v; w]) T_c (camera.py:100-123, frame.py:68-83). */
int gsevt_tr_cams(int B, const gsevt_item_state* state /* host struct */, const float* projraw16 /* host */,
                  float tanfovx, float tanfovy, int W, int H, int tile_w, int tile_h, void* cams, float* vel,
                  float* dts, void* stream);
/* Preprocess of R renders of one map (preprocess of the reference rasterizer,
 * called at gaussian_renderer/__init__.py:362-375), grey intensity folded per
 * Gaussian (frame.py:39-44).  shs: coefficient-major [3K][P] (row k holds
 * coefficient k/3, channel k%3 of every Gaussian), K = (D+1)^2.  Splat
 * records: {x, y, conic_a, conic_b}, {conic_c, opacity, rect, intensity} with
 * rect the bits of x0 | y0 << 8 | x1 << 16 | y1 << 24 (tile bounds; grids of
 * at most 255 x 255 tiles, checked by gsevt_tr_cams); depths go to dkey.
 * rects (x0 | y0 << 16, x1 | y1 << 16 per Gaussian) may be NULL when only
 * packed emission follows.  dval receives the Gaussian index, with its rect's
 * tile count above bit count_shift when count_shift > 0 (2^count_shift >= P;
 * the caller keeps n_tiles < 2^(32 - count_shift)); tiles may then be NULL
 * (a packed emission with counts_packed reads the counts from dval). */
int gsevt_tr_preprocess(int R, int P, int D, int M, const float* means, const float* cov6, const float* opac,
                        const float* shs, const float* intensity_weights3 /* host */, const void* cams, float* splats,
                        uint32_t* rects, uint32_t* tiles, uint32_t* dkey, uint32_t* dval, int count_shift,
                        void* stream);
/* Stable segmented LSD sort of u32 keys/values (R segments, stride apart). */
long long gsevt_seg_sort_scratch_elems(int R, long long stride);
int gsevt_seg_sort_u32(int R, long long stride, const uint32_t* d_counts, long long n_fixed, uint32_t* keys,
                       uint32_t* vals, uint32_t* keys_alt, uint32_t* vals_alt, int key_bits, uint32_t* scratch,
                       void* stream);
/* As gsevt_seg_sort_u32 on key bits [begin_bit, end_bit); vals may be null
 * (keys-only).  Replaces the tile half of the 64-bit key sort
 * (cub::DeviceRadixSort in the reference rasterizer) for packed fragments.
 * kept_counts non-null: the first pass also removes every key equal to
 * 0xFFFFFFFF (a dropped fragment) and kept_counts[r] (device) receives the
 * number of keys segment r keeps; the result is the first kept_counts[r]
 * slots of each segment.  first_hist_ready != 0: the first pass's digit
 * histogram is already at the start of scratch, as gsevt_tr_emit_packed
 * writes it with sort_hist = scratch (that pass then reads no keys to count
 * them); the first pass must take at most 6 bits. */
int gsevt_seg_sort_u32_bits(int R, long long stride, const uint32_t* d_counts, long long n_fixed, uint32_t* keys,
                            uint32_t* vals, uint32_t* keys_alt, uint32_t* vals_alt, int begin_bit, int end_bit,
                            uint32_t* scratch, uint32_t* kept_counts, int first_hist_ready, void* stream);
/* Fragments in depth order (depth sort first, then a stable tile sort gives
 * the order of the 64-bit tile|depth sort). */
long long gsevt_tr_bin_scratch_elems(int R, int P);
int gsevt_tr_emit(int R, int P, const uint32_t* dval_sorted, const uint32_t* tiles, const uint32_t* rects, int W,
                  int tile_w, long long cap, uint32_t* tiles_sorted, uint32_t* incl, uint32_t* scan_scratch,
                  uint32_t* nfrag, int* overflow, uint32_t* ftile, uint32_t* fgid, void* stream);
int gsevt_tr_ranges(int R, long long cap, const uint32_t* ftile_sorted, const uint32_t* nfrag, int n_tiles,
                    uint32_t* ranges, void* stream);
/* Packed fragments: as gsevt_tr_emit with one u32 per fragment,
 * tile << gid_bits | gaussian (needs n_tiles << gid_bits < 2^32), the
 * duplicateWithKeys counterpart; sort them with gsevt_seg_sort_u32_bits on
 * bits [gid_bits, gid_bits + tile bits).  splats: the tr_preprocess records,
 * whose third float of the second float4 holds the tile rect (one byte per
 * bound: x0 | y0 << 8 | x1 << 16 | y1 << 24).  cull != 0: a tile of a
 * Gaussian's rect that the blend's conservative cull rejects on the whole tile
 * (alpha < 1/255 at all its pixels) gets the dropped key 0xFFFFFFFF instead,
 * for the sort to remove (kept_counts); renders are unchanged, as such a
 * fragment contributes to no pixel.  nfrag counts every slot of the rects (the
 * capacity check).  counts_packed != 0: dval_sorted carries the tile counts
 * above bit gid_bits (gsevt_tr_preprocess with count_shift = gid_bits),
 * scanned in place (tiles and tiles_sorted are then not used and may be
 * NULL) instead of
 * gathered from tiles into tiles_sorted.  sort_hist non-null: also writes the
 * histogram of the first pass of the tile sort that follows (the low
 * hist_bits bits of the tile of every kept fragment, per 4096-slot block of
 * each render, at sort_hist[(r << hist_bits | digit) * ceil(cap / 4096) +
 * block]), for gsevt_seg_sort_u32_bits with scratch = sort_hist and
 * first_hist_ready = 1; hist_bits (1..6) must equal the bits of that sort's
 * first pass. */
int gsevt_tr_emit_packed(int R, int P, const uint32_t* dval_sorted, const uint32_t* tiles, const float* splats,
                         int cull, int W, int H, int tile_w, int tile_h, int gid_bits, int counts_packed,
                         long long cap,
                         uint32_t* tiles_sorted, uint32_t* incl, uint32_t* scan_scratch, uint32_t* nfrag,
                         int* overflow, uint32_t* frag, uint32_t* sort_hist, int hist_bits, void* stream);
/* identifyTileRanges counterpart on tile-sorted packed fragments (binary
 * search per tile; empty tiles (0, 0)). */
int gsevt_tr_ranges_packed(int R, long long cap, const uint32_t* frag, const uint32_t* nfrag, int gid_bits,
                           int n_tiles, uint32_t* ranges, void* stream);
/* Grey-intensity blend of every render (frame.py:39-44 folded per Gaussian). */
/* gid_mask: fragment value & gid_mask is the Gaussian (0xFFFFFFFF for plain
 * ids, (1 << gid_bits) - 1 for packed fragments). */
int gsevt_tr_blend_forward(int R, const float* projraw16, float tanfovx, float tanfovy, int W, int H, int tile_w,
                           int tile_h, int P, long long cap, const uint32_t* ranges, const uint32_t* fgid,
                           uint32_t gid_mask, const float* splats, float bg_intensity, float* inten, float* final_T,
                           int* n_contrib, void* stream);
/* || dI/||dI|| - e || per item (frame.py:88-92, tracker.py:93-103); signed or
 * unsigned (|.| of both the render difference and the event frame) images;
 * unsigned_mode = -1 takes it per item from state->stage (0: unsigned). */
int gsevt_tr_loss_partials_count(void);
int gsevt_tr_loss(int B, int W, int H, const float* inten, const float* events, const gsevt_item_state* state,
                  int unsigned_mode, double* partials, float* scal, void* stream);
/* Backward blend of every render: accumulates into grad6 ([R][P][6]: dL/d
 * mean2 x, y, dL/d conic a, b, c, dL/d intensity).  clear_grad6 != 0 clears
 * it first; 0 requires it zero on entry, as gsevt_tr_preprocess_backward with
 * clear_grad6 != 0 leaves it (a cleared buffer stays zero from one iteration
 * to the next without a pass over all of it). */
int gsevt_tr_blend_backward(int R, const float* projraw16, float tanfovx, float tanfovy, int W, int H, int tile_w,
                            int tile_h, int P, long long cap, const uint32_t* ranges, const uint32_t* fgid,
                            uint32_t gid_mask, const float* splats, float bg_intensity, const float* inten,
                            const float* final_T,
                            const int* n_contrib, const float* events, const gsevt_item_state* state,
                            const float* scal, int unsigned_mode, float* grad6, int clear_grad6, void* stream);
/* Persistent forms of the two blends above: `workgroups` workgroups pull
 * (render, tile) items from device counters until all are taken, instead of
 * one workgroup per item.  xcd_queues != 0: 8 queues, queue q holding the
 * renders r = q mod 8, each workgroup starting on the queue of the XCD it
 * runs on (workgroup w on XCD w mod 8) and moving to the next queue when its
 * own is empty; otherwise one queue in render-major order.  queue: device
 * array of gsevt_tr_blend_queue_count() uint32 counters (cleared by the
 * call).  Results equal those of the grid launches (the forward bit for bit;
 * the backward's float atomics within rounding).  Replaces the same render
 * and backward stages of the reference rasterizer as the grid forms. */
int gsevt_tr_blend_queue_count(void);
int gsevt_tr_blend_forward_queue(int R, const float* projraw16, float tanfovx, float tanfovy, int W, int H,
                                 int tile_w, int tile_h, int P, long long cap, const uint32_t* ranges,
                                 const uint32_t* fgid, uint32_t gid_mask, const float* splats, float bg_intensity,
                                 float* inten, float* final_T, int* n_contrib, uint32_t* queue, int workgroups,
                                 int xcd_queues, void* stream);
int gsevt_tr_blend_backward_queue(int R, const float* projraw16, float tanfovx, float tanfovy, int W, int H,
                                  int tile_w, int tile_h, int P, long long cap, const uint32_t* ranges,
                                  const uint32_t* fgid, uint32_t gid_mask, const float* splats, float bg_intensity,
                                  const float* inten, const float* final_T, const int* n_contrib,
                                  const float* events, const gsevt_item_state* state, const float* scal,
                                  int unsigned_mode, float* grad6, int clear_grad6, uint32_t* queue, int workgroups,
                                  int xcd_queues, void* stream);
int gsevt_tr_pose_partials_count(int P);
/* Pose partials of every render from grad6 (preprocessBackward of the
 * reference rasterizer, reduced per workgroup; Gaussians whose six gradients
 * are all zero add nothing).  clear_grad6 != 0: every non-zero entry is set
 * back to zero after it is read, so grad6 leaves all zero. */
int gsevt_tr_preprocess_backward(int R, int P, int D, int M, const float* means, const float* cov6,
                                 const float* opac, const float* shs, const float* intensity_weights3,
                                 const void* cams, float* grad6, float* partials, int clear_grad6, void* stream);
/* Adjoint + torch-default Adam (lr4: theta, rho, w, v; fine-stage ramp of
 * tracker.py:188-202 when opt_vel) + update_vwRT (camera.py:129-155). */
int gsevt_tr_update(int B, int n_partials, const float* partials, const void* cams, const float* vel,
                    const float* dts, const gsevt_item_state* state, const float* lr4 /* host */,
                    int max_optim_iter, int opt_vel, int start_vel_iter, int do_step, void* stream);
/* gsevt_tr_update with the stage logic of one pyramid level of
 * utils/tracker.py:149-240 run per item on the device (state->stage and
 * state->start_vel required): coarse items step theta and rho at the base LR,
 * fine items all four groups with the LR ramp from start_vel; after the step,
 * with the iteration's loss in loss_hist[iter], convergence is more than 10
 * losses with mean |diff| of the last 11 below converged_threshold
 * (tracker.py:65-76); a converged coarse item turns fine (start_vel = iter), a
 * converged fine item is done; a coarse item at iter >= max_optim_iter or a
 * fine one at iter >= start_vel + max_optim_iter is done; done items are not
 * updated again. */
int gsevt_tr_update_staged(int B, int n_partials, const float* partials, const void* cams, const float* vel,
                           const float* dts, const gsevt_item_state* state, const float* lr4 /* host */,
                           int max_optim_iter, float converged_threshold, void* stream);

#ifdef __cplusplus
}  /* extern "C" */
#endif

#endif  /* GSEVT_H_ */
