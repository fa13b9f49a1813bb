/*
 * toed_capi.h -- C ABI of the MI355X (gfx950) TOED detection library
 * (libtoed_hip.so).  Plain pointers and sizes only; device pointers are HIP
 * device allocations owned by the caller unless stated otherwise.
 *
 * Reference interfaces replaced (paths relative to the reference tree):
 *   gpu_kernels.hpp:24-45   gpu_convolve(...) f32/f64 overloads  -> toed_gpu_convolve_f32/_f64
 *   gpu_kernels.hpp:48-63   gpu_nms(...) f32/f64 overloads       -> toed_gpu_nms_f32/_f64
 *   gpu_toed.hpp:79-135     ThirdOrderEdgeDetectionGPU ctor: device buffers,
 *                           events                               -> toed_create / toed_plan
 *   gpu_toed.hpp:340-441    convolve_img + non_maximum_suppresion + host
 *                           build_edge_list (:313-338)           -> toed_detect (one fused
 *                           launch sequence producing the raster-ordered N x 4 list on the GPU)
 *   run_gpu_batch.sh:152-167 one process per image              -> toed_batch_* (one plan,
 *                           one launch sequence, optionally one captured hipGraph)
 *   curvelet/form_curvelet_process.hpp:23-82 form_curvelet_process (host) -> toed_curvelet_*
 * A C ABI cannot overload, so the f32 variants carry a suffix.
 *
 * Every function returns a toed_status (0 = success).  Launch functions are
 * asynchronous on the context's stream unless named *_sync.
 */
#ifndef TOED_CAPI_H
#define TOED_CAPI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    TOED_OK = 0,
    TOED_ERR_ARG = 1,          /* invalid argument (null pointer, bad size, ...) */
    TOED_ERR_SIGMA = 2,        /* sigma outside the supported window radii */
    TOED_ERR_ALLOC = 3,        /* device allocation failed */
    TOED_ERR_HIP = 4,          /* a HIP runtime call failed */
    TOED_ERR_CAPACITY = 5,     /* more edgels than the output capacity (count is still exact) */
    TOED_ERR_STAGING = 6,      /* per-tile staging overflow (re-plan with a larger staging fraction) */
    TOED_ERR_NOT_BUILT = 7     /* no device code for this GPU */
} toed_status;

/* input pixel formats */
enum { TOED_U8 = 0, TOED_F32 = 1 };

/* per-image status bits written by the kernels */
enum { TOED_FLAG_CAPACITY = 1, TOED_FLAG_STAGING = 2 };

typedef struct toed_ctx toed_ctx;
typedef struct toed_batch toed_batch;

/* One image of a detection call.  Edgels are written as rows of 4 floats
 * [x, y, orientation, strength] (reference layout, indices.hpp:40) in the
 * reference's raster order of the 2x interpolated grid (cpu_toed.cpp:510-538). */
typedef struct {
    int in_type;             /* TOED_U8 or TOED_F32 */
    const void *d_img;       /* device pointer, H rows of `pitch` elements */
    int H, W, pitch;         /* pitch in elements, >= W */
    float sigma;             /* Gaussian scale; windows R0 = ceil(4 sigma), R1 = R0 + 1 */
    float thresh;            /* gradient-magnitude threshold (reference: 2, cpu_toed.cpp:403) */
    int border;              /* NMS border on the interp grid (reference: 10, cpu_toed.cpp:400) */
    float *d_edges;          /* device output, capacity x 4 floats */
    long long capacity;      /* rows available in d_edges */
    int *d_index;            /* optional (may be NULL): interp-grid index row*(2W)+col of
                                each edgel, capacity ints (no reference counterpart) */
} toed_image_t;

/* Image files ---------------------------------------------------------------
 * The reference reads its input in main_gpu_cpu.cpp:114-132 (cv::imread, or
 * `>>` on the PGM header and one get() per pixel, gpu_toed.hpp:189-198).
 * toed_image_load decodes an 8-bit PGM (P5, maxval <= 255) or a PNG (8/16-bit
 * grey, grey + alpha, RGB, RGBA, not interlaced; alpha dropped, 16-bit samples
 * keep their high byte, colour to grey with BT.601 weights) into an H x W
 * buffer the caller releases with toed_image_free.  pgm_raw_offset = 1 reads a
 * PGM as the reference's no-OpenCV path does: its `>>` leaves the whitespace
 * byte after the header in the stream, so pixel k is the file's byte k - 1. */
int toed_image_load(const char *path, int pgm_raw_offset, unsigned char **pixels, int *H, int *W);
void toed_image_free(unsigned char *pixels);

/* Library / device queries -------------------------------------------------- */
int toed_version(void);
/* sigma -> window radii as used by the kernels; TOED_ERR_SIGMA if unsupported */
int toed_sigma_radius(float sigma, int *R0, int *R1);

/* Context ------------------------------------------------------------------- */
/* Creates a context on `device` (made current while the context is used) with
 * its own HIP stream.  Replaces the device half of gpu_toed.hpp:79-135. */
int toed_create(int device, toed_ctx **out);
int toed_destroy(toed_ctx *ctx);
/* Launch on an external HIP stream (hipStream_t passed as void*); NULL selects
 * the legacy default stream.  toed_use_own_stream switches back to the
 * context's own (blocking) stream. */
int toed_set_stream(toed_ctx *ctx, void *hip_stream);
int toed_use_own_stream(toed_ctx *ctx);
void *toed_get_stream(toed_ctx *ctx);
int toed_get_device(toed_ctx *ctx);
/* Tile rows per workgroup (8 or 16); 0 = autotuned default. */
int toed_set_tile_rows(toed_ctx *ctx, int tile_rows);
/* Where the third-order orientation is evaluated, for later detections and
 * plans: 0 (default) at the edgel's interp pixel, as cpu_toed.cpp:227-232 and
 * the CUDA path do; 1 at its subpixel location, the nine responses
 * interpolated by cubic convolution (Keys, a = -1/2) from the 4 x 4 interp
 * pixels around it, as the MATLAB detector's QaD path does
 * (third_order_edge_detector.m:91-121).  Positions and strengths are the same
 * in both modes. */
int toed_set_orientation_mode(toed_ctx *ctx, int mode);
/* Staging records per tile and phase as a fraction of the phase's pixels
 * (0 < frac <= 1; 1 can never overflow).  Applies to later plans. */
int toed_set_staging_fraction(toed_ctx *ctx, float frac);
/* Last kernel timings of toed_detect_sync, in ms (reference time_conv /
 * time_nms, gpu_toed.hpp:57): tile kernel, and scan + finalize. */
int toed_last_times(toed_ctx *ctx, float *ms_tiles, float *ms_finalize);

/* Single image ---------------------------------------------------------------
 * Asynchronous: enqueues the fused pipeline; *d_count (device int) receives N
 * and *d_flags (device int, may be NULL) the TOED_FLAG_* bits. */
int toed_detect(toed_ctx *ctx, const toed_image_t *img, int *d_count, int *d_flags);
/* Synchronous convenience: returns N in *n_out and maps flags to a status. */
int toed_detect_sync(toed_ctx *ctx, const toed_image_t *img, long long *n_out);
/* toed_detect without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, kernels and results as toed_detect; its scratch (taps, counts,
 * staging) comes from one array of TOED_NOALLOC_SCRATCH_BYTES that the library
 * holds on each GPU until the process ends, and staging is capped at 64
 * records per tile and phase (TOED_FLAG_STAGING when a tile has more).  That
 * holds a 2160 x 3840 frame at the default 16-row tiles; an image needing more
 * returns TOED_ERR_ARG.  Every call on a GPU uses the same array: calls that
 * overlap on one GPU (two streams, two host threads) race and must be
 * serialised by the caller.  The context must have been created by
 * toed_create, which resolves the array's address on its device. */
#define TOED_NOALLOC_SCRATCH_BYTES (96u << 20)
int toed_detect_noalloc(toed_ctx *ctx, const toed_image_t *img, int *d_count, int *d_flags);

/* Batch: many images of uneven sizes and sigma in one launch sequence ------- */
int toed_batch_plan(toed_ctx *ctx, int n, const toed_image_t *imgs, toed_batch **out);
/* 0: one launch per kernel over all tiles of all images; 1: persistent tile
 * kernel over a device work queue. */
int toed_batch_set_persistent(toed_batch *b, int persistent);
/* Enqueue; per-image counts and flags land in d_counts[n] / d_flags[n] (device). */
int toed_batch_run(toed_batch *b, int *d_counts, int *d_flags);
/* Capture the launch sequence once in a hipGraph and replay it. */
int toed_batch_graph_run(toed_batch *b, int *d_counts, int *d_flags);
int toed_batch_destroy(toed_batch *b);
/* Persistent grid size: workgroups per CU for the persistent tile kernel
 * (0 = as many as fit on a CU, the default).  No reference counterpart. */
int toed_batch_set_grid(toed_batch *b, int wg_per_cu);
/* Point the plan's images at other device buffers of the same type, sizes and
 * pitches (d_imgs[m]: image m, the order of toed_batch_plan), for the next
 * runs: a plan is reused for new inputs without being rebuilt.  No reference
 * counterpart. */
int toed_batch_set_images(toed_batch *b, const void *const *d_imgs);
/* Compact output for the next runs: every image's rows go to one buffer of
 * `capacity` rows, image m's rows right after those of images 0 .. m-1 (its
 * first row is the sum of their counts, computed on the device), so the
 * concatenated list needs no host step; d_index (may be NULL) likewise.  The
 * per-image d_edges / d_index / capacity of the plan are no longer used; an
 * image whose rows would pass `capacity` gets TOED_FLAG_CAPACITY.  Runs then
 * need d_counts.  No reference counterpart. */
int toed_batch_set_compact_output(toed_batch *b, float *d_edges, int *d_index, long long capacity);

/* Autotuner of the batched form (no reference counterpart: the reference runs
 * a process per image, run_gpu_batch.sh:152-167).  Plans the batch with 8 and
 * 16 tile rows (the halo follows from each image's sigma) and times every
 * launch configuration -- one workgroup per tile, and a persistent grid of 1 ..
 * (as many as fit) workgroups of 4 waves per CU -- as reps hipGraph replays
 * on the context's stream; a configuration whose staging overflows is
 * skipped.  The fastest goes to *best and the context keeps its tile rows for
 * later plans; every configuration timed goes to results[0 .. *n_results)
 * (at most max_results; results and n_results may be NULL).  The images'
 * outputs are overwritten. */
typedef struct {
    int tile_rows;           /* 8 or 16 */
    int persistent;          /* 0: one workgroup per tile; 1: persistent grid */
    int wg_per_cu;           /* persistent workgroups per CU (0 when not persistent) */
    float ms;                /* mean time of one replay */
} toed_tune_t;
int toed_batch_autotune(toed_ctx *ctx, int n, const toed_image_t *imgs, int reps, toed_tune_t *best,
                        toed_tune_t *results, int max_results, int *n_results);

/* Differentiable form -------------------------------------------------------
 * No reference counterpart (the reference has no differentiation).  Adjoint
 * of a forward detection with its edgel set frozen: for each image, the
 * forward's interp indices and count, and the upstream gradient of every
 * output row dL/d[x, y, orientation, strength]; accumulates (+=) dL/dimage
 * into d_gimg and dL/dsigma into *d_gsigma (window radii held fixed).
 * Everything is recomputed from the image in FP64 (no stored planes). */
typedef struct {
    int in_type;             /* TOED_U8 or TOED_F32, as in the forward */
    const void *d_img;
    int H, W, pitch;
    float sigma;
    const int *d_index;      /* forward's d_index (interp index per edgel) */
    const int *d_count;      /* forward's edgel count (device int) */
    int max_edgels;          /* launch bound, >= the count */
    const float *d_gout;     /* max_edgels x 4 upstream gradient rows */
    float *d_gimg;           /* H rows of gpitch floats, accumulated */
    int gpitch;
    double *d_gsigma;        /* accumulated, may be NULL */
} toed_grad_t;
int toed_backward(toed_ctx *ctx, int n, const toed_grad_t *items);

/* The same adjoint for the last forward run of a batch plan, tile by tile from
 * the plan's own staging (no index list; LDS-accumulated gradient patches).
 * outs[m] is image m's: upstream rows (its d_edges layout), dL/dimage
 * (accumulated), dL/dsigma (accumulated, may be NULL).  Valid until the plan
 * runs again.  No reference counterpart. */
typedef struct {
    const float *d_gout;
    float *d_gimg;
    int gpitch;
    double *d_gsigma;
} toed_grad_out_t;
int toed_batch_backward(toed_batch *b, const toed_grad_out_t *outs);

/* Subpixel third-order correction of dense edge maps ------------------------
 * Replaces SubpixelTOCorrectionCPU<T> (subpix_TO_correction/
 * subpix_to_correction.hpp:20-96) (interp_n: toed_subpix_plan_n): set_edgemap (:86),
 * set_orientation (:87) or estimate_orientation (:88), then run (:95), which
 * returns the N x 4 list [x, y, orientation, strength] and fills TO_edgemap /
 * TO_orientation (:71-72).  Many maps of any sizes go through one plan and
 * one launch sequence; sigma and the isolation radius are per plan, the NMS
 * threshold per map. */
typedef struct toed_subpix toed_subpix;
typedef struct {
    const float *d_map;        /* H rows of `pitch` floats: dense edge-strength map */
    const float *d_orient;     /* orientation map in radians (rows of `opitch`), or
                                  NULL: estimated from the map (:307-375) */
    int H, W, pitch, opitch;
    double thresh;             /* NMS threshold on the map (reference default 0.2) */
    float *d_edges;            /* device output, capacity x 4 floats */
    long long capacity;
    int *d_index;              /* optional: map pixel i*W+j of each edgel (no reference
                                  counterpart) */
    float *d_to_edgemap;       /* optional H rows of `tpitch` floats (TO_edgemap) */
    float *d_to_orientation;   /* optional, same layout (TO_orientation) */
    int tpitch;
} toed_subpix_map_t;
/* sigma -> filter radius ceil(4 sigma) (:107); TOED_ERR_SIGMA above 12 */
int toed_subpix_radius(double sigma, int *R);
/* nbr_radius: drop edgels with no other edgel within this distance (:558-621,
 * reference default 2; <= 0 keeps all; at most 9). */
int toed_subpix_plan(toed_ctx *ctx, int n, const toed_subpix_map_t *maps, double sigma, double nbr_radius,
                     toed_subpix **out);
/* The constructor's n = interp_n (subpix_to_correction.cpp:92-105): each map
 * (and its given orientation map) is resized to the work grid
 * ((H - 1) 2^n + 1) x ((W - 1) 2^n + 1) by bilinear interpolation with the
 * corners aligned (:69-90, set_edgemap / set_orientation :226-240; FP64, stored
 * as float) at every run; the path runs there with the NMS margin 3 * 2^n
 * (:627-629); x and y of the list are work-grid positions / 2^n (:694-695) and
 * the isolation radius applies to them; TO_edgemap / TO_orientation stay
 * H x W.  capacity counts rows (at most one per work pixel); d_index holds
 * work-grid pixels i * W' + j.  interp_n at most TOED_SUBPIX_MAX_INTERP, and
 * nbr_radius * 2^n at most 9 (the isolation reach on the work grid);
 * toed_subpix_set_grad refuses plans with interp_n > 0.  toed_subpix_plan is
 * this entry point with interp_n = 0. */
#define TOED_SUBPIX_MAX_INTERP 4
int toed_subpix_plan_n(toed_ctx *ctx, int n, const toed_subpix_map_t *maps, double sigma, double nbr_radius,
                       int interp_n, toed_subpix **out);
/* Enqueue on the context's stream; per-map counts and TOED_FLAG_* bits land in
 * d_counts[n] / d_flags[n] (device, either may be NULL). */
int toed_subpix_run(toed_subpix *p, int *d_counts, int *d_flags);
int toed_subpix_destroy(toed_subpix *p);
/* Point the plan's maps at other device buffers of the same sizes and pitches
 * (d_maps[m]: map m) for the next runs.  No reference counterpart. */
int toed_subpix_set_maps(toed_subpix *p, const float *const *d_maps);
/* Per-stage timing of later runs with HIP events (the reference's time_nms,
 * time_conv + time_correct, subpix_to_correction.hpp:76-78), in ms:
 * orientation + NMS kernel; isolation and convolution + correction kernels
 * (the latter writes the list rows); rank scans + TO maps. */
int toed_subpix_set_timing(toed_subpix *p, int on);
int toed_subpix_last_times(toed_subpix *p, float *ms_orient_nms, float *ms_correct, float *ms_list);

/* Differentiable form of the correction (no reference counterpart).
 * toed_subpix_set_grad(p, 1) makes later runs keep what the adjoint needs
 * (per pixel: direction and the derivatives of the estimated angle; per edgel:
 * its direction; maps at least 5 x 5).  toed_subpix_backward then takes, per
 * map, the upstream rows dL/d[x, y, orientation, strength] (d_edges layout)
 * and accumulates (+=) dL/dmap into d_gmap, with every discrete decision of
 * the last run frozen: the edgel set, each edgel's NMS octant and cubic
 * stencil, and per pixel the half-turn of the direction and sign(-Oxy).  For
 * maps with a given orientation, the orientation is a constant input.
 * Valid until the plan runs again. */
typedef struct {
    const float *d_gout;       /* capacity x 4 floats (rows of the map's d_edges) */
    float *d_gmap;             /* H rows of gpitch floats, accumulated */
    int gpitch;
} toed_subpix_grad_t;
int toed_subpix_set_grad(toed_subpix *p, int on);
int toed_subpix_backward(toed_subpix *p, const toed_subpix_grad_t *items);

/* Reference launch layer (unfused, for drop-in and debugging) ---------------
 * The reference's gpu_convolve / gpu_nms overloads (gpu_kernels.hpp:24-63)
 * with their argument lists unchanged; only the name carries the precision
 * suffix (a C ABI cannot overload) and the result is a status code (the
 * reference returns void).  As in the reference, the window is fixed at 17 / 19
 * taps (R0 = 8, R1 = 9) and the taps come in as eight device arrays of 19
 * values: the first, second and third derivative of the Gaussian and the
 * Gaussian at x = k - 9, then the same four at x = k - 8.5 (gpu_toed.hpp:
 * 368-386); NMS uses the reference's threshold 2 and border 10
 * (gpu_nms.cu:97,208).  Launches go to the null stream, like the reference's.
 * toed_gpu_convolve_f32: gpu_kernels.hpp:24-33.  Writes the four interp planes
 * Ix, Iy, |grad I| and TO orientation ((2h) x (2w), row-major) from a float
 * image.  toed_gpu_nms_f32: gpu_kernels.hpp:48-54.  Writes the subpixel
 * position maps (zero where no edgel, reference convention) from the three
 * planes.  The _f64 pair (gpu_kernels.hpp:35-45, :56-63) computes wholly in
 * FP64: image, taps, sums, orientation and the NMS decision. */
int toed_gpu_convolve_f32(int device_id, int h, int w, int interp_h, int interp_w,
                          const float *dev_img, float *dev_Ix, float *dev_Iy,
                          float *dev_I_grad_mag, float *dev_I_orient,
                          const float *dev_Gx, const float *dev_Gxx,
                          const float *dev_Gxxx, const float *dev_G_of_x,
                          const float *dev_Gx_sh, const float *dev_Gxx_sh,
                          const float *dev_Gxxx_sh, const float *dev_G_of_x_sh);
int toed_gpu_convolve_f64(int device_id, int h, int w, int interp_h, int interp_w,
                          const double *dev_img, double *dev_Ix, double *dev_Iy,
                          double *dev_I_grad_mag, double *dev_I_orient,
                          const double *dev_Gx, const double *dev_Gxx,
                          const double *dev_Gxxx, const double *dev_G_of_x,
                          const double *dev_Gx_sh, const double *dev_Gxx_sh,
                          const double *dev_Gxxx_sh, const double *dev_G_of_x_sh);
int toed_gpu_nms_f32(int device_id, int interp_h, int interp_w,
                     const float *dev_Ix, const float *dev_Iy, const float *dev_I_grad_mag,
                     float *dev_subpix_pos_x_map, float *dev_subpix_pos_y_map);
int toed_gpu_nms_f64(int device_id, int interp_h, int interp_w,
                     const double *dev_Ix, const double *dev_Iy, const double *dev_I_grad_mag,
                     double *dev_subpix_pos_x_map, double *dev_subpix_pos_y_map);
/* The same layer for any supported sigma (no reference counterpart: the
 * reference's taps are literals for sigma = 2): taps from sigma (windows
 * R0 = ceil(4 sigma), R1 = R0 + 1), and NMS with a given threshold and border. */
int toed_gpu_convolve_sigma_f32(int device_id, int h, int w, int interp_h, int interp_w,
                                const float *dev_img, float *dev_Ix, float *dev_Iy,
                                float *dev_I_grad_mag, float *dev_I_orient, float sigma);
int toed_gpu_convolve_sigma_f64(int device_id, int h, int w, int interp_h, int interp_w,
                                const double *dev_img, double *dev_Ix, double *dev_Iy,
                                double *dev_I_grad_mag, double *dev_I_orient, double sigma);
int toed_gpu_nms_thresh_f32(int device_id, int interp_h, int interp_w,
                            const float *dev_Ix, const float *dev_Iy, const float *dev_I_grad_mag,
                            float *dev_subpix_pos_x_map, float *dev_subpix_pos_y_map,
                            float thresh, int border);
int toed_gpu_nms_thresh_f64(int device_id, int interp_h, int interp_w,
                            const double *dev_Ix, const double *dev_Iy, const double *dev_I_grad_mag,
                            double *dev_subpix_pos_x_map, double *dev_subpix_pos_y_map,
                            double thresh, int border);

/* Curvelet formation (host code) ---------------------------------------------
 * The consumer the reference runs on its FP64 edgel list (main_gpu_cpu.cpp:
 * 211-281): curvelet/form_curvelet_process.hpp:23-82 form_curvelet_process.
 *   toed_curvelet_create      the constructor (:23-31, :101-113): edgeinfo is
 *                             the list in column major, n_edgels rows of
 *                             [x, y, orientation, strength] (Array.hpp layout,
 *                             as main_gpu_cpu.cpp:231-236 hands it over);
 *                             rad, gap, dx, dt (radians), token_len, max_k,
 *                             max_size_to_group and the centered /
 *                             bidirectional grouping flags as there.  The
 *                             list is copied.
 *   toed_curvelet_execute     execute() (:120-163): grouping, link graph,
 *                             curve fragments, poly-arcs.  nthreads = 0 uses
 *                             every hardware thread; results do not depend on it.
 *   toed_curvelet_output_size get_output_size (:437-457); info_w is the info
 *                             width the reference's driver allocates (10, 1, 12).
 *   toed_curvelet_output      get_output_arrary (:459-485): chain (h x w ints)
 *                             and info (h x info_w doubles) in column major;
 *                             either may be NULL.  Cells the reference leaves
 *                             unwritten are 0.  Type 2 (poly-arcs) is written
 *                             in full; the reference's own build stops at the
 *                             assert in curvelet::set_output for that type.
 * Results are bit-identical to the reference's on the same list. */
typedef struct toed_curvelets toed_curvelets;
int toed_curvelet_create(const double *edgeinfo, unsigned n_edgels, unsigned nrows, unsigned ncols, double rad,
                         double gap, double dx, double dt, double token_len, double max_k,
                         unsigned max_size_to_group, int centered, int bidirectional, toed_curvelets **out);
int toed_curvelet_execute(toed_curvelets *h, int nthreads);
int toed_curvelet_output_size(const toed_curvelets *h, unsigned output_type, unsigned *out_h, unsigned *out_w,
                              unsigned *info_w);
int toed_curvelet_output(const toed_curvelets *h, unsigned output_type, int *chain, double *info);
int toed_curvelet_destroy(toed_curvelets *h);

#ifdef __cplusplus
}
#endif
#endif /* TOED_CAPI_H */
