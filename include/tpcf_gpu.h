/*
 * tpcf_gpu.h -- C-ABI of the MI355X (gfx950) pair-counting library.
 *
 * Drop-in for the correlation call of tilmantroester/2pcf:
 *   call shear_2pcf(field, bins, results, spherical_coords, num_threads, use_kd_tree, verbose)
 *   call kappa_2pcf(field, bins, results, spherical_coords, num_threads, use_kd_tree, verbose)
 * (src/tpcf.f90:134-160 and :162-188), whose work is done by
 * tangential_shear_tpcf (src/tpcf.f90:190-378) and scalar_kappa_tpcf
 * (src/tpcf.f90:380-545).  The reference has no FFI; the Fortran module
 * tpcf_gpu (2pcf_amd/fortran/tpcf_gpu.f90) binds these entry points with
 * ISO_C_BINDING and keeps the reference's subroutine signatures.
 *
 * Conventions
 *  - Arrays are caller-owned, contiguous, 0-based.  Two-dimensional outputs
 *    use Fortran's column-major order: marks are [n_A][n_bin] with the bin
 *    index fastest (the reference's real(n_bin, n_A), src/tpcf.f90:204-205),
 *    per-patch-pair sums are [P_B][P_A][n_bin] (real(n_bin, P_A, P_B)).
 *  - Every entry point is synchronous: it returns after the outputs are
 *    written.  A context must not be used by two host threads at once.
 *  - Status codes replace the reference's print/stop error handling; the
 *    library never exits the process.
 *  - Outputs are overwritten, never accumulated (the reference zeroes its
 *    intent(out) arrays, src/tpcf.f90:242-252).
 */
#ifndef TPCF_GPU_H
#define TPCF_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Largest n_bin (the reference takes any --n-bin, src/main.f90:354-357).  One
 * pass of the pair kernel handles up to 63 consecutive bins (one squared edge
 * per wavefront lane); a call with more bins runs one pass per group of 63.
 * tpcf_adjoint_* take at most 63 bins. */
#define TPCF_MAX_BIN 1024

/* Bin spacing enum of src/utils.f90:7. */
enum tpcf_spacing { TPCF_LIN = 0, TPCF_LOG = 1, TPCF_IHS = 2, TPCF_SQRT = 3 };

enum tpcf_status {
    TPCF_OK = 0,
    TPCF_ERR_ARG = 1,          /* bad size, NULL required pointer, bad flag */
    TPCF_ERR_BINS = 2,         /* ill-defined binning (src/utils.f90:161-206 stops) */
    TPCF_ERR_ALLOC = 3,        /* device allocation failed */
    TPCF_ERR_HIP = 4,          /* a HIP runtime call failed */
    TPCF_ERR_UNSUPPORTED = 5,  /* option this build does not implement */
    TPCF_ERR_NONFINITE = 6,    /* a coordinate is NaN or Inf */
    TPCF_ERR_NODEVICE = 7,     /* no usable GPU */
    TPCF_ERR_PATCH = 8,        /* patch label outside [0, n_patch) */
    TPCF_ERR_GRID = 9,         /* catalogue extent too large for the cell grid */
    TPCF_ERR_MASK = 10         /* tpcf_batch_run: new values mask an object the batch was planned with
                                  (a kept lens now has w == 0 or z < 0, a kept source z < 0): plan a new batch */
};

/* Flags of the correlation calls. */
enum tpcf_flags {
    TPCF_DEVICE_INPUTS = 1u << 0,  /* input arrays (and patch labels) are device pointers */
    TPCF_SPHERICAL = 1u << 1,      /* spherical_coords = .true.: x, y are longitude and latitude and the bin
                                      edges angles, all in radians (src/tpcf.f90:313-320, :337-345); computed
                                      in FP64 whatever the entry point's precision.  Also taken by
                                      tpcf_adjoint_* and tpcf_adjoint_record_* (the shear angle of the
                                      source-side pass stays the one at the lens). */
    TPCF_VERBOSE = 1u << 2,        /* verbose = .true.: print a summary to stdout */
    TPCF_NO_KD_TREE = 1u << 3,     /* use_kd_tree = .false.: accepted; the pair set is identical */
    TPCF_BATCH_AUTOTUNE = 1u << 4, /* batch: time candidate cell / work-item sizes first */
    TPCF_BATCH_NO_GRAPH = 1u << 5, /* batch: run eagerly instead of capturing a HIP graph */
    TPCF_DEVICE_MARKS = 1u << 6    /* tpcf_shear_* / tpcf_kappa_*: the marks outputs (xi_E_marks, xi_B_marks,
                                      K_marks, w_marks, marks_weight) are device pointers, written in place with
                                      no copy to the host (the marks of a large catalogue are 4 n_bin n_A reals:
                                      a consumer on the GPU does not pay the PCIe transfer) */
};

/* Catalogue kinds. */
enum tpcf_kind { TPCF_SHEAR = 0, TPCF_KAPPA = 1 };

/* type(binning), src/utils.f90:18-24, plus the FP64 edges used for binning. */
typedef struct tpcf_bins {
    float x_min, x_max, bin_size;
    int32_t n_bin;
    int32_t spacing;               /* enum tpcf_spacing */
    float ihs_theta;
    double edges[TPCF_MAX_BIN + 1];    /* edges[0] = x_min, edges[n_bin] = x_max */
    float bin_centers[TPCF_MAX_BIN];   /* arithmetic midpoints, src/utils.f90:229 */
} tpcf_bins;

/* Outputs.  Each pointer may be NULL (that output is skipped).  "real"
 * outputs have real_bytes bytes per element: 4 for default real, 8 for a
 * -fdefault-real-8 build.  Pointers are host pointers. */
typedef struct tpcf_out {
    int32_t real_bytes;
    /* type(shear_2pcf_results), src/tpcf.f90:17-22, per bin [n_bin] */
    void *theta;          /* bin centres (src/tpcf.f90:151 / :179) */
    void *mean_theta;     /* Theta/W, NaN for an empty bin (src/tpcf.f90:369) */
    void *g_t;            /* G_t/K where K>0 else 0 (kappa: xi_E) */
    void *g_x;            /* G_x/K where K>0 else 0 (kappa: xi_B) */
    void *weights;        /* W */
    int64_t *n_pairs;
    /* raw FP64 sums per bin [n_bin]: Theta, G_t, G_x, W and K (K is internal
     * in the reference, exported here for stacking and bootstrap) */
    double *sum_theta, *sum_gt, *sum_gx, *sum_w, *sum_k;
    /* per-lens marks [n_A][n_bin] and marks_weight [n_A] (src/tpcf.f90:354-358) */
    void *xi_E_marks, *xi_B_marks, *K_marks, *w_marks, *marks_weight;
    /* per-patch-pair raw sums [P_B][P_A][n_bin]; summed over P_B they are the
     * per-block sums of src/tpcf.f90:1130-1137 */
    int64_t *pp_n;
    double *pp_theta, *pp_gt, *pp_gx, *pp_w, *pp_k;
} tpcf_out;

/* Timings and counters of the last correlation call on a context. */
typedef struct tpcf_stats {
    double ms_total;        /* whole call, host wall clock */
    double ms_upload;       /* H2D copies (0 with TPCF_DEVICE_INPUTS) */
    double ms_prepare;      /* bbox, sort into cells, segment and chunk tables (HIP events) */
    double ms_pairs;        /* the pair kernel (HIP events on the library stream) */
    double ms_finish;       /* reduction, normalisation and D2H */
    int64_t n_pairs;        /* in-bin pairs counted */
    int64_t n_lens_used;    /* lenses with w != 0 (and z >= 0 for shear) */
    int64_t n_source_used;  /* sources kept (z >= 0 for shear) */
    int64_t n_chunks;       /* lens chunks of up to 64 lenses */
    int64_t n_segments;     /* non-empty (cell, patch) source segments */
    double cell_size;       /* source cell side used */
    int64_t seg_visits[3];  /* fast / edge / multi segment visits (only with TPCF_OPT_COUNTERS) */
    int64_t pair_visits[3]; /* lens-source pairs examined per class (same) */
    int64_t n_items;        /* pair-kernel work items (chunks, heavy ones split) */
} tpcf_stats;

/* Options for tpcf_set_option. */
enum tpcf_option {
    TPCF_OPT_CELL_SIZE = 1,     /* > 0: source cell side; 0: automatic (default) */
    TPCF_OPT_COUNTERS = 2,      /* != 0: collect seg_visits / pair_visits (slower) */
    TPCF_OPT_CELL_TARGET = 3,   /* automatic cell size: mean sources per cell (default: 128 for
                                   FP32 geometry, 64 for FP64) */
    TPCF_OPT_SPLIT = 4          /* pair-kernel work-item band height in units of x_max (default 0.2) */
};

typedef struct tpcf_ctx tpcf_ctx;

/* create_bins(x_min=, x_max=, n_bin=, p=, spacing=), src/utils.f90:139-230. */
int tpcf_create_bins(float x_min, float x_max, int32_t n_bin, int32_t spacing, float ihs_theta,
                     tpcf_bins *bins);

/* find_bin_index (src/utils.f90:101-137) under the library's bin definition:
 * 0-based bin k with edges[k]^2 <= x^2 < edges[k+1]^2 evaluated in FP64,
 * x == x_max -> n_bin-1, -1 outside [x_min, x_max]. */
int32_t tpcf_find_bin_index(double x, const tpcf_bins *bins);

/* Context: device, stream and reusable device buffers.  device < 0 picks the
 * current device. */
int tpcf_ctx_create(int32_t device, tpcf_ctx **ctx);
int tpcf_ctx_destroy(tpcf_ctx *ctx);
int tpcf_set_option(tpcf_ctx *ctx, int32_t option, double value);
/* Use a caller-owned hipStream_t (NULL restores the library's own stream). */
int tpcf_set_stream(tpcf_ctx *ctx, void *hip_stream);
int tpcf_get_stats(const tpcf_ctx *ctx, tpcf_stats *stats);
const char *tpcf_last_error(const tpcf_ctx *ctx);

/* shear_2pcf / tangential_shear_tpcf (src/tpcf.f90:134-160, :190-378), FP32
 * inputs: lenses x,y,z,val,w [n_A]; sources x,y,z,e1,e2,w,m [n_B].  Geometry
 * and per-pair products in FP32, pairs within 1e-5 (relative, on d^2) of a bin
 * edge re-binned in FP64, sums in FP64.  patch_A / patch_B (nullable) are
 * 0-based labels in [0, n_patch_A) / [0, n_patch_B). */
int tpcf_shear_f32(tpcf_ctx *ctx,
                   int64_t n_A, const float *x_A, const float *y_A, const float *z_A,
                   const float *val_A, const float *w_A,
                   int64_t n_B, const float *x_B, const float *y_B, const float *z_B,
                   const float *e1_B, const float *e2_B, const float *w_B, const float *m_B,
                   const tpcf_bins *bins, uint32_t flags,
                   const int32_t *patch_A, const int32_t *patch_B,
                   int32_t n_patch_A, int32_t n_patch_B, tpcf_out *out);

/* kappa_2pcf / scalar_kappa_tpcf (src/tpcf.f90:162-188, :380-545), FP32
 * inputs: A x,y,val,w [n_A]; B x,y,kappa_E,kappa_B,w [n_B]. */
int tpcf_kappa_f32(tpcf_ctx *ctx,
                   int64_t n_A, const float *x_A, const float *y_A, const float *val_A, const float *w_A,
                   int64_t n_B, const float *x_B, const float *y_B, const float *kappa_E,
                   const float *kappa_B, const float *w_B,
                   const tpcf_bins *bins, uint32_t flags,
                   const int32_t *patch_A, const int32_t *patch_B,
                   int32_t n_patch_A, int32_t n_patch_B, tpcf_out *out);

/* Allocation-free variant of tpcf_shear_f32 (no counterpart in the reference):
 * the same arguments, outputs and kernels, for callers that cannot allocate
 * device memory (serving loops that must not call hipMalloc, harnesses that
 * forbid it).  It calls no hipMalloc, hipFree or other allocating device
 * routine: every device buffer of the call is carved from a 256 MiB scratch
 * array that each GPU holds from the library's load until the process ends.
 * A call that needs more (about 2M x 2M objects with device inputs and 16 x 16
 * patches fit) returns TPCF_ERR_ALLOC with a message naming the limit.  All
 * calls on one GPU share its scratch array: overlapping calls on one GPU race
 * and must be serialised by the caller. */
int tpcf_shear_f32_noalloc(tpcf_ctx *ctx,
                           int64_t n_A, const float *x_A, const float *y_A, const float *z_A,
                           const float *val_A, const float *w_A,
                           int64_t n_B, const float *x_B, const float *y_B, const float *z_B,
                           const float *e1_B, const float *e2_B, const float *w_B, const float *m_B,
                           const tpcf_bins *bins, uint32_t flags,
                           const int32_t *patch_A, const int32_t *patch_B,
                           int32_t n_patch_A, int32_t n_patch_B, tpcf_out *out);

/* FP64 twins (the reference built with -fdefault-real-8): FP64 geometry,
 * binning and sums throughout. */
int tpcf_shear_f64(tpcf_ctx *ctx,
                   int64_t n_A, const double *x_A, const double *y_A, const double *z_A,
                   const double *val_A, const double *w_A,
                   int64_t n_B, const double *x_B, const double *y_B, const double *z_B,
                   const double *e1_B, const double *e2_B, const double *w_B, const double *m_B,
                   const tpcf_bins *bins, uint32_t flags,
                   const int32_t *patch_A, const int32_t *patch_B,
                   int32_t n_patch_A, int32_t n_patch_B, tpcf_out *out);

int tpcf_kappa_f64(tpcf_ctx *ctx,
                   int64_t n_A, const double *x_A, const double *y_A, const double *val_A, const double *w_A,
                   int64_t n_B, const double *x_B, const double *y_B, const double *kappa_E,
                   const double *kappa_B, const double *w_B,
                   const tpcf_bins *bins, uint32_t flags,
                   const int32_t *patch_A, const int32_t *patch_B,
                   int32_t n_patch_A, int32_t n_patch_B, tpcf_out *out);

/* Reverse mode (no counterpart in the reference, which has no derivative):
 * given the gradient of a scalar loss with respect to the per-bin outputs
 * g_t, g_x of tpcf_shear_* (gbar_t, gbar_x: host arrays [n_bin]), the
 * gradient with respect to each lens' val and w [n_A] and each source's e1,
 * e2, w and m [n_B].  Inputs are those of the forward call; with
 * TPCF_DEVICE_INPUTS the inputs and the gradient arrays are device pointers.
 * Any gradient pointer may be NULL.  Lenses with w == 0 (skipped by the
 * reference, src/tpcf.f90:285) receive the gradient of the smooth extension;
 * objects masked by z < 0 receive zero.  Gradients are FP64.  Flags:
 * TPCF_DEVICE_INPUTS, TPCF_SPHERICAL, TPCF_VERBOSE, TPCF_NO_KD_TREE.  On the
 * sphere the source-side pass takes the position angle at the lens as the
 * forward does, from the lens' direction scaled by cos(lat_A), so a source's
 * gradient terms from a lens within ~1e-8 rad of a pole lose precision (at
 * the pole itself they are 0). */
int tpcf_adjoint_shear_f32(tpcf_ctx *ctx,
                           int64_t n_A, const float *x_A, const float *y_A, const float *z_A,
                           const float *val_A, const float *w_A,
                           int64_t n_B, const float *x_B, const float *y_B, const float *z_B,
                           const float *e1_B, const float *e2_B, const float *w_B, const float *m_B,
                           const tpcf_bins *bins, uint32_t flags, const double *gbar_t, const double *gbar_x,
                           double *val_A_bar, double *w_A_bar,
                           double *e1_B_bar, double *e2_B_bar, double *w_B_bar, double *m_B_bar);
int tpcf_adjoint_shear_f64(tpcf_ctx *ctx,
                           int64_t n_A, const double *x_A, const double *y_A, const double *z_A,
                           const double *val_A, const double *w_A,
                           int64_t n_B, const double *x_B, const double *y_B, const double *z_B,
                           const double *e1_B, const double *e2_B, const double *w_B, const double *m_B,
                           const tpcf_bins *bins, uint32_t flags, const double *gbar_t, const double *gbar_x,
                           double *val_A_bar, double *w_A_bar,
                           double *e1_B_bar, double *e2_B_bar, double *w_B_bar, double *m_B_bar);
/* The same for tpcf_kappa_*: gradients of xi_E (gbar_E) and xi_B (gbar_B)
 * with respect to val, w of A and kappa_E, kappa_B, w of B. */
int tpcf_adjoint_kappa_f32(tpcf_ctx *ctx,
                           int64_t n_A, const float *x_A, const float *y_A, const float *val_A, const float *w_A,
                           int64_t n_B, const float *x_B, const float *y_B, const float *kappa_E,
                           const float *kappa_B, const float *w_B,
                           const tpcf_bins *bins, uint32_t flags, const double *gbar_E, const double *gbar_B,
                           double *val_A_bar, double *w_A_bar,
                           double *kE_B_bar, double *kB_B_bar, double *w_B_bar);
int tpcf_adjoint_kappa_f64(tpcf_ctx *ctx,
                           int64_t n_A, const double *x_A, const double *y_A, const double *val_A,
                           const double *w_A,
                           int64_t n_B, const double *x_B, const double *y_B, const double *kappa_E,
                           const double *kappa_B, const double *w_B,
                           const tpcf_bins *bins, uint32_t flags, const double *gbar_E, const double *gbar_B,
                           double *val_A_bar, double *w_A_bar,
                           double *kE_B_bar, double *kB_B_bar, double *w_B_bar);

/* Recorded adjoint (no counterpart in the reference): the adjoint above runs
 * two pair passes, a lens-side pass whose per-lens sums also give the forward
 * outputs, and a source-side pass.  tpcf_adjoint_record_* runs the lens-side
 * pass alone, returns the forward outputs (g_t, g_x or xi_E, xi_B per bin,
 * host arrays [n_bin], nullable; 0 where K or W is 0) and keeps that pass's
 * per-lens sums (4 n_bin doubles per used lens, device memory) in *tape.
 * tpcf_adjoint_replay then gives the gradient for any upstream gradient with
 * the source-side pass only, as often as wanted; an autodiff forward records,
 * its backward replays, so forward + backward is two passes instead of three.
 * Inputs must be device pointers (TPCF_DEVICE_INPUTS; the tape keeps the
 * pointers, which must stay valid and unchanged until the last replay);
 * gradient outputs of the replay are device pointers, any may be NULL
 * (s1/s2 = e1/e2 for shear, kappa_E/kappa_B for kappa; m_B_bar is ignored for
 * kappa).  A tape belongs to the device it was recorded on and is freed by
 * tpcf_tape_destroy. */
typedef struct tpcf_tape tpcf_tape;
int tpcf_adjoint_record_shear_f32(tpcf_ctx *ctx,
                                  int64_t n_A, const float *x_A, const float *y_A, const float *z_A,
                                  const float *val_A, const float *w_A,
                                  int64_t n_B, const float *x_B, const float *y_B, const float *z_B,
                                  const float *e1_B, const float *e2_B, const float *w_B, const float *m_B,
                                  const tpcf_bins *bins, uint32_t flags, double *g_t, double *g_x,
                                  tpcf_tape **tape);
int tpcf_adjoint_record_shear_f64(tpcf_ctx *ctx,
                                  int64_t n_A, const double *x_A, const double *y_A, const double *z_A,
                                  const double *val_A, const double *w_A,
                                  int64_t n_B, const double *x_B, const double *y_B, const double *z_B,
                                  const double *e1_B, const double *e2_B, const double *w_B, const double *m_B,
                                  const tpcf_bins *bins, uint32_t flags, double *g_t, double *g_x,
                                  tpcf_tape **tape);
int tpcf_adjoint_record_kappa_f32(tpcf_ctx *ctx,
                                  int64_t n_A, const float *x_A, const float *y_A, const float *val_A,
                                  const float *w_A,
                                  int64_t n_B, const float *x_B, const float *y_B, const float *kappa_E,
                                  const float *kappa_B, const float *w_B,
                                  const tpcf_bins *bins, uint32_t flags, double *xi_E, double *xi_B,
                                  tpcf_tape **tape);
int tpcf_adjoint_record_kappa_f64(tpcf_ctx *ctx,
                                  int64_t n_A, const double *x_A, const double *y_A, const double *val_A,
                                  const double *w_A,
                                  int64_t n_B, const double *x_B, const double *y_B, const double *kappa_E,
                                  const double *kappa_B, const double *w_B,
                                  const tpcf_bins *bins, uint32_t flags, double *xi_E, double *xi_B,
                                  tpcf_tape **tape);
int tpcf_adjoint_replay(tpcf_ctx *ctx, const tpcf_tape *tape, const double *gbar_1, const double *gbar_2,
                        double *val_A_bar, double *w_A_bar,
                        double *s1_B_bar, double *s2_B_bar, double *w_B_bar, double *m_B_bar);
int tpcf_tape_destroy(tpcf_tape *tape);

/* ---- many catalogues (mock fields) as one HIP graph ----
 * No counterpart in the reference beyond its field loop (src/main.f90:596-660,
 * one shear_2pcf / kappa_2pcf call per field, results stacked as in
 * src/main.f90:678-695).  tpcf_batch_create_* prepares every catalogue once
 * (cells, segments, chunks: positions and masks w_A == 0, z < 0) and, with
 * TPCF_BATCH_AUTOTUNE, first times candidate cell sizes and work-item sizes on
 * a sample of up to 12 catalogues of each kind spread over the batch's sizes.
 * Each tpcf_batch_run re-reads the values (val, w,
 * z, s1, s2, m) from the same device arrays, runs every catalogue and writes
 * each catalogue's outputs (per-bin outputs, raw sums, per-patch-pair sums,
 * and the per-lens marks where out's marks pointers are not NULL: those are
 * device arrays of out->real_bytes per element, written in place each run as
 * with TPCF_DEVICE_MARKS; the reference writes every field's marks,
 * src/main.f90:704-707).  The first run captures the work into one HIP
 * graph (catalogues on parallel branches); later runs replay it.  Device
 * arrays only (TPCF_DEVICE_INPUTS); they must stay valid while the batch
 * exists.  m_B may be NULL (m == 0).  With TPCF_SPHERICAL every catalogue is
 * on the sphere (x, y longitude and latitude, bins angles, in radians; FP64
 * geometry as for a single call).  A run whose new values mask an object the
 * batch was planned with returns TPCF_ERR_MASK.  The context must stay valid
 * while the batch is created and run; tpcf_batch_destroy does not use it. */
typedef struct tpcf_catalog {
    int32_t kind;                  /* enum tpcf_kind */
    int32_t n_patch_A, n_patch_B;
    int64_t n_A, n_B;
    const void *x_A, *y_A, *z_A, *val_A, *w_A;          /* z_A: shear only */
    const void *x_B, *y_B, *z_B, *s1_B, *s2_B, *w_B, *m_B;   /* s1, s2: e1, e2 or kappa_E, kappa_B */
    const int32_t *patch_A, *patch_B;                   /* nullable */
    tpcf_out *out;                 /* outputs of this catalogue: host arrays, except the marks (device arrays) */
} tpcf_catalog;

typedef struct tpcf_batch_stats {
    double ms_plan;            /* preparation, autotuning included */
    double ms_run;             /* last run (host wall clock, graph launch to outputs written) */
    double cell_target, split; /* the settings used (TPCF_OPT_CELL_TARGET, TPCF_OPT_SPLIT) */
    double tune_ms[5];         /* autotuning: time of each candidate on the sample */
    int64_t n_pairs;           /* last run, all catalogues */
    int64_t n_items;           /* pair-kernel work items, all catalogues */
    int64_t graph_nodes;       /* nodes of the captured graph (0 before the first run or without a graph) */
    int64_t runs;
    int32_t n_catalogs, autotuned;
} tpcf_batch_stats;

typedef struct tpcf_batch tpcf_batch;

int tpcf_batch_create_f32(tpcf_ctx *ctx, int32_t n_cat, const tpcf_catalog *cats, const tpcf_bins *bins,
                          uint32_t flags, tpcf_batch **batch);
int tpcf_batch_create_f64(tpcf_ctx *ctx, int32_t n_cat, const tpcf_catalog *cats, const tpcf_bins *bins,
                          uint32_t flags, tpcf_batch **batch);
int tpcf_batch_run(tpcf_batch *batch);
int tpcf_batch_get_stats(const tpcf_batch *batch, tpcf_batch_stats *stats);
int tpcf_batch_destroy(tpcf_batch *batch);

/* ---- block bootstrap and covariance ------------------------------------------
 * Replaces the resampling loop of survey_marked_point_block_bootstrap
 * (src/tpcf.f90:1165-1191; survey_marked_point_bootstrap :1230-1262 is the
 * same with one object per block) and covariance_matrix (src/utils.f90:745-764)
 * as the reference's main program applies it to the resampled g_t and g_x
 * (src/main.f90:524-537).
 *
 * Blocks are lens patches: g_t, g_x, k hold each block's per-bin sums of the
 * G_t, G_x and K marks over its lenses (the per-patch-pair sums of a
 * correlation call summed over source patches), weight its lenses' summed w.
 * Sample i with block multiplicities M[i][p] gives
 *   g_t[i][b] = sum_p M g_t[p][b] / sum_p M k[p][b]  (0 where that K <= 0),
 *   g_x likewise, weights[i] = sum_p M weight[p]
 * and cov_t[a][b] = sum_i (g_t[i][a] - mean_a)(g_t[i][b] - mean_b) / (n_boot - 1)
 * (cov_x likewise).  M comes from `multiplicity` ([n_boot][n_block] int32) or,
 * when it is NULL, from n_draw = int(n_block * supersampling) uniform draws of a
 * block per sample, from a counter-based generator keyed by `seed` (written to
 * `multiplicity_out` when that is not NULL).  With TPCF_DEVICE_INPUTS every
 * array argument is a device address, else a host address.  FP64 throughout;
 * n_bin <= TPCF_MAX_BIN. */
typedef struct tpcf_blocks {
    int32_t n_block, n_bin;
    const double *g_t, *g_x, *k;   /* [n_block][n_bin] */
    const double *weight;          /* [n_block], nullable */
} tpcf_blocks;

int tpcf_bootstrap(tpcf_ctx *ctx, const tpcf_blocks *blocks, int32_t n_boot, const int32_t *multiplicity,
                   uint64_t seed, double supersampling, uint32_t flags, double *g_t, double *g_x,
                   double *weights, double *cov_t, double *cov_x, int32_t *multiplicity_out);

/* ---- k-means patch labels ----------------------------------------------------
 * Replaces partition_data (src/kmeans.f90:12-263), the clustering of the
 * voronoi_blocks and equal_weight_blocks schemes (src/tpcf.f90:1100-1113):
 * rounds of nearest-centroid assignment (first minimum kept) and
 * weighted-mean centroids until a round moves no point, moves fewer than
 * reassignment_tol * n, or max_iterations + 1 rounds have run.  init 0: the
 * fixed starting grid of kmeans_compute_fixed_initial_centroids
 * (src/kmeans.f90:85-110), labels int(i/n k); init 1: `centroids` holds the k
 * starting centroids.  labels are 0-based.  TPCF_SPHERICAL: longitude /
 * latitude in radians, great-circle distances.
 * weight_equalizer_scaling s > 0: from the second round the distance to
 * cluster j is scaled by 1 + (W_j / W_mean - 1) s, with the cluster weights
 * W_j moved after every step of 256 consecutive points (every point of a step
 * sees the weights from before it; the reference moves them after every
 * point).  The result does not depend on any thread order.  Such rounds run
 * in one workgroup.  equal_size 1: the runner-up clusters are
 * tracked from the second round and, when (max W - min W) / W_mean exceeds
 * weight_tol after the rounds, postprocess_weights (src/kmeans.f90:214-262)
 * hands points to their runner-up clusters (equal runner-up ratios are taken
 * in point order).  centroids are those of the last round; cluster_weights
 * those after the hand-overs.  2 <= k <= 1024; init, equal_size in {0, 1};
 * s finite and >= 0; FP64 arithmetic. */
typedef struct tpcf_kmeans_opts {
    int32_t k;
    int32_t max_iterations;            /* the reference's default: 1000 */
    int32_t init;                      /* 0: fixed grid, 1: given centroids */
    int32_t equal_size;                /* 1: runner-up tracking and postprocess_weights */
    double reassignment_tol;           /* stop when moves / n < tol (0: run to convergence) */
    double weight_equalizer_scaling;   /* s >= 0 (0: plain distances) */
    double weight_tol;                 /* equal_size: the post-processing's relative weight tolerance */
} tpcf_kmeans_opts;

int tpcf_kmeans_f32(tpcf_ctx *ctx, int64_t n, const float *x, const float *y, const float *w,
                    const tpcf_kmeans_opts *opts, uint32_t flags, int32_t *labels, double *centroids,
                    double *cluster_weights, int32_t *rounds, int64_t *last_moves);
int tpcf_kmeans_f64(tpcf_ctx *ctx, int64_t n, const double *x, const double *y, const double *w,
                    const tpcf_kmeans_opts *opts, uint32_t flags, int32_t *labels, double *centroids,
                    double *cluster_weights, int32_t *rounds, int64_t *last_moves);

/* ---- rectangular bootstrap blocks ------------------------------------------
 * Replaces the block assignment of the rectangular_blocks scheme of
 * survey_marked_point_block_bootstrap (src/tpcf.f90:1017-1049), the
 * reference's default bootstrap scheme (src/main.f90:100), for one field: the
 * bounding box of the field's n objects is cut into n_x_blocks x n_y_blocks
 * linear bins (create_bins / find_bin_index in single precision,
 * src/utils.f90:101-230), and object j gets the 0-based label
 * (iy - 1) * n_y_blocks + ix - 1 + offset, ix and iy its 1-based bins (the
 * reference's row stride is n_y_blocks).  offset: the labels of the fields
 * before this one (the reference adds n_x_blocks n_y_blocks per field).
 * Double coordinates are rounded to single precision first.  A label outside
 * [offset, offset + n_x_blocks n_y_blocks) (n_x_blocks < n_y_blocks, where the
 * reference writes past the field's blocks) returns TPCF_ERR_ARG.  Flags:
 * TPCF_DEVICE_INPUTS (x, y and labels are device arrays). */
int tpcf_rect_blocks_f32(tpcf_ctx *ctx, int64_t n, const float *x, const float *y, int32_t n_x_blocks,
                         int32_t n_y_blocks, int32_t offset, uint32_t flags, int32_t *labels);
int tpcf_rect_blocks_f64(tpcf_ctx *ctx, int64_t n, const double *x, const double *y, int32_t n_x_blocks,
                         int32_t n_y_blocks, int32_t offset, uint32_t flags, int32_t *labels);

#ifdef __cplusplus
}
#endif

#endif /* TPCF_GPU_H */
