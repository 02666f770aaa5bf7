/*
 * barbie.h -- C ABI of the MI355X bispectrum-multipole estimator library
 * (libbarbie_amd.so).
 *
 * The reference program (barbie.c) has no library interface: its stages talk
 * through globals and files.  The entry points below replace those stages one
 * for one; each declaration names the reference code it stands in for.
 * Plain C types only.  Every function returns 0 on success or a negative
 * BB_E* code, and bb_last_error() gives a message for the calling thread.
 * The library never calls exit().
 *
 * Threading: a bb_plan is immutable after creation and may be shared between
 * threads.  A bb_estimator owns device memory and one HIP stream; calls on
 * one estimator must be serialised by the caller.  Different estimators may
 * be used from different threads.
 */
#ifndef BARBIE_H
#define BARBIE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BB_NPOLES 16

enum {
    BB_OK = 0,
    BB_E_INVALID = -1,   /* invalid argument or geometry */
    BB_E_NOMEM = -2,     /* host allocation failed */
    BB_E_DEVICE = -3,    /* HIP error (no device, launch failure, ...) */
    BB_E_DEVMEM = -4,    /* device allocation failed / does not fit in HBM */
    BB_E_IO = -5,        /* file error */
    BB_E_RANGE = -6      /* a triangle falls outside the bin grid */
};

/* Ring-edge rules. */
enum {
    BB_RINGS_REFERENCE = 0, /* barbie.c:110-130: blocks of ten, width b+1 in block b,
                               -0.1 shift, last ring closed at Nmax; truncated at Nr */
    BB_RINGS_UNIFORM = 1,   /* same construction with a fixed width */
    BB_RINGS_EXPLICIT = 2   /* caller-supplied low/high edges */
};

/* Arithmetic of the contraction stage (the FFTs run in FP64 for BB_FP64 and
 * in FP32 for the others; the per-bin sums, triangle counts and Y_lm
 * projection are FP64 in every mode). */
typedef enum {
    BB_FP64 = 0,    /* FP64 fields, FP64 MFMA contraction (the FP64 GPU oracle) */
    BB_FP32 = 1,    /* FP32 fields, FP32 MFMA contraction, FP64 across LDS chunks */
    BB_BF16X3 = 2   /* FP32 fields, BF16x3 split products on MFMA (hi*hi + hi*lo +
                       lo*hi), FP32 within an LDS chunk, FP64 across chunks */
} bb_precision;

/* Execution options of an estimator (bb_estimator_create_opts). */
typedef struct bb_options {
    bb_precision precision;
    int full_grid;  /* 1: run the plane-ring transforms and triple sums on the
                       NN grid, as the reference does.  0: when NN >= 3*kmax+1
                       (kmax = largest |k_x| of any ring mode, so that no sum
                       of three ring modes wraps around on the NN grid), run
                       each contraction on a smaller grid instead: a
                       contraction of rings r1, r2, r3 sums exactly (times
                       (NN/NG)^2) on any grid NG > K1 + K2 + K3 (K_r = largest
                       |k_x| of ring r) that also holds each ring's modes
                       (NG >= 2 K_r + 2).  Below 3*kmax+1 the NN grid is used. */
    int x_splits;   /* partial sums per contraction over the points (0: automatic) */
    int grid_levels; /* with full_grid 0: the largest number of distinct grids
                       (0: automatic, 1: one grid).  The grids are chosen
                       among the even 2,3,5-smooth sizes to minimise the
                       summed points over all contractions. */
    int max_batch;  /* fields one launch sequence processes together (0 or 1:
                       one).  Device buffers per field are held max_batch
                       times; the work lists are shared.  With x_splits 0 the
                       contraction work of a field is cut into about
                       2048 / max_batch x-ranges, so a field's bins are the
                       same bits in every batch of one estimator, and agree
                       with those of an estimator of another max_batch to
                       rounding (FP64: ~1e-15; split paths: ~1e-7). */
    int waves;      /* waves per contraction workgroup: 4 or 8 (0: 8, or the
                       environment variable BARBIE_CONTRACT_WAVES) */
    int mix_fp32;   /* BF16X3 only: percent (0-90) of the contraction's cost given
                       to the FP32-MFMA kernel, which runs concurrently with the
                       BF16x3 kernel on a second stream (the two load the matrix
                       cores and the vector ALUs differently); 0: none */
    int use_graph;  /* 1: capture the launch sequence of a batch into a HIP
                       graph the first time it runs with a set of device
                       buffers, and replay the graph when the same buffers
                       come again (bb_estimate and bb_estimate_batch use the
                       estimator's own buffers, so they always replay) */
    int occupancy;  /* BF16X3 with 4 waves only: contraction workgroups per
                       compute unit the kernel is sized for, 2 (0: default) or
                       3 (6 instead of 8 tile windows per wave and at most 96
                       instead of 120 staged rings per work item, so that three
                       workgroups' registers and LDS fit on one CU) */
} bb_options;

/* Geometry and binning of one estimate.  Replaces the constants hard-coded at
 * barbie.c:58-84 and the ring construction at barbie.c:110-130. */
typedef struct bb_geometry {
    int NN;              /* grid size: even, factors 2,3,5 only */
    int Nmax;            /* kz planes 0..Nmax-1 (barbie.c:71) */
    int Nr;              /* number of rings (barbie.c:72) */
    int ring_rule;       /* BB_RINGS_* */
    double ring_width;   /* BB_RINGS_UNIFORM only */
    int last_ring_to_nmax; /* BB_RINGS_UNIFORM: 1 closes the last ring at Nmax */
    const double *ring_low;  /* BB_RINGS_EXPLICIT: Nr lower edges */
    const double *ring_high; /* BB_RINGS_EXPLICIT: Nr upper edges */
    int n_mtbin;         /* shape bins per axis (barbie.c:84) */
    int UL_k1, LL_k1, BW_k1; /* size bins (argv[3..5], barbie.c:63-65) */
    double LL;           /* cell size (barbie.c:67) */
    double ff, beta;     /* analytic-comparison constants (barbie.c:79-80) */
} bb_geometry;

typedef struct bb_plan_stats {
    int NN, Nmax, Nr, n_mtbin, n_k1bin;
    int rcount;          /* ring triples (calCR, funcs.c:389-408) */
    long long ncontr;    /* contractions visited by barbie.c:247-292 */
    long long npairs;    /* distinct (p1,r1,p2,r2) pair fields */
    int ngroups;         /* distinct (p1,p2) plane groups */
    int nfields_used;    /* distinct (plane, ring) fields read by some contraction */
    int nbins;           /* n_k1bin * n_mtbin * n_mtbin */
} bb_plan_stats;

typedef struct bb_plan bb_plan;
typedef struct bb_estimator bb_estimator;

const char *bb_last_error(void);
const char *bb_version(void);

/* Fills *g with the reference's geometry for the given CLI arguments:
 * Nmax=100, Nr=40, piecewise rings, n_mtbin=5, LL=1, ff=0.1, beta=1
 * (barbie.c:61-84). */
int bb_geometry_reference(int NN, int UL_k1, int LL_k1, int BW_k1, bb_geometry *g);

/* Host-side planning; no GPU is touched.  Builds the ring edges
 * (barbie.c:110-130), ring masks as integer radius bounds (barbie.c:135-148),
 * the ring triples (calCR's predicate, funcs.c:389-395) and the full list of
 * contractions in the reference's enumeration order (barbie.c:247-292) with
 * their bins (barbie.c:343-345).  Validates what the reference leaves
 * unchecked (docs/REVIEW.md:10): even NN, BW>0, divisible size range,
 * Nmax <= NN/2, rings inside the grid, every bin index in range. */
int bb_plan_create(const bb_geometry *g, bb_plan **out);
void bb_plan_destroy(bb_plan *p);
int bb_plan_stats_get(const bb_plan *p, bb_plan_stats *s);
/* k_perp_Low/High/avg of barbie.c:89-130 (Nr doubles each). */
int bb_plan_ring_edges(const bb_plan *p, double *low, double *high, double *avg);
/* Ring triples in calCR order (funcs.c:389-404): rcount * 3 ints. */
int bb_plan_triples(const bb_plan *p, int *r123);
/* Contractions in barbie.c:247-292 order: ncontr * 8 ints
 * {p1, r1, p2, r2, p3, r3, rp, flat_bin}. */
int bb_plan_contractions(const bb_plan *p, int *rec);

/* Device side.  Creates an estimator for a plan on the current HIP device:
 * allocates the plane-ring cache and work buffers in HBM and computes the
 * triangle counts Num (calCR, funcs.c:328-424) on the GPU in FP64. */
int bb_estimator_create(const bb_plan *p, bb_precision prec, bb_estimator **out);
int bb_estimator_create_opts(const bb_plan *p, const bb_options *opt, bb_estimator **out);
void bb_estimator_destroy(bb_estimator *e);
/* Grid levels an estimator with these options uses (host only, no GPU):
 * *nlevels grids in ascending size, and for each contraction (plan order)
 * the index of its level (clevel may be NULL). */
int bb_plan_grid_levels(const bb_plan *p, const bb_options *opt, int *nlevels, int *grids,
                        int *clevel);
/* Largest grid the estimator's transforms and sums run on (NG; NN with
 * full_grid). */
int bb_estimator_grid(const bb_estimator *e, int *NG);
/* Grid levels in ascending size: *nlevels and up to BB_MAX_LEVELS grids. */
#define BB_MAX_LEVELS 16
int bb_estimator_levels(const bb_estimator *e, int *nlevels, int *grids);
/* Autotuner (no counterpart in the reference): times an estimate for each
 * candidate layout -- waves per contraction workgroup (8, 4) x number of grid
 * levels (4, 5, 6, 8; BF16X3: 4, 6; those not above base->grid_levels when it
 * is set) x, for BF16X3, mix_fp32 (0, 30, 45) -- on the current device and
 * returns
 * in *best the fastest options (base with waves, grid_levels, mix_fp32 set) and its
 * time per estimate in *best_ms (may be NULL).  d_field: a device field to
 * time with, or NULL for zeros (the work does not depend on the values).
 * Each candidate runs once untimed and `repeats` times timed (best kept).
 * The bins of every candidate are the same to FP64 rounding. */
int bb_tune(const bb_plan *p, const bb_options *base, const double *d_field, int repeats,
            bb_options *best, double *best_ms);
/* Device bytes an estimator of this plan and precision holds (default
 * options); the layout it measures is cached in the plan and reused by
 * bb_estimator_create. */
int bb_estimator_footprint(const bb_plan *p, bb_precision prec, size_t *bytes);
int bb_estimator_footprint_opts(const bb_plan *p, const bb_options *opt, size_t *bytes);

/* One estimate (barbie.c:176-373): field_r2c is the FFTW in-place r2c buffer
 * [NN][NN][NN+2] doubles in host memory (the contents of
 * field_N<NN>_R<NRi>.txt).  Outputs are the un-normalised FP64 bin sums
 * bin_BS_real, bin_normalization and bin_Biana of barbie.c:87, each
 * [n_k1bin][n_mtbin][n_mtbin][16] (any may be NULL). */
int bb_estimate(bb_estimator *e, const double *field_r2c,
                double *bin_bs, double *bin_norm, double *bin_biana);
/* Same with the field already in device memory (pointer from hipMalloc) and
 * device output pointers; enqueued on the estimator's stream, which is
 * synchronised before return. */
int bb_estimate_device(bb_estimator *e, const double *d_field_r2c,
                       double *d_bin_bs, double *d_bin_norm, double *d_bin_biana);

/* Batches (no counterpart in the reference, which reads one field per
 * process): nfields <= max_batch fields back to back ([nfields][NN][NN][NN+2]
 * doubles), outputs back to back per field ([nfields][n_k1bin][n_mtbin]
 * [n_mtbin][16] each).  All fields of a batch share every launch. */
int bb_estimate_batch(bb_estimator *e, int nfields, const double *fields_r2c,
                      double *bin_bs, double *bin_norm, double *bin_biana);
int bb_estimate_batch_device(bb_estimator *e, int nfields, const double *d_fields_r2c,
                             double *d_bin_bs, double *d_bin_norm, double *d_bin_biana);
/* Asynchronous form (device pointers): enqueues the batch on the estimator's
 * stream and returns; bb_estimator_query sets *done without blocking,
 * bb_estimator_wait blocks.  Device buffers must stay valid until done.
 * Batches enqueued while others are pending run after them in order (the
 * estimator's stage times then describe the last one); any other call on
 * the estimator waits for all pending batches first. */
int bb_estimate_batch_async(bb_estimator *e, int nfields, const double *d_fields_r2c,
                            double *d_bin_bs, double *d_bin_norm, double *d_bin_biana);
/* Mixed geometries (no counterpart in the reference, which runs one field
 * of one geometry per process): the estimates of n BF16x3 estimators of
 * different plans -- nfields[k] fields back to back at d_fields[k] for
 * estimator k, outputs as bb_estimate_batch_device's -- with the contraction
 * work of all of them in one persistent launch: a round of resident
 * workgroups takes the units of every field of every geometry from one
 * queue (the estimators in order of decreasing unit size, each one's fields
 * in turn, each field's units largest first).  The plane gathers and transforms run on each
 * estimator's stream before it, the binning after it; the call returns when
 * all are done.  The estimators must share one 4-wave layout shape (the
 * same occupancy, mix_fp32 0); each field's bins are the same bits as its
 * own estimator gives it alone. */
int bb_estimate_mixed(int n, bb_estimator *const *ests, const int *nfields,
                      const double *const *d_fields, double *const *d_bs, double *const *d_norm,
                      double *const *d_biana);
int bb_estimator_query(bb_estimator *e, int *done);
int bb_estimator_wait(bb_estimator *e);
/* The estimator's HIP stream (hipStream_t), for callers that order their own
 * copies against it. */
int bb_estimator_stream(const bb_estimator *e, void **stream);
/* Runs the estimator's work on `stream` (hipStream_t, owned by the caller and
 * valid while used) from now on; NULL returns to the estimator's own stream.
 * Waits for a pending batch first. */
int bb_estimator_set_stream(bb_estimator *e, void *stream);

/* Diagnostics of the last estimate (first field of the last batch): per-contraction triple-product sums in
 * the reference's enumeration order (ncontr doubles, scaled to the NN grid),
 * the GPU triangle counts Num per ring triple (rcount doubles, NN grid), and
 * one plane-ring field D_{plane,ring} on the grid of a level (NG*NG complex,
 * interleaved, NG = grids[level]) converted to FP64; BB_E_INVALID when the
 * level holds no such field. */
int bb_estimator_contraction_sums(bb_estimator *e, double *S);
int bb_estimator_triangle_counts(bb_estimator *e, double *num);
int bb_estimator_field(bb_estimator *e, int level, int plane, int ring, double *out);
/* Contraction work lists: out[0] work items, [1] workgroups (units) of the
 * launch, [2] 16x16 output tiles, [3] largest staged ring count of an item,
 * [4] sum of staged rings over items, [5] bytes the contraction reads from
 * the cache per estimate, [6] largest grid NG, [7] points per LDS chunk,
 * [8] waves per workgroup, [9] wave balance x1000 (tile windows / (waves x
 * busiest wave), summed over items), [10] grid levels, [11] partial tiles
 * (output tiles x their x-splits), [12] sum over contractions of NG^2 of
 * their level (points summed), [13] tile-window slots (items x waves x
 * windows per wave), [14] cost of the busiest waves (windows x chunks, summed
 * over items), [15] points of all plane-ring fields of all levels (the
 * cache), [16] band rows of the fields (the rows whose transforms along j the
 * ring pass sums, over all fields),
 * [17] pair-product points formed (16 rows per pair block, times the points
 * of its grid), [18] the part of [17] in rows that hold a pair, [19] tile
 * windows times the points of their grid.  The environment variables
 * BARBIE_CONTRACT_WAVES (4 or 8) and BARBIE_GRID_LEVELS (overrides
 * grid_levels) are read when the estimator is created. */
#define BB_WORKLIST_N 20
int bb_estimator_worklist(const bb_estimator *e, long long *out);
/* The same work-list figures for an estimator with these options, computed
 * on the host without a device ([13]: tile window slots executed, i.e.
 * items x waves x windows per wave; [14]: the units' summed cost, tile
 * windows of each item's busiest wave x chunks). */
int bb_plan_worklist(const bb_plan *p, const bb_options *opt, long long *out);
/* Device times of the last estimate in milliseconds: [0] plane gather,
 * [1] plane-ring FFTs, [2] contraction, [3] binning, [4] total.  A replayed
 * graph (use_graph) has no stage boundaries: [0]-[3] read 0. */
int bb_estimator_timings(bb_estimator *e, double *ms5);

/* Request server (no counterpart in the reference, which runs one field per
 * process: barbie.c:58-397 started by run:1 / many_fields.sh).  One scheduler
 * thread owns the current HIP device; fields of any geometry are submitted,
 * batched with queued fields of the same geometry (up to max_batch) and run
 * on up to max_inflight estimators at once, each on its own stream.  The
 * geometry with the shortest measured time per field starts first, except
 * that a request waiting longer than max(50 ms, 8 x its expected time) goes
 * first, oldest first; fields expected to take over 2 ms hold at most
 * max_inflight - 1 lanes, and a batch holds about 8 ms of fields.  An
 * estimator ("lane") is created only when its device bytes fit the HBM
 * budget; idle lanes of other geometries are evicted least recently used
 * first.  A field's bins are the same bits whatever else runs. */
typedef struct bb_server bb_server;
typedef struct bb_server_options {
    size_t hbm_budget;       /* device bytes for lanes (0: 80% of the free memory) */
    int max_inflight;        /* lanes running at once (0: 4; at most 16) */
    int max_batch;           /* fields per lane launch (0: 4; at most 256) */
    bb_precision precision;
    int tune;                /* 1: run bb_tune once per geometry, before its first lane */
    int reserved_cus;        /* CUs for short requests (expected under 2 ms): half of
                                the max_inflight streams run on them and take short
                                requests only, so that these never wait behind the
                                workgroups of long estimates; the other streams run
                                on the remaining CUs (0: 1/8 of the CUs; < 0: none) */
} bb_server_options;
enum { BB_REQ_QUEUED = 0, BB_REQ_RUNNING = 1, BB_REQ_DONE = 2, BB_REQ_FAILED = 3 };
int bb_server_create(const bb_server_options *opt, bb_server **out);
/* Finishes running batches, drops queued requests, frees everything. */
void bb_server_destroy(bb_server *s);
/* field_r2c: [NN][NN][NN+2] doubles; host memory (copied before return) or,
 * with on_device != 0, device memory that must stay valid until the request
 * is done.  *ticket identifies the request. */
int bb_submit(bb_server *s, const bb_geometry *g, const double *field_r2c, int on_device,
              long long *ticket);
/* Creates up to nlanes lanes (at most max_inflight) of a geometry now
 * (autotuning it first when the server tunes), so that traffic does not pay
 * for lane creation; the lanes count against the HBM budget like any other. */
int bb_server_prepare(bb_server *s, const bb_geometry *g, int nlanes);
/* paused != 0: start no new batches (running ones finish), so that a burst
 * of submissions queues up and batches together; 0 resumes. */
int bb_server_pause(bb_server *s, int paused);
/* *state = BB_REQ_*; does not block. */
int bb_poll(bb_server *s, long long ticket, int *state);
/* Blocks until the request is done, copies its bins (any pointer may be
 * NULL; [n_k1bin][n_mtbin][n_mtbin][16] each) and releases the ticket.
 * latency_ms (may be NULL): [0] queued, [1] running, [2] total, in ms. */
int bb_wait(bb_server *s, long long ticket, double *bin_bs, double *bin_norm,
            double *bin_biana, double *latency_ms);
/* out[13]: submitted, completed, failed, batches, fields in batches, lanes
 * created, lanes evicted, peak lane bytes, queued, running lanes, lanes,
 * lane bytes, budget. */
int bb_server_stats(bb_server *s, long long *out);

/* Normalisation and the 16 output files of barbie.c:383-397:
 * <dir>/BS_test_ii<ii>_N<NN>_R<NRi>.txt with the reference's row filter and
 * "%d\t%d\t%d\t%e\t%e\t%f\n" format. */
int bb_write_outputs(const bb_plan *p, const double *bin_bs, const double *bin_norm,
                     const double *bin_biana, const char *dir, int NRi);
/* Same files with the denominator checks the reference lacks
 * (barbie.c:388-393 divides every multipole by its own bin_normalization but
 * filters rows on the monopole's only): in a written row, a multipole whose
 * own denominator is not positive and finite, or whose B or Biana is not
 * finite, is written as "nan" in both value columns (the reference prints
 * whatever 0/0 or x/0 gives there).  *n_unsupported (may be NULL): the
 * number of such (bin, multipole) entries over the 16 files. */
int bb_write_outputs_checked(const bb_plan *p, const double *bin_bs, const double *bin_norm,
                             const double *bin_biana, const char *dir, int NRi,
                             long long *n_unsupported);

/* The reference's on-disk caches of a field, written into dir (which must
 * exist), in place of calCR's and wrapid_fast's files (funcs.c:328-415,
 * 537-566; read by read_rings, funcs.c:429-452, and barbie.c:312-320):
 * closed_rings_<NN> (rcount x {r1, r2, r3, Num} doubles), rcountFor_<NN>
 * (text), and dataPlaneRing_<NN>_<p>_<r>_R<NRi> for every p < Nmax, r < Nr
 * (D_{p,r} on the NN grid, 2*NN*NN doubles, row-major interleaved complex,
 * FP64).  Also cache_config_<NN>_R<NRi>.txt: the geometry and a 64-bit
 * FNV-1a key of it.  field_r2c: host [NN][NN][NN+2] doubles.  The fields are
 * built on the device (FP64, the estimator's transforms on the NN grid) in
 * batches of about 256 MiB. */
int bb_export_cache(bb_estimator *e, const double *field_r2c, const char *dir, int NRi);
/* Reads an existing ring-triple cache back, as read_rings does
 * (funcs.c:429-452: rcountFor_<NN> and closed_rings_<NN> in dir, written by
 * calCR, funcs.c:402-415, or by bb_export_cache), with the checks the
 * reference lacks (docs/REVIEW.md:9): when cache_config_<NN>_R<NRi>.txt is
 * present its key must be this geometry's; the count and the ring triples
 * must be this plan's, in calCR order; every Num must equal the estimator's
 * within rel_tol of the largest (*nmismatch, may be NULL: how many do not).
 * Any failure is BB_E_INVALID (BB_E_IO for missing or short files) and
 * changes nothing.  use_num != 0: the estimator then bins with the file's
 * Num values, as the reference does after read_rings. */
int bb_import_rings(bb_estimator *e, const char *dir, int NRi, double rel_tol, int use_num,
                    long long *nmismatch);
/* Checks an existing plane-ring cache (dataPlaneRing_<NN>_<p>_<r>_R<NRi> in
 * dir, wrapid_fast's files, funcs.c:537-566) against the fields of
 * field_r2c (host [NN][NN][NN+2] doubles): every stride-th (p, r) in (p, r)
 * order is built on the device (FP64, NN grid) and compared; *max_rel gets
 * the largest max |file - field| / max |field|.  The reference reuses such
 * files whenever its sentinel file exists (barbie.c:193-196), also when they
 * belong to another field or geometry; a large *max_rel says so.
 * BB_E_IO when a file is missing or short. */
int bb_check_plane_cache(bb_estimator *e, const double *field_r2c, const char *dir, int NRi,
                         int stride, double *max_rel);

/* Power-spectrum multipoles of a field: the program calpow_mom
 * (calpow_mom.c:22-117) and its calpow_mom function (funcs.c:860-1097) on the
 * current device.  field_r2c: [NN][NN][NN+2] doubles, host memory, or device
 * memory when on_device != 0.  Nbin logarithmic |k| bins (1..10; the
 * reference uses 10), cell size LL (the reference uses 0.56).  out: Nbin rows
 * of {kmean, P0, P2, P4, P6, num}, normalised as funcs.c:1055-1071 does (the
 * columns of psoutN<NN>_R<NRi>.dat, calpow_mom.c:104-113).  Mode bins match
 * the reference exactly; the Legendre weights are float, as there.  The work
 * runs on a stream of the library's own, so a device field must be complete
 * before the call.  The first call on a device keeps 0.7 MB of partial sums
 * and that stream until the process ends; calls are serialised. */
int bb_power_multipoles(const double *field_r2c, int on_device, int NN, int Nbin, double LL,
                        double *out);
/* No counterpart in the reference: bb_power_multipoles for callers that
 * cannot allocate device memory (code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc).  It calls no hipMalloc, hipFree,
 * stream creation or allocating library routine: its partial sums live in a
 * static device array of the library (about 0.4 MB per device, held until
 * the process ends), and its kernels run on the caller's `stream`
 * (hipStream_t; NULL: the null stream), which is synchronised before return.
 * d_field_r2c: [NN][NN][NN+2] doubles in device memory, complete on `stream`
 * (or before the call).  Limits: Nbin <= 10 (the array holds 10 bins), NN
 * even in 4..4096; larger sizes fail with BB_E_INVALID.  out: as
 * bb_power_multipoles, bit for bit.  Every call on a device shares that
 * device's array (the geometry sums are recomputed each call, about 40 us at
 * NN=384); the library serialises calls, so calls from several threads do
 * not overlap on the array. */
int bb_power_multipoles_noalloc(const double *d_field_r2c, int NN, int Nbin, double LL,
                                void *stream, double *out);

/* Seeded input field on the GPU: the step before the estimator, the
 * reference's generator (generate_field.c:22-107: Fill_Modes, Make_Image,
 * Non_Gauss, Get_Modes, RSD of ngrf_2d_funcs.c:341-577, 734-765) with a
 * counter-based random stream (Philox4x64-10, one block per Fourier mode,
 * Box-Muller) in place of GSL's mt19937 + ziggurat, which this build cannot
 * reproduce.  Output: the FFTW in-place r2c buffer [NN][NN][NN+2] doubles
 * that barbie reads (host memory, or device memory when on_device != 0).
 * Fourier modes get Re, Im ~ N(0, V P(k) / 2), V = (NN LL)^3, P(k) = amp
 * k^index with k = kunit |(ia, ja, kz)| (kunit 1: grid units; the
 * reference's 2 pi / (NN LL)), zero at k = 0 and for k >= kcut (kcut <= 0:
 * no cut); the kz = 0 and NN/2 planes are Hermitian.  Then delta = c2r / V,
 * delta += fnl (delta^2 - <delta^2>), times LL^3, r2c, times 1 + beta mu^2.
 * stage 1 stops after the Gaussian modes (before the division by V).
 * NN: even, 4..1024, prime factors 2, 3, 5.  The work runs on a stream of
 * the library's own, synchronised before return; calls are serialised. */
typedef struct bb_field_params {
    double amp;     /* P(k) amplitude (the workload: 1e3) */
    double index;   /* P(k) exponent (-1.5) */
    double kunit;   /* k per grid index (1) */
    double kcut;    /* modes with k >= kcut are zero (NN/3); <= 0: none */
    double fnl;     /* non-Gaussian coefficient (0.3) */
    double beta;    /* Kaiser factor (1) */
    double LL;      /* cell size (1; the reference's generator uses 0.56) */
    int stage;      /* 0: the field; 1: the Gaussian Fourier modes only */
} bb_field_params;
int bb_generate_field(int NN, unsigned long long seed, const bb_field_params *p,
                      double *field_r2c, int on_device);

#ifdef __cplusplus
}
#endif
#endif /* BARBIE_H */
