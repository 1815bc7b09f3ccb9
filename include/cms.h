/*
 * cms.h — C ABI of the MI355X-native all-pairs Meta-Storms distance library
 * (libcms.so).
 *
 * The reference has no library boundary: its GPU path is a pair of CUDA kernels
 * launched from a host loop, and its CPU path is a member function. Each entry
 * point below names the reference interface it replaces (file:line in
 * qdu-bioinfo/Coordinated-Meta-Storms).
 *
 * Conventions
 *   - Every function returns 0 (CMS_OK) on success or a negative CMS_E_* code.
 *     Nothing in the library calls exit(); cms_last_error() gives the message
 *     of the last failure on a context (or of the last context-less failure
 *     when passed NULL).
 *   - Host buffers are owned by the caller; the library copies what it needs.
 *     Device memory is owned by the library (or by the caller for *_device
 *     entry points, which take device pointers).
 *   - One context per GPU. A context is not thread-safe; callers serialise.
 *   - Similarities are float in [0, 1]; the packed upper triangle uses the
 *     reference's index p = m*N + n - (m+1)(m+2)/2 for m < n
 *     (cms_cuda.cu:1801-1811, comp_sam.cpp:346).
 */
#ifndef CMS_H
#define CMS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CMS_REG_SIZE 150 /* register-slot space of order.txt (comp.h:26) */

enum {
    CMS_OK = 0,
    CMS_E_INVAL = -1,      /* bad argument */
    CMS_E_NOMEM = -2,      /* host or device allocation failed */
    CMS_E_HIP = -3,        /* HIP runtime error (no device, launch failure) */
    CMS_E_SCHEDULE = -4,   /* order.txt schedule invalid (slot out of range, read before write, leaf out of range) */
    CMS_E_STATE = -5,      /* call out of order (no tree / no samples loaded) */
    CMS_E_IO = -6,         /* file could not be opened or parsed */
    CMS_E_UNSUPPORTED = -7, /* valid input that the selected kernel does not handle */
    CMS_E_RANGE = -8,       /* sample mass outside the fixed-point range of the resident scale */
    CMS_E_TIMEOUT = -9      /* cms_server_wait: the request has not finished yet */
};

/* Kernel families. */
enum {
    CMS_KERNEL_WALK = 0,     /* drop-in: the reference's FP32 arithmetic per pair, bit-exact with comp:
                                the post-order walk (-M 0/1), a merge of the two leaf lists (-M 2..5;
                                -M 4 with the host C library's logf) */
    CMS_KERNEL_SHAREDMIN = 1 /* fast: decayed-mass embedding + fixed-point shared-minimum tiles (-M 0/1),
                                leaf tiles with FP64 sums (-M 2..5) */
};

/* Distance metric (comp -M). */
enum {
    CMS_METRIC_WEIGHTED = 0,   /* -M 0, Meta-Storms (comp.h:332-375) */
    CMS_METRIC_UNWEIGHTED = 1, /* -M 1, unweighted Meta-Storms (comp.h:378-446); walk (bit-exact) or shared-minimum */
    CMS_METRIC_COS = 2,        /* -M 2..5: leaf-level metrics, dist.h:13-159 via Calc_sim(a, b, mode) */
    CMS_METRIC_EU = 3,         /*   (comp.h:296-330), similarity = 1 - dist; cms_all_pairs and */
    CMS_METRIC_JSD = 4,        /*   cms_all_pairs_device only, no tree schedule needed (order_n may */
    CMS_METRIC_BC = 5          /*   be 0) */
};

/* Abundance storage of the walk kernel (arithmetic is always FP32). */
enum {
    CMS_STORE_F32 = 0,
    CMS_STORE_BF16 = 1, /* per-sample power-of-two scale, Kahan-compensated total */
    CMS_STORE_F16 = 2
};

typedef struct cms_opts {
    int kernel;       /* CMS_KERNEL_* (default WALK) */
    int storage;      /* CMS_STORE_* (default F32; WALK only) */
    int tile_a;       /* SHAREDMIN row-block samples per wave tile: 0 or 32 (others CMS_E_INVAL) */
    int tile_b;       /* SHAREDMIN column-block samples per wave tile: 0 or 32 */
    int chunk;        /* SHAREDMIN entries staged per LDS chunk: 0 (any) or 8, 16, 32; restricts
                         the variants the default choice and the autotuner use */
    int use_graph;    /* capture per-band tile launches in a hipGraph (SHAREDMIN) */
    int metric;       /* CMS_METRIC_* (default WEIGHTED) */
    int bin;          /* SHAREDMIN: store samples binned by sparsity (densest first) */
    int autotune;     /* SHAREDMIN: time the persistent-kernel variants once, keep the fastest */
    int variant;      /* SHAREDMIN: 1..16 force a persistent-kernel variant (0 = default/tuned) */
    int embed_slice;  /* SHAREDMIN: samples per embedding slice, multiple of 64 (0 = by a 16 GiB layout) */
    int embed_input;  /* SHAREDMIN embedding input: 0 = the samples' leaf lists where the leaf stream
                         has at most 253,952 positions, else the F32 walk layout; 1 = the walk layout */
    int wide_rows;    /* SHAREDMIN split tiles: 0 = rows read as 32-bit offsets from each bitmap window's
                         first row below 699,051 padded samples, 64-bit row addresses above; 1 = always
                         64-bit row addresses (same results) */
    int reserved;
} cms_opts;

typedef struct cms_ctx cms_ctx;

/* ---------------------------------------------------------------- context */

void cms_opts_default(cms_opts *o);

/* Replaces the per-GPU setup of Multi_GPU_split_MetaStorm (cms_cuda.cu:1470-1652). */
int cms_create(int device, const cms_opts *o, cms_ctx **out);
void cms_destroy(cms_ctx *c);
const char *cms_last_error(const cms_ctx *c);
int cms_set_opts(cms_ctx *c, const cms_opts *o);

/* Tree schedule, arrays exactly as loaded from tree/order.txt
 * (_Comp_Tree::Load_order, comp.h:161-192; getters comp.h:62-84).
 * Copied; validated (slot range, read-before-write, leaf range). */
int cms_set_tree(cms_ctx *c, int leaf_n, int order_n,
                 const int32_t *order_1, const float *dist_1,
                 const int32_t *order_2, const float *dist_2,
                 const int32_t *order_d);

/* Dense abundances, n x leaf_n row-major, percent scale as produced by
 * _Comp_Tree::Load_abd (comp.h:243-279). Replaces the per-sample device
 * rows of cms_cuda.cu:1617-1652. */
int cms_set_samples(cms_ctx *c, int n, const float *abd);
/* Same data in CSR form: row s holds (leaf, value) for its nonzero leaves. */
int cms_set_samples_csr(cms_ctx *c, int n, const int64_t *rowptr,
                        const int32_t *leaf, const float *val);

/* All N(N-1)/2 similarities into a caller-owned host buffer (packed upper
 * triangle). Replaces the tile loop of Multi_GPU_split_MetaStorm
 * (cms_cuda.cu:1523-2196) and the CPU loop of Multi_Comp_Table
 * (comp_sam.cpp:338-349). */
int cms_all_pairs(cms_ctx *c, float *sim_packed);
/* Same, into a device buffer of N(N-1)/2 floats, enqueued on `stream`
 * (hipStream_t, NULL = the context's stream). No host synchronisation.
 * d_sim_packed NULL = a device buffer owned by the context (read it back with
 * cms_copy_result). */
int cms_all_pairs_device(cms_ctx *c, float *d_sim_packed, void *stream);
/* Copy the first `count` floats of the context-owned result buffer to host
 * memory (waits for the context's stream). */
int cms_copy_result(cms_ctx *c, float *host, int64_t count);
/* Wait for all work enqueued by the context. */
int cms_synchronize(cms_ctx *c);
/* Number of HIP devices visible to the library (negative on error). */
int cms_device_count(void);
/* Diagnostic, no counterpart in the reference: run a plain device-to-device
 * copy kernel of `bytes` bytes `reps` times on the context's stream and return
 * its average duration in *ms. Used to calibrate hardware traffic counters
 * (FETCH_SIZE / WRITE_SIZE) against a known byte count. */
int cms_calibrate_copy(cms_ctx *c, int64_t bytes, int reps, double *ms);

/* Streamed all-pairs for sample sets whose triangle does not fit in host
 * memory at once (the reference keeps it whole, cms_cuda.cu:1485-1487, and
 * walks tiles, 1523-2196). The triangle of STORED positions (see
 * cms_sample_order; the identity unless opts.bin) is computed in bands of
 * about band_pairs pairs; each band is copied to pinned host memory while the
 * next one computes, and fn(user, first, count, sims) is called on the
 * calling thread with sims[k] = sim of stored packed index first + k. The
 * pointer is valid during the call only. With opts.use_graph the band's
 * launches are replayed from a hipGraph. */
typedef void (*cms_band_fn)(void *user, int64_t first, int64_t count, const float *sims);
int cms_all_pairs_stream(cms_ctx *c, int64_t band_pairs, cms_band_fn fn, void *user);
/* order[p] = caller's sample index stored at position p (n entries). */
int cms_sample_order(cms_ctx *c, int32_t *order);

/* Explicit pair list: sim[i] = sim(row[i], col[i]). Host arrays. The
 * reference kernels' boundary (CalcSimOfDiagonalSquare cms_cuda.cu:52-54 and
 * CalcSimOfNormalRectangle cms_cuda.cu:495-497 with d_order_row/d_order_col). */
int cms_sim_pairs(cms_ctx *c, int64_t n_pairs, const int32_t *row,
                  const int32_t *col, float *sim);

/* The same pair list with no device allocation. No counterpart in the
 * reference: for callers that cannot allocate device memory (tight serving
 * loops, harnesses that forbid hipMalloc). The device layouts must already be
 * built for the context's current kernel and options (cms_prepare, or an
 * earlier computing call), else CMS_E_STATE; the pair lists and results are
 * staged in a static per-device scratch of CMS_NOALLOC_MAX_PAIRS pairs
 * (12 MiB, held until the process ends); more pairs are CMS_E_INVAL. All
 * contexts on one device share that scratch: calls that overlap on one
 * device race. */
#define CMS_NOALLOC_MAX_PAIRS (1 << 20)
int cms_sim_pairs_noalloc(cms_ctx *c, int64_t n_pairs, const int32_t *row,
                          const int32_t *col, float *sim);

/* Rectangular tile rows [r0, r0+nr) x cols [c0, c0+nc) into a host buffer of
 * nr*nc floats, row-major. Diagonal entries (same sample) are 1. Replaces one
 * iteration of the tile loop (cms_cuda.cu:1523-1539, defineOrder 961-985). */
int cms_tile(cms_ctx *c, int r0, int nr, int c0, int nc, float *out);

/* Rebuild the device layouts of the selected kernel from the uploaded
 * samples (force != 0: even if current), reusing device allocations: the
 * block-leaf-major walk layout, and for CMS_KERNEL_SHAREDMIN the fixed-point
 * embedding and its block bitmaps. Runs implicitly on first use; a caller
 * timing the whole path from resident inputs calls it once per pass. */
int cms_prepare(cms_ctx *c, int force);

/* Diagnostics of the last all-pairs run (waits for its kernel to finish). */
typedef struct cms_stats {
    double kernel_ms;          /* device time of the pair kernels (HIP events) */
    double prep_ms;            /* device time of layout / embedding kernels */
    double pair_node_updates;  /* node-steps executed (pruned count for SHAREDMIN) */
    double dense_updates;      /* pairs x OrderN */
    double valu_ops;           /* algorithmic VALU lane-ops executed */
    double bytes_moved;        /* algorithmic HBM bytes of the pair kernels */
    int64_t tiles;             /* tiles launched */
    int nslots;                /* register slots after re-allocation */
    int scale_bits;            /* SHAREDMIN fixed-point scale (2^scale_bits per percent) */
    int entries;               /* SHAREDMIN active embedding entries */
    int variant;               /* SHAREDMIN persistent-kernel variant in use (1..16) */
    double tune_ms[16];        /* autotuner timings per variant (0 = not timed) */
    int64_t bands;             /* bands of the last streamed run */
    int64_t entries_visited;   /* SHAREDMIN: entries set in both blocks' bitmaps, summed over tiles */
    int64_t entries_staged;    /*   of those, accumulated for all 32 x 32 pairs of the tile */
    int64_t entries_sparse;    /*   of those, taken by the sparse path (at most 3 samples with mass in each block) */
    int64_t sparse_pairs;      /*   pairs with mass on both sides in the sparse entries */
    int64_t entries_half;      /*   of those, taken by the half path (at most 3 samples with mass in one block) */
    int64_t half_pairs;        /*   pairs with mass on both sides in the half entries */
} cms_stats;
int cms_get_stats(cms_ctx *c, cms_stats *s);

/* ------------------------------------------------------------- serving
 *
 * No counterpart in the reference (its only mode is one all-pairs run per
 * process, cms_cuda.cu:1470-2235). A server keeps the shared-minimum layout
 * of a context's samples (the references) resident in HBM and compares
 * query sets against all of them:
 *   - query sets are queued FIFO and admitted while query slots and a
 *     device-memory budget (an arena allocated at creation) have room;
 *   - admitted sets are embedded on their own stream while the device runs
 *     steps; a step compares every active query with one band of reference
 *     blocks; a set joins at the next step boundary, is compared with each
 *     band exactly once over `bands` consecutive steps, and leaves (its slots
 *     and memory are reused) — continuous batching over the band cycle;
 *   - results are exact integer sums of the fixed-point embedding with the
 *     references' scale, so a query's output does not depend on what else
 *     is in flight.
 * The context must outlive the server and is read-only while it exists
 * (cms_set_* and cms_prepare return CMS_E_STATE). */
typedef struct cms_server cms_server;

typedef struct cms_server_opts {
    int slots;             /* query slots resident at once, multiple of 32 (default 1024) */
    int bands;             /* reference bands per cycle (default 16) */
    int64_t memory_budget; /* bytes of the per-request arena (0 = 8 GiB or less if the device has less free) */
    int variant;           /* cross-kernel variant: 1..6 plain tiles as cms_opts.variant 1..6, 7..8 split tiles
                              (8-row chunk, 1 or 2 waves per workgroup); 0 = 7 */
    int reserved[3];
} cms_server_opts;

typedef struct cms_request_info {
    int rc;                /* CMS_OK or the request's error */
    int nq;                /* queries */
    int64_t join_step;     /* step at which it joined the cycle (-1 = never) */
    double t_submit;       /* seconds on the server's clock */
    double t_admit;        /* slots and memory granted, embedding queued */
    double t_join;         /* first step launched */
    double t_done;         /* output delivered to the caller's buffer */
    double t_computed;     /* last step finished on the device (seen by the server within ~0.1 ms; -1 if
                              the set was delivered before the server saw it) */
} cms_request_info;

typedef struct cms_server_stats {
    int64_t submitted, completed, failed;
    int64_t queued, active;        /* now */
    int64_t steps;                 /* steps launched */
    double kernel_ms;              /* device time of the cross kernels (finished steps) */
    double pairs;                  /* query x reference pairs of finished requests */
    double pair_entries;           /* entries visited by the cross tiles (x 1024 pairs) */
    int64_t arena_bytes, arena_peak;
    int slots, slots_peak, bands;
    double uptime_s;
    double entries_staged;         /* of the visited entries: staged for all 32 x 32 pairs */
    double entries_sparse, sparse_pairs;  /* split cross tiles: light in both blocks, pairs with mass there */
    double entries_half, half_pairs;      /*   light in one block, pairs evaluated there */
} cms_server_stats;

void cms_server_opts_default(cms_server_opts *o);
/* Build (if needed) the context's shared-minimum layout for its metric and
 * start the serving threads. The walk layout of the context is released. */
int cms_server_create(cms_ctx *ref, const cms_server_opts *o, cms_server **out);
/* Queue nq query samples (CSR over leaves, abundances as for
 * cms_set_samples_csr; copied before return). When the request finishes,
 * out (host, nq x n_ref floats) holds row q = similarity of query q to every
 * reference in the context's sample order. Fails at once (and queues
 * nothing) on bad input, on a set that can never fit (nq > slots or memory
 * beyond the budget), on negative values, or on a query whose mass exceeds
 * the references' fixed-point range (CMS_E_RANGE). */
int cms_server_submit(cms_server *s, int nq, const int64_t *rowptr, const int32_t *leaf, const float *val,
                      float *out, int64_t *id);
/* Wait up to timeout_s (< 0: forever) for request id; CMS_E_TIMEOUT if it
 * is still running, else its rc. info may be NULL. A finished request is
 * forgotten after the first successful wait. */
int cms_server_wait(cms_server *s, int64_t id, double timeout_s, cms_request_info *info);
/* Wait until every submitted request has finished. */
int cms_server_drain(cms_server *s, double timeout_s);
int cms_server_get_stats(cms_server *s, cms_server_stats *st);
/* Stops accepting requests, finishes the ones in flight, frees the arena. */
void cms_server_destroy(cms_server *s);

/* ------------------------------------------------- host loader (no GPU) */

/* Validate a schedule without a device (the checks of cms_set_tree):
 * register range, read-before-write, leaf range. Outputs (optional): live
 * register slots after re-allocation, whether it is a tree (every leaf and
 * value read at most once), whether all 1 - w >= 0. */
int cms_validate_tree(int leaf_n, int order_n, const int32_t *order_1, const float *dist_1,
                      const int32_t *order_2, const float *dist_2, const int32_t *order_d,
                      int *nslots, int *is_tree, int *nonneg_beta);

/* Tree database as configured by $ParallelMETA/databases/db.config
 * (_PMDB, db.h:21-194, 287-358; _Comp_Tree::Init comp.h:122-137). */
typedef struct cms_db cms_db;

/* root = the ParallelMETA directory (NULL: read the environment variable,
 * utility.h:27-41). db_letter as for -D; unknown letters fall back to 'G'
 * (db.h:44). */
int cms_db_open(const char *root, char db_letter, cms_db **out);
void cms_db_close(cms_db *db);
int cms_db_leaf_n(const cms_db *db);
int cms_db_order_n(const cms_db *db);
int cms_db_has_tree(const cms_db *db);
int cms_db_has_cp(const cms_db *db);
/* Pointers into the database (valid until cms_db_close). */
int cms_db_order(const cms_db *db, const int32_t **order_1, const float **dist_1,
                 const int32_t **order_2, const float **dist_2,
                 const int32_t **order_d);
const char *cms_db_leaf_name(const cms_db *db, int i);

/* OTU count table (-T), _Table_Format::Load_Table (table_format.h:110-165):
 * header = title + feature ids, rows = sample id + exactly F numbers; rows of
 * the wrong width are skipped. */
typedef struct cms_table cms_table;
int cms_table_load(const char *path, cms_table **out);
/* Build a table from memory: n samples, f features, dense counts n x f. */
int cms_table_from_dense(int n, int f, const char *const *sample_names,
                         const char *const *feature_names, const float *counts,
                         cms_table **out);
void cms_table_free(cms_table *t);
int cms_table_n(const cms_table *t);
int cms_table_f(const cms_table *t);
const char *cms_table_sample_name(const cms_table *t, int i);
const char *cms_table_feature_name(const cms_table *t, int j);
int cms_table_skipped(const cms_table *t); /* rows rejected by the width check */
/* Sparse table file: the same table (feature and sample names, counts) in
 * CSR form, little-endian binary ("CMSTCSR1", u64 samples, features,
 * entries; u32-length-prefixed names; i64 row offsets, i32 ascending column
 * indices, f32 counts). No counterpart in the reference, whose only table is
 * the dense text of table_format.h:110-165: a 10,000 x 99,322 text table is
 * about 2 GB, its sparse file about 90 MB. cms_table_load reads either form
 * (the magic selects it) and every cms_table_* function treats them alike. */
int cms_table_save_csr(const cms_table *t, const char *path);
int cms_table_is_sparse(const cms_table *t);
/* Build a sparse table from memory: n samples, f features, row s =
 * (col[k], counts[k]) for k in [rowptr[s], rowptr[s+1]), columns strictly
 * ascending. */
int cms_table_from_csr(int n, int f, const char *const *sample_names, const char *const *feature_names,
                       const int64_t *rowptr, const int32_t *col, const float *counts, cms_table **out);

/* Dense percent-scale abundances n x leaf_n, _Comp_Tree::Load_abd
 * (comp.h:243-279) with Check_OTU (utility.h:181-190) and copy-number
 * correction (db.h:251-280). mapped[i] (optional) = mapped OTU count. */
int cms_table_abundance(const cms_db *db, const cms_table *t, int cp_correct,
                        float *abd, int *mapped);
/* The same abundances as CSR over leaves (ascending leaf ids per sample;
 * arrays allocated by the library, free each with cms_free): the input of
 * cms_set_samples_csr without the n x leaf_n dense matrix. */
int cms_table_abundance_csr(const cms_db *db, const cms_table *t, int cp_correct,
                            int64_t **rowptr, int32_t **leaf, float **val);

/* ---------------------------------------- sample files and lists (-i / -l)
 *
 * One sample per classification file. Replaces _Comp_Tree::Load_abd(const
 * char *, float *, bool) (comp.h:194-236) over
 * _OTU_Parser::Load_file_to_hash (otu_parser.h:104-157): a title line
 * starting "#S" selects the per-read format ("<read> <otu> ..." per line,
 * one count each), any other title the count format ("<otu> <count> ...",
 * the last line of an id wins); count / copy number per OTU, FP32 total in
 * the reference's hash-map order, a = x / (total / 100) for the OTUs in the
 * tree. abd: leaf_n floats (zeroed first); *mapped = OTUs found in the tree
 * (the number `cuda-comp -i` prints). A file that cannot be opened is
 * CMS_E_IO; counts summing to zero give an all-zero sample. */
int cms_sample_file_abundance(const cms_db *db, const char *path, int cp_correct,
                              float *abd, int *mapped);

/* The same for n files as CSR over leaves (zero abundances dropped; loaded
 * in parallel; arrays allocated by the library, free each with cms_free);
 * mapped: n ints or NULL. Replaces the Load_abd loop of Multi_Comp
 * (cms_cuda.cu:1362-1375, comp_sam.cpp:240-258). */
int cms_sample_files_csr(const cms_db *db, int n, const char *const *paths, int cp_correct,
                         int64_t **rowptr, int32_t **leaf, float **val, int *mapped);

/* Sample list of `-l list [-p prefix]`: Load_List(list, files, ids, prefix)
 * (utility.h:344-391). Per non-empty line: "<id> <file> ..." with two or
 * more fields, else the whole line is the file and the id is the directory
 * holding it. The prefix is prepended to every file. */
typedef struct cms_list cms_list;
int cms_list_load(const char *path, const char *prefix, cms_list **out);
int cms_list_n(const cms_list *l);
const char *cms_list_path(const cms_list *l, int i);
const char *cms_list_name(const cms_list *l, int i);
void cms_list_free(cms_list *l);

/* .dist writer, Output_Matrix (cms_cuda.cu:1305-1342 = comp_sam.cpp:184-219):
 * header "\t"+names; row i = name then "\t"+value per column, value
 * = 1.0 - (double)sim (0 on the diagonal), or sim (1 on the diagonal) when
 * is_sim; values printed like std::ostream default (6 significant digits). */
int cms_write_dist(const char *path, int n, const char *const *names,
                   const float *sim_packed, int is_sim);

/* ---------------------------------------- database builder (no GPU)
 *
 * The reference only reads tree/id.txt and tree/order.txt
 * (_Comp_Tree::Load_id / Load_order, comp.h:139-192); these build them from
 * a Newick tree. Leaves are numbered in order of appearance; a node with one
 * child is merged into its child's edge; a node with k > 2 children becomes
 * a chain of binary nodes joined by zero-length edges; nodes are evaluated in
 * Sethi-Ullman order with lowest-free registers written as -(slot + 1). */
typedef struct cms_newick cms_newick;
int cms_newick_parse(const char *text, int64_t len, cms_newick **out);
int cms_newick_load(const char *path, cms_newick **out);
void cms_newick_free(cms_newick *t);
int cms_newick_leaf_n(const cms_newick *t);
int cms_newick_order_n(const cms_newick *t);
const char *cms_newick_leaf_name(const cms_newick *t, int i);
/* Schedule arrays (valid until cms_newick_free), as cms_set_tree takes them. */
int cms_newick_order(const cms_newick *t, const int32_t **order_1, const float **dist_1,
                     const int32_t **order_2, const float **dist_2, const int32_t **order_d);
int cms_newick_info(const cms_newick *t, int *nslots, int *multifurcations, int *unary_merged);
/* root/databases/db.config (entry 'G' -> name), root/databases/name/tree/
 * {id,order}.txt, and name/copy_number.txt copied from cp_path if not NULL. */
int cms_newick_write_db(const cms_newick *t, const char *root, const char *name, const char *cp_path);

/* ------------------------------------------- synthetic inputs (no GPU) */

/* Seeded synthetic tree: random binary tree (Yule process), branch lengths
 * ~ Exponential(bl_mean) clipped to [0, bl_max] and rounded to 6 decimals,
 * order.txt emitted in Sethi-Ullman post-order with lowest-free slot
 * allocation. Arrays sized leaf_n - 1 (order) are written by the callee into
 * caller buffers; leaf names are "T<index>". */
int cms_synth_tree(uint64_t seed, int leaf_n, double bl_mean, double bl_max,
                   int32_t *order_1, float *dist_1, int32_t *order_2,
                   float *dist_2, int32_t *order_d, int *max_slots);

typedef struct cms_synth_samples_opts {
    uint64_t seed;
    int n_samples;
    int leaf_n;
    int nnz_lo, nnz_hi;   /* nonzero OTUs per sample */
    int nnz_log_uniform;  /* 0: uniform integer in [lo, hi]; 1: log-uniform */
    double zipf_s;        /* OTU popularity exponent over a random rank order */
    double nb_mean;       /* NegBinomial mean of counts (Gamma-Poisson) */
    double nb_dispersion; /* NegBinomial dispersion r (shape of the Gamma) */
} cms_synth_samples_opts;

/* Counts in CSR form, columns ascending. rowptr has n+1 entries; col/cnt
 * are allocated by the library (free with cms_free). */
int cms_synth_samples(const cms_synth_samples_opts *o, int64_t **rowptr,
                      int32_t **col, float **cnt);
/* Copy numbers uniform integer in [1, cp_max] per leaf (seeded). */
int cms_synth_copy_numbers(uint64_t seed, int leaf_n, int cp_max, float *cp);
/* Percent-scale abundances from CSR counts with the Load_abd arithmetic
 * (comp.h:252-275): x = count / cp, FP32 total in column order, a = x / total.
 * cp may be NULL (no correction). Output CSR values in place of counts. */
int cms_counts_to_abundance_csr(int n, const int64_t *rowptr, const int32_t *col,
                                const float *cnt, const float *cp, float *abd_val);
/* Write a ParallelMETA-style database directory: root/databases/db.config
 * (entry 'G' -> name), root/databases/name/tree/{id,order}.txt and
 * copy_number.txt (if cp != NULL). */
int cms_write_db(const char *root, const char *name, int leaf_n, int order_n,
                 const int32_t *order_1, const float *dist_1,
                 const int32_t *order_2, const float *dist_2,
                 const int32_t *order_d, const float *cp);
/* Write a dense OTU count table (text) from CSR counts over leaf_n columns. */
int cms_write_table(const char *path, int n, int leaf_n, const int64_t *rowptr,
                    const int32_t *col, const float *cnt);
/* The same table as a sparse table file (see cms_table_save_csr). */
int cms_write_table_csr(const char *path, int n, int leaf_n, const int64_t *rowptr,
                        const int32_t *col, const float *cnt);

void cms_free(void *p);
const char *cms_version(void);

#ifdef __cplusplus
}
#endif
#endif /* CMS_H */
