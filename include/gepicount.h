/*
 * gepicount.h -- C ABI of the MI355X gEpiCount scan engine (libgepicount.so).
 *
 * Plain C types only: pointers, sizes, status codes.  The library owns all
 * device memory; callers own every host buffer they pass in.  Every entry
 * point returns GEPI_OK (0) or a negative gepi_status and never throws;
 * gepi_last_error() gives the message of the last failure on the calling
 * thread.  There is no CPU fallback: without a usable gfx950 GPU the scan
 * entry points fail with GEPI_ERR_DEVICE.
 */
#ifndef GEPICOUNT_H
#define GEPICOUNT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GEPI_NUM_FUNCS 5
#define GEPI_NUM_CLASSES 2
/* Upper bound on samples per class layout; the count kernels keep all table
 * arithmetic in 32-bit integers up to this size.  The reference caps samples
 * at 8192 (CONST_MASK_MAX_SIZE, inc/gEpiCount.h:51; kernel1.cu:329-333).    */
#define GEPI_MAX_SAMPLES 46340
#define GEPI_DEFAULT_TOPMOST 1000000 /* src/gEpiCount.cpp:34 */

/* Boolean functions: enum BinaryFunction, inc/gEpiCount.h:58-66.  The names
 * follow the reference: NAND = ~A & B, ANDN = A & ~B, NANDN = ~(A | B).      */
typedef enum gepi_bfunc {
    GEPI_BFUNC_AND = 0,
    GEPI_BFUNC_NAND = 1,
    GEPI_BFUNC_ANDN = 2,
    GEPI_BFUNC_NANDN = 3,
    GEPI_BFUNC_XOR = 4
} gepi_bfunc;

/* Objectives: enum FitnessFunction, inc/gEpiCount.h:68-73. */
typedef enum gepi_ffunc { GEPI_FFUNC_ENTROPY = 0, GEPI_FFUNC_FISHERPV = 1 } gepi_ffunc;

typedef enum gepi_status {
    GEPI_OK = 0,
    GEPI_ERR_ARG = -1,      /* invalid argument                               */
    GEPI_ERR_LAYOUT = -2,   /* rows not paired as (allele 1, allele 2) sharing valid bits */
    GEPI_ERR_LIMIT = -3,    /* size beyond a documented limit                 */
    GEPI_ERR_NOMEM = -4,    /* device or host allocation failed               */
    GEPI_ERR_DEVICE = -5,   /* no usable GPU / HIP runtime error              */
    GEPI_ERR_IO = -6,       /* file missing or unreadable                     */
    GEPI_ERR_FORMAT = -7,   /* malformed .bed/.bim/.fam                       */
    GEPI_ERR_NOCLASS = -8,  /* phenotype class 1 or 2 absent                  */
    GEPI_ERR_INTERNAL = -9
} gepi_status;

/* One ranked interaction.  Replaces InteractionResult (inc/gEpiCount.h:75-82):
 * sA >= sB are allele-row indices, (w, x, y, z) is count_type4 n, obj the
 * objective (entropy delta -I(F;C) in bits, or log10 Fisher p), fun the
 * gepi_bfunc.  48 bytes, 8-byte aligned.                                     */
typedef struct gepi_result {
    int64_t sA;
    int64_t sB;
    int32_t w, x, y, z;
    double obj;
    int32_t fun;
    int32_t reserved;
} gepi_result;

/* Scan parameters.  Replaces struct _paramP1 (inc/gEpiCount.h:116-132);
 * the file name and QA thresholds stay with the caller.                      */
typedef struct gepi_params {
    int32_t gpu_num;       /* -G: HIP device ordinal                           */
    int32_t ffun;          /* -f: gepi_ffunc                                   */
    int64_t topmost;       /* -N: results to keep                              */
    int32_t count_path;    /* 0 = fastest (3 from 10 K steps of 128 samples,
                            * else 4; 4 for Fisher in gepi_scan_noalloc),
                            * 1 = popcount (VALU), 2 = int8 MFMA, 3 = fp4
                            * (e2m1) MFMA on 32 x 32 wave tiles, 4 = fp4 MFMA
                            * on LDS-staged 64 x 32 tiles                    */
    int32_t intermediate;  /* -D: report progress and intermediate results
                            * through the thread's progress hook              */
    int32_t blk_x, blk_y;  /* --block-sizes: accepted for compatibility        */
    int32_t max_grid_area; /* --max-grid-area: cap on evals per launch, in Mi  */
    int32_t reserved;
} gepi_params;

/* Per-scan measurements filled by the scan entry points (may be NULL). */
typedef struct gepi_scan_stats {
    double ms_scan;          /* host wall time of the scan, panel resident     */
    double ms_count_kernels; /* sum of count-kernel durations (HIP events)     */
    double ms_prepare;       /* class layout build (gather kernel)             */
    double ms_select;        /* top-k maintenance kernels + final extraction   */
    int64_t evals;           /* row pairs x 5 functions                        */
    int64_t exact_evals;     /* evaluations admitted to the exact score        */
    int32_t bands;           /* count-kernel launches                          */
    int32_t count_path;      /* path actually used                             */
    int32_t n_class0, n_class1;
    int32_t pend_overflows;  /* bands redone: filter survivors overflowed      */
    int32_t cand_overflows;  /* bands redone: admitted candidates overflowed   */
    int64_t bytes_allocated; /* device memory by which this call grew the
                                scan workspace; always 0 for gepi_scan_noalloc */
} gepi_scan_stats;

/* Defaults of src/gEpiCount.cpp:34-39 and :100-115. */
void gepi_params_init(gepi_params *par);

/* Progress of a running scan, for the reference's -D output: its tile loop
 * reports elapsed and remaining time and prints the current best records
 * every so often (src/kernels/kernel1.cu:420-438,
 * src/proc_gold_1.h:140-163).  Here a scan with par->intermediate set calls
 * the calling thread's progress hook after every band of tiles.           */
typedef struct gepi_progress {
    int32_t bands;           /* bands completed                               */
    int32_t reserved;
    int64_t tiles_done, tiles_total;
    double ms_elapsed;       /* since the scan started                        */
    double ms_remaining;     /* estimate: the same time per tile for the rest */
    const gepi_result *tops; /* the best records so far, sorted ascending     */
    int64_t n_tops;          /* min(topmost, records held)                    */
} gepi_progress;
typedef void (*gepi_progress_fn)(const gepi_progress *progress, void *user);
/* Installs (fn != NULL) or removes the calling thread's progress hook; it
 * applies to gepi_scan and gepi_process_gpu_1 calls made by this thread
 * with par->intermediate != 0.  Copying the records to the host costs a
 * few ms per band on large top-k, so it is a diagnostic, not for timing.  */
int gepi_set_progress(gepi_progress_fn fn, void *user);

/* Plane-level scan on device `par->gpu_num`.  Mirrors
 *   void process_gpu_1(PlinkReader<ui8> *pr, InteractionResult *tops,
 *                      const struct _paramP1 &par)   (src/proc_gpu.h:134,
 *                                                     src/kernels/kernel1.cu:508-511)
 * bit_data/valid_data are the reader's BitData()/ValidData() [n_rows][words]
 * with words = ceil(n_samples/64) (src/reader/PlinkReader.h:199-225); rows
 * 2i and 2i+1 are the two alleles of one variant and share valid bits.
 * mask_ctrl/mask_case are classData(1)/classData(2)->mask() [words].
 * tops receives min(par->topmost, 5 * n_rows*(n_rows+1)/2) results sorted by
 * (obj, sA, sB, fun) ascending; their number goes to *n_tops.               */
int gepi_process_gpu_1(const uint64_t *bit_data, const uint64_t *valid_data,
                       const uint64_t *mask_ctrl, const uint64_t *mask_case, int64_t n_rows,
                       int32_t n_samples, const gepi_params *par, gepi_result *tops,
                       int64_t *n_tops, gepi_scan_stats *stats);

/* ---- resident panels (serving / repeated scans) --------------------------
 * A panel keeps the packed allele planes on one device across scans; each
 * scan supplies its own class masks.  No counterpart in the reference, which
 * re-uploads the panel per call (src/kernels/setup1.h:72-102).               */
typedef struct gepi_panel gepi_panel;

/* From reader planes (same layout and pairing rule as gepi_process_gpu_1). */
int gepi_panel_create(const uint64_t *bit_data, const uint64_t *valid_data, int64_t n_rows,
                      int32_t n_samples, int32_t device, gepi_panel **out);
/* From raw SNP-major .bed genotype bytes (the file after its 3-byte header):
 * n_variants blocks of ceil(n_samples/4) bytes, decoded on the device with the
 * rules of src/reader/BinPlinkReader.h:171-201.                              */
int gepi_panel_create_bed(const uint8_t *bed_payload, int64_t n_variants, int32_t n_samples,
                          int32_t device, gepi_panel **out);
int gepi_panel_destroy(gepi_panel *panel);
/* Scan a resident panel with the given class masks ([words] each). */
int gepi_scan(gepi_panel *panel, const uint64_t *mask_ctrl, const uint64_t *mask_case,
              const gepi_params *par, gepi_result *tops, int64_t *n_tops,
              gepi_scan_stats *stats);
/* gepi_scan without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, allocation-free serving loops).  Every buffer a
 * scan needs (score tables, scan state, class layout, filter survivors,
 * candidate lists) lives in one fixed array of GEPI_NOALLOC_SCRATCH_MIB MiB
 * per GPU, held until the process ends.  Fp4 count paths only (count_path
 * 0, 3 or 4); a scan whose buffers do not fit fails with GEPI_ERR_LIMIT
 * (with top-k up to 100,000: about 43,000 variants at 2,048 samples and
 * 12,000 at 8,192 on path 3, 52,000 and 15,000 on path 4).
 * Results equal gepi_scan's.  Calls on one GPU share the array: calls that
 * overlap on one GPU (from several threads) race and must not be made.      */
#define GEPI_NOALLOC_SCRATCH_MIB 64
int gepi_scan_noalloc(gepi_panel *panel, const uint64_t *mask_ctrl, const uint64_t *mask_case,
                      const gepi_params *par, gepi_result *tops, int64_t *n_tops,
                      gepi_scan_stats *stats);
/* Reads the panel's planes back in reader layout (testing aid). */
int gepi_panel_download(gepi_panel *panel, uint64_t *bit_data, uint64_t *valid_data);
/* Missing-call QA on the device: the counts of checkQuality
 * (src/reader/PlinkReader.h:237-335).  variant_missing[n_variants] and
 * sample_missing[n_samples] (either may be NULL) receive missing calls; a
 * variant counts as bad when (float)missing / n_samples > max_miss_snps, a
 * sample when its fraction of missing allele rows > max_miss_samples, the
 * reference's tests (PlinkReader.h:318-319, :281-282).                               */
int gepi_panel_qa(gepi_panel *panel, float max_miss_snps, float max_miss_samples,
                  int32_t *variant_missing, int32_t *sample_missing, int64_t *n_bad_variants,
                  int64_t *n_bad_samples);
/* checkQuality(remove = true), which the reference declares but does not
 * implement (PlinkReader.h:236): *out is a new resident panel holding only
 * the variants within max_miss_snps, in order; kept_variants[n_variants]
 * (may be NULL) receives their original indices and *n_kept their number;
 * keep_samples[ceil(n_samples/64)] (may be NULL) the mask of samples within
 * max_miss_samples, to be ANDed into the class masks of later scans.  Both
 * fractions are taken on the unfiltered panel.                               */
int gepi_panel_filter(gepi_panel *panel, float max_miss_snps, float max_miss_samples,
                      gepi_panel **out, int64_t *kept_variants, int64_t *n_kept,
                      uint64_t *keep_samples);

/* ---- batched scans ---------------------------------------------------------
 * Many scans (panels of any size, phenotype vectors) advance together in
 * rounds of three launches, so small scans share the GPU instead of
 * running one after another.  No counterpart in the reference, which runs
 * one process_gpu_1 per process (src/gEpiCount.cpp:84-87); each item's
 * result is identical to gepi_scan's on the same input (fp4 count path).
 * An engine owns a stream and device workspace on one device; one batch at a
 * time per engine, and an item's panel must not be scanned elsewhere while
 * the batch runs.                                                          */
typedef struct gepi_engine gepi_engine;

typedef struct gepi_batch_item {
    gepi_panel *panel;          /* resident panel on the engine's device        */
    const uint64_t *mask_ctrl;  /* [ceil(n_samples/64)] class 0 (phenotype 1)   */
    const uint64_t *mask_case;  /* class 1 (phenotype 2)                        */
    int64_t topmost;            /* results wanted                               */
    int32_t ffun;               /* GEPI_FFUNC_ENTROPY or GEPI_FFUNC_FISHERPV    */
    int32_t status;             /* out: GEPI_OK or this item's error code       */
    gepi_result *tops;          /* out: [topmost], caller owned                 */
    int64_t n_tops;             /* out                                          */
    gepi_scan_stats stats;      /* out: ms_scan = batch start to item completion */
} gepi_batch_item;

typedef struct gepi_batch_stats {
    double ms_batch;            /* host wall time of the batch                  */
    double ms_prepare;          /* class layouts of all items                   */
    double ms_count_kernels;    /* count launches (HIP events)                  */
    int64_t evals;              /* summed over items                            */
    int32_t rounds;             /* (count, exact, top-k) rounds                 */
    int32_t items_failed;
    int32_t graph_rounds;       /* rounds replayed from captured HIP graphs     */
    int32_t reserved;
} gepi_batch_stats;

int gepi_engine_create(int32_t device, gepi_engine **out);
int gepi_engine_destroy(gepi_engine *engine);

/* Schedule of the batched engine for one size class: scans with at most
 * max_ksteps 128-sample K steps (both classes together; the last class
 * takes every larger scan).  No counterpart in the reference.              */
typedef struct gepi_tune_class {
    int32_t max_ksteps;
    int32_t wg_per_cu;   /* persistent count workgroups per CU (0: one tile
                            per workgroup)                                    */
    int32_t tile;        /* count job shape of the class's entropy scans:
                            0 = a 64 x 32-variant tile (four waves of
                            32 x 16), 1 = half of one, 32 x 32 variants (four
                            waves of 16 x 16), 2 = two tiles as four 32 x 32
                            wave tiles (count path 3's kernel, one unit per
                            workgroup).  Fisher scans run wave tiles in every
                            class (they park flagged pairs); captured rounds
                            (GEPI_ENGINE_GRAPH=1) have no wave-tile launch and
                            run whole tiles in its place.                    */
    int32_t reserved;
    double band_evals;   /* evaluations per scan and round                    */
    double ms_best;      /* autotune trial of the class: chosen schedule      */
    double ms_default;   /* the same trial under the default schedule         */
} gepi_tune_class;

/* Autotuner: for each of 4 size classes (at most 2, 8, 32 and any number of
 * K steps), runs a synthetic batch of 4 scans of that class (40,000 variants
 * x 256 samples, 28,000 x 1,024, 18,000 x 4,096, 12,000 x 8,192; top-100)
 * under every candidate schedule (count job shape 64 x 32 or 32 x 32
 * variants x one tile per workgroup or 1, 2 or 3 persistent workgroups per
 * CU, or wave tiles; x 5e8, 2e9 or 8e9 evaluations per band), repeats times
 * each, and keeps the fastest (the default unless another is more than 3%
 * faster).  The default schedule: 2 persistent workgroups per CU, 2e9
 * evaluations per band, entropy scans of at least 48 K steps on wave tiles
 * and smaller ones on 64 x 32 tiles.  The engine uses the table from then
 * on: a scan's job shape and band size by its class, a round's launch form
 * by the class with the most count work in it.  A table that differs from
 * the default is kept only if it is also 1% faster on all 16 trial scans
 * together.  Results never depend on the schedule.  Writes min(n_out, 4)
 * classes; *n_classes = 4.  On failure the engine is left with the default
 * schedule.                                                                */
int gepi_engine_autotune(gepi_engine *engine, int32_t repeats, gepi_tune_class *out,
                         int32_t n_out, int32_t *n_classes);
/* Installs a table (from an earlier autotune): 1-16 classes, max_ksteps
 * strictly increasing, wg_per_cu in [0, 16], tile 0, 1 or 2, band_evals >= 1e6.
 * n = 0 restores the defaults.                                               */
int gepi_engine_set_tuning(gepi_engine *engine, const gepi_tune_class *classes, int32_t n);
/* Runs every item to completion.  Returns GEPI_OK when the batch ran (an
 * item's own failure is in its status) or the device error that stopped it. */
int gepi_engine_scan_batch(gepi_engine *engine, gepi_batch_item *items, int32_t n_items,
                           gepi_batch_stats *stats);

/* ---- serving loop -----------------------------------------------------------
 * A server owns an engine and a worker thread.  Requests (a resident panel
 * and class masks) are admitted between rounds while the device memory they
 * need fits the budget (continuous batching: a request submitted while
 * others run joins them at the next round).  Panels are resident and keyed
 * by a 128-bit content hash of their genotype bytes: adding the same bytes
 * again returns the same id.  A re-scan with a new phenotype rebuilds only
 * the class layout, on the device.  Each result is identical to gepi_scan's
 * (fp4 path) on the same input.  No counterpart in the reference (one scan
 * per process, src/gEpiCount.cpp:84-87).  All functions are thread safe.   */
typedef struct gepi_server gepi_server;

typedef struct gepi_server_params {
    int32_t device;
    int32_t autotune;        /* nonzero: the server's engine runs the batch
                                autotuner (gepi_engine_autotune, 2 repeats)
                                before the first request                      */
    int64_t hbm_budget;      /* scan workspace admitted at once, bytes; 0 = 3/4
                                of the device memory free at creation; a
                                request larger than the budget runs alone      */
    int64_t round_wg_budget; /* workgroups per count round; 0 = 262,144: bounds
                                how long a new request waits to join          */
} gepi_server_params;

typedef struct gepi_server_stats {
    int64_t submitted, completed, failed;
    int64_t evals;           /* of completed requests                          */
    int64_t rounds;
    int32_t max_active;      /* most requests in one round                     */
    int32_t panels;
    double p50_ms, p99_ms, mean_ms; /* submit -> completion                    */
    double busy_ms;          /* time spent in rounds                           */
} gepi_server_stats;

int gepi_server_create(const gepi_server_params *par, gepi_server **out);
/* Waits for the requests in flight, then releases everything. */
int gepi_server_destroy(gepi_server *server);
/* Resident panel from SNP-major .bed bytes (as gepi_panel_create_bed); the
 * id is the content hash. */
int gepi_server_add_panel_bed(gepi_server *server, const uint8_t *bed_payload, int64_t n_variants,
                              int32_t n_samples, uint64_t *panel_id);
/* Drops one reference; the panel goes when no request uses it. */
int gepi_server_release_panel(gepi_server *server, uint64_t panel_id);
/* Queues a scan; masks are copied.  *ticket identifies it for _wait. */
int gepi_server_submit(gepi_server *server, uint64_t panel_id, const uint64_t *mask_ctrl,
                       const uint64_t *mask_case, int64_t topmost, int32_t ffun,
                       uint64_t *ticket);
/* Reports whether a ticket is done (1) or not (0) without blocking. */
int gepi_server_done(gepi_server *server, uint64_t ticket, int32_t *done);
/* Blocks until the request is done; copies min(capacity, n) results and
 * returns the request's status.  A ticket can be waited for once. */
int gepi_server_wait(gepi_server *server, uint64_t ticket, gepi_result *tops, int64_t capacity,
                     int64_t *n_tops, gepi_scan_stats *stats);
int gepi_server_get_stats(gepi_server *server, gepi_server_stats *out);

/* ---- host-side PLINK reader and report ---------------------------------- */
/* Replaces BinPlinkReader<ui8> (src/reader/BinPlinkReader.h, PlinkReader.h).
 * Fields are read-only for the caller; release with gepi_plink_close.        */
typedef struct gepi_plink {
    int64_t n_rows;       /* numSNPs(): 2 x .bim lines                        */
    int32_t n_samples;    /* numSamples()                                     */
    int32_t words;        /* elementsSNP()                                    */
    uint64_t *bit_data;   /* BitData()   [n_rows][words]                      */
    uint64_t *valid_data; /* ValidData() [n_rows][words]                      */
    uint64_t *mask_ctrl;  /* classData(1)->mask(), NULL if no sample has phenotype 1 */
    uint64_t *mask_case;  /* classData(2)->mask(), NULL if no sample has phenotype 2 */
    int32_t *pheno;       /* per sample phenotype (Sample::Class())           */
    char **sample_name;   /* per sample (Sample::Name(): first .fam token)    */
    char **snp_id;        /* per row (SNP::Id())                              */
    char *allele;         /* per row (SNP::Alleles())                         */
    int32_t *chr, *pos;   /* per row                                          */
    int64_t total_missing;
    uint8_t *bed_payload; /* raw genotype bytes after the header              */
} gepi_plink;

/* checkConsistency() + readData(): 0 or negative gepi_status. */
int gepi_plink_open(const char *root, gepi_plink **out);
void gepi_plink_close(gepi_plink *p);
/* checkQuality(remove=false, ...) (PlinkReader.h:237-335): reports samples and
 * rows above the missing fractions on stderr; counts go to the out args.    */
int gepi_plink_check_quality(const gepi_plink *p, float max_miss_snps, float max_miss_samples,
                             int32_t *n_bad_samples, int32_t *n_bad_rows);
/* printInteractions(pr, tops, numT, stride, context) (src/misc/print_misc.h:56-67). */
int gepi_print_interactions(const gepi_plink *p, const gepi_result *tops, int64_t n,
                            int32_t stride, int32_t context);
/* fisher_pval (src/misc/fisher_p.c:52-76, declared inc/gEpiCount.h:137-139):
 * log10 two-sided Fisher p of the table (n11, n12; n21, n22), as float.      */
float gepi_fisher_pval(int32_t n11, int32_t n12, int32_t n21, int32_t n22);
/* The exact objective of one class x function table, as a result's obj holds
 * it: w, x = function-true samples of class 0 and class 1, y, z = class
 * sizes; ffun GEPI_FFUNC_ENTROPY gives -I(F;C) in bits (the per-evaluation
 * entropy of src/kernels/entropy_inline.h:30-80), GEPI_FFUNC_FISHERPV log10 of
 * the two-sided Fisher p (src/kernels/fisher_d.cu:249-285).  Computed on the
 * host by the code of the scan's exact stage, so it equals the obj of any
 * result with these counts bit for bit.  GEPI_ERR_ARG unless 0 <= w <= y and
 * 0 <= x <= z; GEPI_ERR_LIMIT past GEPI_MAX_SAMPLES.                        */
int gepi_score_table(int32_t w, int32_t x, int32_t y, int32_t z, int32_t ffun, double *score);

/* ---- misc ---------------------------------------------------------------- */
const char *gepi_last_error(void);
const char *gepi_version(void);
/* Number of visible HIP devices (0 when none); status as return value. */
int gepi_device_count(int32_t *count);

#ifdef __cplusplus
}
#endif
#endif /* GEPICOUNT_H */
