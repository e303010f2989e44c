/*
 * codon_nsga3.h -- C ABI of the MI355X NSGA-III codon-design engine.
 *
 * The reference (CAU-HPCL/Protein_encoding_v2_NSGA-III) has no library API:
 * its boundary is the host loop of src/main.cu, which uploads tables and
 * scalars to device symbols (main.cu:666-678), launches one initialisation
 * kernel (main.cu:703) and then two cooperative kernels per generation
 * (main.cu:708-728), and finally copies the pool back (main.cu:772-787).
 * Each entry point below replaces one of those steps for one job whose state
 * lives on the device; the batched entry points advance many such jobs per
 * launch sequence.
 *
 * Conventions: plain pointers and sizes only; the caller owns every host
 * buffer (read during the call, written only through output pointers); the
 * library owns all device memory behind the opaque handles.  Every function
 * returns CD_OK (0) or a negative CD_E_* code, with a message available from
 * cd_last_error().  A cd_job is not thread-safe; distinct jobs may be driven
 * from distinct threads.
 *
 * Genome encoding: design d, CDS c, codon k is the int8 synonym index
 * genome[(d*m + c)*L + k] into the amino acid's synonym block of the codon
 * table (the reference stores ASCII nucleotides, utils.cuh:1072-1077).
 *
 * Objective record (main.cu:642,645,654; common.cuh:23-29), minimised:
 *   obj[6] = { -min CAI, -min CPB, -min HSC, -min HD, max |GC3-ref|, max SL }
 *   idx[12]: [0] CAI CDS, [2] CPB CDS, [4] HSC CDS, [6],[7] HD pair,
 *            [8] GC3 CDS, [9] GC3 direction (+1 up, -1 down), [10] SL CDS
 *   pql[3] = stem (p, q, l): first base of the left arm, innermost base of the
 *            right arm, length.
 */
#ifndef CODON_NSGA3_H
#define CODON_NSGA3_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CD_OBJECTIVES 6

enum {
    CD_OK = 0,
    CD_E_INVALID = -1, /* bad argument (mirrors main.cu:485-489, utils.cuh:84-85) */
    CD_E_NOMEM = -2,   /* device allocation failed */
    CD_E_DEVICE = -3,  /* HIP runtime error */
    CD_E_STATE = -4,   /* call out of order (e.g. step before init) */
    CD_E_BUDGET = -5   /* batch admission would exceed its memory budget */
};

/* Codon table (include/info.cuh:6-54). */
typedef struct cd_tables {
    const char *codons;   /* 192 letters A/C/G/U: codon g = codons[3g..3g+2]; blocks per amino
                             acid (A C D E F G H I K L M N P Q R S T V W Y Z), synonyms by
                             ascending weight, the three stops last */
    const float *weights; /* [64] CAI weights in (0, 1] */
    const float *cps;     /* [61*61] codon-pair scores, cps[a*61 + b] */
    float gc3_ref;        /* ref_GC3_percent (info.cuh:7) */
} cd_tables;

/* One job: main.cu's argv (N, G via cd_job_step, m, Pm), the FASTA protein
 * after findAminoIndex and the appended stop (main.cu:497-562), the
 * reference-point divisions H (main.cu:564-577) and an explicit seed. */
typedef struct cd_job_desc {
    int32_t L;            /* codons including the final stop (>= 3) */
    int32_t m;            /* CDS copies (>= 2) */
    int32_t N;            /* population (> 0); the pool holds 2N designs */
    int32_t H;            /* Das-Dennis divisions (>= 1): C(H+5, 5) reference points */
    float Pm;             /* mutation probability in [0, 1] */
    uint64_t seed;        /* Philox key */
    const int8_t *aa;     /* [L] amino-acid indices 0..20; 20 (stop) exactly at L-1 */
    cd_tables tables;
} cd_job_desc;

typedef struct cd_job cd_job;
typedef struct cd_batch cd_batch;

/* Allocate a job's device state and upload its tables and reference points
 * (replaces main.cu:558-678, including combinationRepetition).  The memory
 * comes from the device's stream-ordered pool, which keeps what
 * cd_job_destroy returns for later jobs. */
int cd_job_create(const cd_job_desc *desc, cd_job **out);

/* Random initial population, design 0 with the best-CAI codons everywhere,
 * evaluated (replaces initializationKernel, main.cu:26-119 and its launch at
 * main.cu:703).  Enqueued on the job's stream. */
int cd_job_init(cd_job *job);

/* Enqueue `generations` generations: mutate + evaluate N children, then
 * NSGA-III selection of N survivors from the 2N pool (replaces the loop body
 * main.cu:708-728: mutationKernel and sortingKernel).  Asynchronous. */
int cd_job_step(cd_job *job, int32_t generations);

/* As cd_job_step, and wait; *ms receives the elapsed time measured with HIP
 * events on the job's stream around the enqueued generations. */
int cd_job_step_timed(cd_job *job, int32_t generations, float *ms);

/* Wait for the job's stream. */
int cd_job_sync(cd_job *job);

/* Generations completed so far. */
int cd_job_generation(cd_job *job, int32_t *generation);

/* Copy the current pool out (replaces main.cu:772-787).  Any pointer may be
 * NULL.  genome [2N][m][L], obj [2N][6], idx [2N][12], pql [2N][3],
 * sorted [N] (survivor pool indices, ascending), front_of [2N] (front index
 * of every ranked design, -1 otherwise), state [54] = extreme points [6][6],
 * ideal [6], worst [6], nadir [6] (sorting.cuh's estimated_* globals). */
int cd_job_read(cd_job *job, int8_t *genome, float *obj, int8_t *idx, int32_t *pql,
                int32_t *sorted, int32_t *front_of, float *state);

/* Run `generations` generations launching every kernel separately and
 * report the mean duration in milliseconds of each kernel, measured with HIP
 * events on the job's stream: kernel_ms[0] offspring (survivor copy,
 * mutation, objectives 0-4), [1] stem scan of the changed CDS, [2] stem
 * objective and dominance matrix, [3] fronts + normalisation,
 * [4] association, [5] niching. */
#define CD_PROFILE_KERNELS 6
int cd_job_profile(cd_job *job, int32_t generations, float *kernel_ms);

/* Number of CDS whose objectives were recomputed since cd_job_init (children
 * reuse the statistics of the CDS their operator left unchanged). */
int cd_job_cds_scanned(cd_job *job, int64_t *count);

int cd_job_destroy(cd_job *job);

/* Evaluate `count` designs [count][m][L] (the six objective reductions of
 * utils.cuh:271-888).  desc->N, H, Pm and seed are ignored. */
int cd_eval_batch(const cd_job_desc *desc, const int8_t *genome, int32_t count,
                  float *obj, int8_t *idx, int32_t *pql);

/* Non-dominated sorting (sorting.cuh:974-1089) of P objective vectors until
 * at least N are ranked: front_of [P] (-1 for unranked), *n_fronts. */
int cd_nds(const float *obj, int32_t P, int32_t N, int32_t *front_of, int32_t *n_fronts);

/* No counterpart in the reference: cd_nds for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops).  The
 * same sort over device buffers -- obj [P][6] in, front_of [P] and
 * n_fronts [1] out -- enqueued on `stream` (a hipStream_t; NULL for the null
 * stream) without waiting.  It calls no allocator: its scratch is a
 * fixed-size device array (about 8.2 MiB) that every GPU the library runs on
 * holds until the process ends, so P is at most CD_NOALLOC_MAX_POOL, and
 * calls that overlap on one GPU (other streams or threads) race on it;
 * calls in one stream are ordered. */
#define CD_NOALLOC_MAX_POOL 8192
int cd_nds_noalloc(const float *obj, int32_t P, int32_t N, int32_t *front_of, int32_t *n_fronts, void *stream);

/* One NSGA-III selection (sortingKernel, main.cu:409-454) of N survivors
 * from 2N objective vectors.  extremes [36] is read and updated;
 * sorted [N]; front_of [2N]; norm [18] = ideal, worst, nadir; flags [2] =
 * (exact cut, hyperplane exception).  gen and seed key the niching draws. */
int cd_select(const float *obj, int32_t N, int32_t H, uint32_t gen, uint64_t seed,
              float *extremes, int32_t *sorted, int32_t *front_of, float *norm, int32_t *flags);

/* Batched form: the listed jobs (already initialised, distinct, with their
 * own tables and sizes) advance together, one generation per launch
 * sequence of persistent kernels over all of them.  While a batch holds a
 * job, cd_job_step, cd_job_profile and cd_job_destroy of that job fail with
 * CD_E_STATE, and cd_job_read, cd_job_sync and cd_job_generation first wait
 * for the batch's stream; cd_batch_destroy releases the jobs. */
int cd_batch_create(cd_job *const *jobs, int32_t count, cd_batch **out);
int cd_batch_step(cd_batch *batch, int32_t generations);
int cd_batch_step_timed(cd_batch *batch, int32_t generations, float *ms);
int cd_batch_sync(cd_batch *batch);
/* As cd_job_profile, for one generation of every job of the batch: the two
 * launch classes of the offspring kernel are timed together as kernel_ms[0]. */
int cd_batch_profile(cd_batch *batch, int32_t generations, float *kernel_ms);
int cd_batch_destroy(cd_batch *batch);

/* Launch shapes of a generation's kernels (no counterpart in the reference,
 * which sizes its cooperative launches from occupancy queries,
 * main.cu:603-613).  Every shape gives the same results; only the time
 * differs. */
typedef struct cd_shapes {
    int32_t design_split;   /* k_design LDS bytes up to which children launch apart (0: one launch) */
    int32_t design_threads; /* k_design workgroup size of the other children (a lone job: all): 256 or 512 */
    int32_t stem_waves;     /* k_stem_list workgroups per CU: 5 to 8 */
    int32_t dom_rows;       /* k_dominance tile rows per workgroup (64 x 64 tiles): 64, 128 or 256 */
    int32_t fronts_threads; /* k_fronts workgroup size: 256, 512 or 1024 */
    int32_t assoc_designs;  /* k_associate designs per workgroup: 4, 8 or 16 (a wave each), 32 or 64 (8 lanes each) */
    int32_t niche_threads;  /* niching workgroup size (k_place_niche, or k_place and k_niche above 2048 designs): 256, 512 or 1024 */
} cd_shapes;

/* Autotuner: a coordinate search over the shapes, each candidate timed
 * (HIP events around each kernel) over `gens_per_candidate` ordinary
 * generations of the job or batch -- which therefore advances by
 * *generations_run generations, with exactly the results it would have had
 * in any shape.  The chosen shapes are installed and cached for the device
 * under a coarse signature of the work (job count, children, codons per
 * child, pool size, in power-of-two buckets); jobs, batches and the
 * server's ticks created later with the same signature start from them. */
int cd_job_tune(cd_job *job, int32_t gens_per_candidate, cd_shapes *chosen, int32_t *generations_run);
int cd_batch_tune(cd_batch *batch, int32_t gens_per_candidate, cd_shapes *chosen, int32_t *generations_run);
/* Read (get) and/or install (set) the shapes of a job or batch; either
 * pointer may be NULL. */
int cd_job_shapes(cd_job *job, cd_shapes *get, const cd_shapes *set);
int cd_batch_shapes(cd_batch *batch, cd_shapes *get, const cd_shapes *set);

/* Served form (no counterpart in the reference, which runs one job per
 * process): continuous batching of jobs of different protein lengths,
 * population sizes and CDS counts.  Requests queue in arrival order; an
 * admitting thread admits the oldest one whenever its device bytes
 * (cd_job_device_bytes) fit the budget beside the jobs holding memory and
 * fewer than max_jobs hold memory, and initialises it on its own stream
 * while the resident jobs keep running; a worker thread advances every
 * resident job one generation per tick (one captured hipGraph per resident
 * set); a collecting thread copies a finished job's final pool to host
 * memory and frees its device memory.  Populations, fronts and generator
 * state stay on the device until the job's generation budget is spent.  A
 * job's result is the same alone and in any mix.  Thread-safe: submit, wait
 * and result may be called from any thread. */
typedef struct cd_server cd_server;
enum { CD_SERVER_QUEUED = 0, CD_SERVER_RUNNING = 1, CD_SERVER_DONE = 2, CD_SERVER_FAILED = 3 };
int cd_server_create(int64_t hbm_budget_bytes, int32_t max_jobs, cd_server **out);
/* Copies the description and its arrays; CD_E_BUDGET if the job alone
 * exceeds the budget.  *ticket identifies the request. */
int cd_server_submit(cd_server *server, const cd_job_desc *desc, int32_t generations, int64_t *ticket);
/* Wait up to timeout_s seconds (negative: no limit) for the request to end;
 * *state = CD_SERVER_*. */
int cd_server_wait(cd_server *server, int64_t ticket, double timeout_s, int32_t *state);
/* Copy out a finished request's final pool (as cd_job_read; any pointer may
 * be NULL) and release it.  times [3] = seconds queued, seconds resident,
 * seconds from submission to result (time to front). */
int cd_server_result(cd_server *server, int64_t ticket, int8_t *genome, float *obj, int8_t *idx, int32_t *pql,
                     int32_t *sorted, int32_t *front_of, double *times);
/* stats [11] = queued requests, resident jobs, device bytes in use, ticks
 * run, designs produced, GPU time of the ticks (microseconds, HIP events
 * around each tick), worker time while jobs were resident (microseconds),
 * resident-set changes (each captures the tick again), time spent on them,
 * worker time launching ticks, worker time waiting for the tick two back
 * (microseconds). */
int cd_server_stats(cd_server *server, int64_t *stats);
/* Stops the worker; unfinished requests are dropped. */
int cd_server_destroy(cd_server *server);

/* Quality indicators on the host (replace the reference's prebuilt `hv`
 * program, main.cu:839-858, and MinEuclid, utils.cuh:42-58).
 * cd_hypervolume: exact hypervolume (minimisation) of n points in d
 * dimensions against ref (WFG algorithm).
 * cd_normalize: main.cu:789-835's normalisation of raw records [n][6].
 * cd_min_distance_spread: minimum distance to the origin and maximum spread
 * of n normalised points [n][6]. */
int cd_hypervolume(const double *pts, int32_t n, int32_t d, const double *ref, double *out);
int cd_normalize(const float *obj, int32_t n, float *out);
int cd_min_distance_spread(const float *pts, int32_t n, float *min_dist, float *spread);

/* Device bytes a job with these sizes holds (for admission by HBM budget). */
int64_t cd_job_device_bytes(int32_t L, int32_t m, int32_t N, int32_t H);

const char *cd_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
