/*
 * blz_prelim.h -- C ABI of the MI355X BLASTP preliminary-search engine.
 *
 * This is the drop-in boundary.  It replaces the hook BLAZE places in NCBI's prelim stage:
 *   int BLAZE_GPU(BlastSeqSrc*, uint32_t threadIndex, uint32_t totalThreads, uint32_t queryLen,
 *                 uint32_t* survivingSeqs, uint32_t* num_survivors,
 *                 int ungapped_cutoff, int gapped_cutoff)
 *   declared BLAZE_gpu/gpu_cpu_common.h:43, defined BLAZE_gpu/gpuBLAZE.cu:1428, called from
 *   Blast_RunPreliminarySearchWithInterruptMT at BLAZE_ncbi/blast_engine.c:2983.
 * BLAZE's hook only filters subjects and lets the CPU re-run the prelim core on survivors
 * (blast_engine.c:2988-2996).  This ABI returns the per-subject HSP lists themselves, the
 * lists that core writes to the HSP stream (blast_engine.c:2499), so the caller can hand them
 * to the unchanged traceback stage.
 *
 * Everything is explicit: no side channel through BlastSeqSrc (blast_seqsrc.h:341-451), no
 * __constant__ per-query state (gpuBLAZE.cu:81-82), so several queries can be in flight.
 * Errors are returned as status codes; nothing here calls exit().
 * All pointers are host pointers unless a name says otherwise.  Plain C types only.
 */
#ifndef BLZ_PRELIM_H
#define BLZ_PRELIM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    BLZ_OK = 0,
    BLZ_EINVAL = 1,     /* bad argument */
    BLZ_ENOMEM = 2,     /* host or device allocation failed */
    BLZ_EDEVICE = 3,    /* HIP runtime error or no usable gfx950 device */
    BLZ_ECAPACITY = 4,  /* caller's sink too small: counts are filled, *total says how many */
    BLZ_EOVERFLOW = 5,  /* some (query, subject) pairs were handed back unprocessed (blz_overflow_drain) */
    BLZ_ETIMEOUT = 6    /* blz_server_wait: the request is not done yet */
} blz_status;

/* Search parameters.  The reference passes the two cutoffs at run time
 * (gpu_cpu_common.h:43; computed at blast_engine.c:2971-2979) and hard-codes the rest
 * (gpuBLAZE.cu:37,59-72; prelim_stage.cpp:144). */
typedef struct {
    int32_t word_threshold;  /* neighbourhood threshold T, default 11 */
    int32_t window;          /* two-hit window, default 40 */
    int32_t x_ungapped;      /* raw ungapped X-drop, default 16 */
    int32_t x_gapped;        /* raw prelim gapped X-drop, default 38 */
    int32_t gap_open;        /* default 11 */
    int32_t gap_extend;      /* default 1 */
    int32_t ungapped_cutoff; /* init-hit cutoff (raw), default 41 */
    int32_t gapped_cutoff;   /* HSP cutoff (raw), default 67 */
    int32_t reap_score;      /* prelim e-value reap as a score: HSPs scoring below it are dropped
                                after the purge and sort (blast_engine.c:1436-1457,896-930);
                                0 keeps every HSP.  blz_search_setup computes it. */
} blz_params;

/* One HSP.  0-based offsets, *_end one past the last aligned residue; q_gs/s_gs is the
 * gapped start the score-only DP was anchored at (NCBI BlastHSP query/subject
 * offset, end, gapped_start; raw score). */
typedef struct {
    int32_t score;
    int32_t q_off, q_end;
    int32_t s_off, s_end;
    int32_t q_gs, s_gs;
} blz_hsp;

/* Per-run counters and device timings (HIP events on the run's stream). */
typedef struct {
    uint64_t subjects;         /* subjects scanned */
    uint64_t residues;         /* subject residues scanned */
    uint64_t word_hits;        /* lookup hits */
    uint64_t ungapped_ext;     /* two-hit ungapped extensions */
    uint64_t init_hits;        /* ungapped HSPs at or above the ungapped cutoff */
    uint64_t gapped_subjects;  /* subjects that reached the gapped stage */
    uint64_t gapped_ext;       /* semi-gapped extensions (left+right counted once) */
    uint64_t dp_cells;         /* DP cells evaluated inside the X-drop bands */
    uint64_t hsps;             /* HSPs returned */
    uint64_t overflow_subjects;/* (query, subject) pairs past a per-subject cap of the main
                                  passes: run by the overflow queue, or handed back */
    double ms_seed;            /* seed/two-hit/ungapped kernel time */
    double ms_gapped;          /* gapped-stage kernel time */
    double ms_total;           /* device time from first launch to results on device */
    int32_t seed_launches;     /* kernel launches of the seed stage */
    int32_t gapped_launches;
    uint64_t overflow_left;    /* pairs handed back (blz_overflow_drain); 0 unless BLZ_EOVERFLOW */
    int32_t reruns;            /* runs repeated because a work buffer filled up */
    int32_t drain_rounds;      /* overflow-queue passes on the device */
    double ms_drain;           /* wall time of the overflow queue's passes */
} blz_run_stats;

/* Result sink (caller-owned).  Subjects with no HSP are omitted, as NCBI skips the stream
 * write for an empty list (blast_engine.c:812-815, "hspcnt == 0 -> continue").
 * On return: n_subjects entries of oids/counts (ascending oid) and their HSPs subject-major in
 * score order.  If a capacity is short, BLZ_ECAPACITY is returned and the totals say how
 * much room is needed. */
typedef struct {
    int64_t *oids;          /* [subject_cap] */
    int32_t *counts;        /* [subject_cap] */
    int64_t subject_cap;
    blz_hsp *hsps;          /* [hsp_cap] */
    int64_t hsp_cap;
    int64_t n_subjects;     /* out */
    int64_t n_hsps;         /* out */
} blz_hsp_sink;

typedef struct blz_ctx blz_ctx;     /* device, stream, reusable device work buffers */
typedef struct blz_db blz_db;       /* resident database (residues + offsets in HBM) */
typedef struct blz_query blz_query; /* prepared query: lookup tables in HBM */

/* Context: binds one HIP device (gfx950) and owns a stream plus work buffers that are grown
 * on demand and reused.  Replaces the per-call pinned/device rings BLAZE_GPU allocates and
 * never frees (gpuBLAZE.cu:1528-1555). */
blz_status blz_ctx_create(int32_t device, blz_ctx **out);
void blz_ctx_destroy(blz_ctx *ctx);
const char *blz_last_error(const blz_ctx *ctx);
/* The hipStream_t all work of this context runs on (for callers that time or order work). */
void *blz_ctx_stream(blz_ctx *ctx);

void blz_default_params(blz_params *p);

/* Launch shape and variant choices of a context (picked by the autotuner in blz_tune.py; no
 * reference counterpart: the reference hard-codes 64 threads per subject, gpuBLAZE.cu:1639).
 * Zero fields keep the built-in choice. */
typedef struct {
    int32_t seed_variant;      /* 1: global tables, in-place extension; 2: LDS tables, batched */
    int32_t seed_waves;        /* waves per workgroup of the seed kernel */
    int32_t seed_blocks_per_cu;/* resident workgroups per CU the seed launch is sized for */
    int32_t gapped_blocks;     /* workgroups of the gapped kernel */
    int32_t gapped_threads;    /* threads per workgroup of the wide gapped kernel */
    int32_t seed_table;        /* lookup tables: 1 global, 2 entries in LDS, 3 all in LDS (0 auto) */
    int32_t gapped_wave_classes;/* cost classes 0..k-2 run one wave per subject, the rest one lane;
                                  k = 1..9 (0: all classes on waves) */
    int32_t query_packing;     /* batches: 1 scans each query on its own, else short queries share scans */
    int32_t pack_len;          /* letters of a query pack (concatenation); 0: 1024 - 128 - window */
    int32_t bin_ring;          /* query-size bins of a batch: diagonal-ring lengths up to bin_ring,
                                  2x, 4x, beyond (one seed launch per bin and table placement); 0: 1152 */
    int32_t graph;             /* 1 launches on the stream every time; else a run's launch sequence
                                  is captured in hipGraphs and replayed while the same queries and
                                  database come again */
    int32_t gapped_lanes;      /* lanes per subject in the wave-parallel gapped DP: 16 (four
                                  subjects per wave) or 64; 0: 16 */
    int32_t subject_groups;    /* 1: a wave scans one subject at a time; else short subjects of a
                                  single-query scan share a wave's scan (up to 8, 1,024 letters) */
    int32_t overflow_mode;     /* 0: (query, subject) pairs past a per-subject device cap go to the
                                  overflow queue, which runs them again on the device with buffers
                                  sized to them, and a run that filled one of its work buffers is
                                  run again with larger ones; 1: such pairs are handed back
                                  (BLZ_EOVERFLOW, blz_overflow_drain) and the buffers keep their
                                  size (only the overflow list grows, so every pair is reported) */
    /* Capacity limits, for tests of the overflow paths (0: built-in size) */
    int32_t cap_stage;         /* init hits per subject staged by the seed pass (<= 4096) */
    int32_t cap_retry;         /* init hits per pair staged by the retry pass (<= 65536) */
    int32_t cap_ovf;           /* overflow-list entries */
    int32_t cap_wide_ring;     /* DP band ring of the wide gapped pass (power of two, <= 4096) */
    int32_t cap_wide_saved;    /* HSPs per subject before the purge in the wide gapped pass */
    int64_t cap_work;          /* (query, subject) work items of a run */
    int64_t cap_pool;          /* init hits of a run */
    int64_t cap_hsps;          /* HSPs of a run */
    int64_t cap_tb_bytes;      /* traceback main pass: traceback bytes per wave (pairs needing
                                  more go to the wide pass) */
    int64_t cap_tb_runs;       /* traceback: edit-script runs a call starts with */
    int32_t tb_lanes;          /* traceback main pass: lanes per pair, 64 (one pair per wave) or 16
                                  (four pairs per wave, slower: 25.3 against 20.4 ms on the
                                  traceback workload); 0: 64 */
    int32_t gapped_width;      /* bits per DP score in the 16-lane gapped pass: 32 (one column per
                                  lane) or 16 (two columns per lane, packed int16 arithmetic; DPs
                                  past its bounds go to the wide pass; slower: 3.12 against 2.77 ms
                                  on the scan workload); 0: 32 */
} blz_tuning;

blz_status blz_ctx_set_tuning(blz_ctx *ctx, const blz_tuning *t);
blz_status blz_ctx_get_tuning(const blz_ctx *ctx, blz_tuning *t);

/* Upload a database once; it stays resident.  residues: concatenated NCBIstdaa letters;
 * offsets[nsubj+1]: subject i is residues[offsets[i] .. offsets[i+1]), at most 16,777,215
 * letters long (BLZ_EINVAL otherwise).  Replaces the host
 * gather and per-chunk H2D copies of BLAZE_GPU (gpuBLAZE.cu:1587-1618) and the DB arrays
 * BLAZE allocates in prelim_stage.cpp:1204-1215. */
blz_status blz_db_upload(blz_ctx *ctx, const uint8_t *residues, const uint64_t *offsets,
                         int64_t nsubj, blz_db **out);
void blz_db_free(blz_db *db);
int64_t blz_db_num_subjects(const blz_db *db);
uint64_t blz_db_num_residues(const blz_db *db);

/* Build a query's neighbourhood lookup table on the host and upload it; qlen 1..16,000 (the
 * seed kernel keeps a query's diagonal state, qlen + window + 128 entries, in LDS).  Replaces the
 * query preprocessing of prelim_stage.cpp:340-571 and copyConstantMemory (gpuBLAZE.cu:1823-1842). */
blz_status blz_query_prepare(blz_ctx *ctx, const uint8_t *query, int32_t qlen,
                             const blz_params *params, blz_query **out);
void blz_query_free(blz_query *q);

/* Prelim stage of one query over the whole database, one launch per stage.
 * Replaces BLAZE_GPU + BLAST_PreliminarySearchEngineGPU (blast_engine.c:2983-2996).
 * Queries and databases may come from any context on the same device. */
blz_status blz_prelim_run(blz_ctx *ctx, const blz_query *q, const blz_db *db,
                          blz_hsp_sink *sink, blz_run_stats *stats);

/* Prelim stage of nq queries over the database in one persistent work-stealing launch per
 * query-size bin and stage; short queries share scans (packs).  sinks[i] receives query i's
 * lists; sinks may be NULL, then the lists are fetched afterwards with blz_result_*.  stats
 * (optional) sums over the batch.  Unless tuning.graph is 1 the launch sequence of a run (this
 * one and blz_prelim_run) is captured into two hipGraphs, one per stage, and replayed while the
 * same queries and database come again. */
blz_status blz_prelim_run_batch(blz_ctx *ctx, const blz_query *const *queries, int32_t nq,
                                const blz_db *db, blz_hsp_sink *sinks, blz_run_stats *stats);

/* blz_prelim_run without any allocation.  No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  Every device work buffer lies in one fixed 64 MiB array that the
 * library holds on each GPU it runs on until the process ends; the call makes no hipMalloc,
 * hipFree, pinned or host-heap allocation.  Databases of more than 131,072 subjects are rejected
 * (BLZ_EINVAL).  The lists go straight into the caller's sink (BLZ_ECAPACITY, sizes reported, if
 * it is too small) and are not kept for blz_result_*.  Subjects past its per-subject caps (256
 * init hits, 1,024 HSPs) or its run caps (524,288 init hits, 262,144 HSPs) have no list: the call
 * returns BLZ_EOVERFLOW with stats->overflow_left set, and those need blz_prelim_run.  Every call
 * on one GPU shares that GPU's array: calls that overlap on one GPU (from several threads or
 * contexts) race and must be serialised by the caller. */
blz_status blz_prelim_run_noalloc(blz_ctx *ctx, const blz_query *q, const blz_db *db, blz_hsp_sink *sink,
                                  blz_run_stats *stats);

/* Lists of query qi (its index in the last run's batch; 0 after blz_prelim_run) kept by the
 * context until its next run: their sizes, and a copy into a caller-sized sink (BLZ_ECAPACITY,
 * sizes still reported, if it is too small).  No reference counterpart: the reference hands
 * back survivor OIDs only (gpuBLAZE.cu:1673-1691). */
blz_status blz_result_sizes(const blz_ctx *ctx, int32_t qi, int64_t *n_subjects, int64_t *n_hsps);
blz_status blz_result_copy(const blz_ctx *ctx, int32_t qi, blz_hsp_sink *sink);

/* (query index << 40 | subject) of the pairs the last run handed back (it returned
 * BLZ_EOVERFLOW: tuning.overflow_mode 1, or the overflow queue could not size its buffers);
 * *n receives their number.  By default the overflow queue finishes such pairs on the device
 * itself, where BLAZE flags capped subjects for re-evaluation on the CPU (README.md:36-37,
 * flag 255 at gpuBLAZE.cu:1088,1243-1247). */
blz_status blz_overflow_drain(const blz_ctx *ctx, int64_t *pairs, int64_t cap, int64_t *n);

/* ---- search statistics ---------------------------------------------------------------
 * What NCBI computes before the prelim stage and after it, for the lists this engine returns:
 * the cutoffs the reference passes to its hook (blast_engine.c:2971-2981; the ungapped X-drop 16,
 * gapped X-drop 38 and gap trigger 41 it hard-codes at gpuBLAZE.cu:37,59-63 are 7, 15 and 22 bits),
 * the e-values (Blast_HSPListGetEvalues, blast_engine.c:1441), the prelim e-value reap
 * (s_Blast_HSPListReapByPrelimEvalue, blast_engine.c:896-930,1457) and the per-query hit list
 * the HSP stream hands to traceback (prelim_stage.cpp:1462-1498). */
typedef struct {
    double lambda, K, logK, H;
} blz_karlin;

typedef struct {
    double expect;              /* e-value threshold of the search, default 10 */
    double prelim_evalue;       /* reap threshold after the prelim stage; 0: expect */
    double x_ungapped_bits;     /* default 7 */
    double x_gapped_bits;       /* default 15 (prelim gapped X-drop) */
    double x_final_bits;        /* default 25 (traceback X-drop) */
    double gap_trigger_bits;    /* default 22 */
    int32_t gap_open, gap_extend;   /* default 11, 1 (BLOSUM62 costs with published statistics) */
    int32_t word_threshold, window; /* default 11, 40 */
    int32_t query_composition;  /* 1: ungapped statistics from the query's composition against the
                                   background; 0: background on both sides */
    int32_t max_target_seqs;    /* default 500 */
    int32_t composition_based_stats; /* default 2; nonzero sizes the prelim hit list 2x + 50 */
    int32_t reserved;
} blz_stats_options;

typedef struct {
    blz_karlin ungapped, gapped;
    double alpha, beta;         /* length-adjustment constants of the gapped statistics */
    int32_t length_adjustment;
    int32_t gap_trigger;        /* raw */
    int64_t eff_query_len, eff_db_len;
    double eff_searchsp;
    double prelim_evalue;
    int32_t reap_score;         /* lowest score whose e-value is within prelim_evalue */
    int32_t x_final;            /* raw traceback X-drop */
    int32_t hitlist_size;       /* prelim hit-list size */
    int32_t reserved;
} blz_search_stats;

void blz_stats_default_options(blz_stats_options *o);
/* Karlin-Altschul parameters of BLOSUM62 for letter frequencies freq[28] (NULL: the
 * Robinson-Robinson background) against the background. */
blz_status blz_karlin_ungapped(const double *freq, blz_karlin *out);
/* Published gapped parameters of BLOSUM62 for the gap costs (BLZ_EINVAL if none are published). */
blz_status blz_karlin_gapped(int32_t gap_open, int32_t gap_extend, blz_karlin *out, double *alpha, double *beta);
/* Statistics of one query against a database of db_residues letters in db_seqs sequences, and
 * the search parameters they imply (params may be NULL). */
blz_status blz_search_setup(const blz_stats_options *o, const uint8_t *query, int32_t qlen, uint64_t db_residues,
                            int64_t db_seqs, blz_params *params, blz_search_stats *stats);
/* E-value of a raw gapped score. */
double blz_evalue(const blz_search_stats *s, int32_t score);
/* The hit list of query qi of the last run (NCBI Blast_HSPResultsFromHSPStreamWithLimitEx and
 * Blast_HSPResultsSortByEvalue): at most s->hitlist_size subjects, the best by best e-value
 * (e-values below 1e-180 equal), then best score, then higher ordinal id first (for a database
 * uploaded length-sorted by blz_db_upload_blastdb, the id it has in the BLAST database), in that order;
 * each subject's HSPs keep their score order.  best_evalue[subject_cap] (optional) receives each
 * listed subject's best e-value.  BLZ_ECAPACITY (sizes reported) if the sink is too small. */
blz_status blz_result_hitlist(const blz_ctx *ctx, int32_t qi, const blz_search_stats *s, blz_hsp_sink *sink,
                              double *best_evalue);

/* ---- traceback stage ---------------------------------------------------------------------
 * The gapped alignment with traceback of the prelim HSPs of listed (query, subject) pairs,
 * normally each query's hit list (blz_result_hitlist): NCBI Blast_TracebackFromHSPList for blastp
 * without composition-based rescoring (NCBI's -comp_based_stats 0), the stage downstream of the
 * lists the reference's prelim stage writes (blast_engine.c:2499) and which the reference leaves
 * to NCBI's CPU code.  Per pair, the HSPs in their score order: one contained in an HSP already
 * traced (score no higher, both end points inside it) is skipped; otherwise ALIGN_EX runs left
 * and right from its gapped start with the final X-drop and the path is read back.  HSPs scoring
 * at least the cutoff are kept, then the common-endpoint purge and the score sort. */
enum { BLZ_OP_SUB = 0, BLZ_OP_GAP_Q = 1, BLZ_OP_GAP_S = 2 };  /* aligned pair; subject letter
                                      against a gap in the query; query letter against a gap */

/* One traced HSP: the blz_hsp fields (score and extent of the final alignment, the gapped start
 * it was anchored at), identical pairs, alignment columns, gap letters, gap openings, and its
 * edit script, runs[op_first, op_first + op_count) left to right, each run op << 28 | length. */
typedef struct {
    int32_t score;
    int32_t q_off, q_end;
    int32_t s_off, s_end;
    int32_t q_gs, s_gs;
    int32_t num_ident, align_len, gaps, gap_opens;
    int32_t op_first, op_count;
} blz_tb_hsp;

/* A pair to trace: query index into the call's query array, subject oid, and how many of the
 * call's prelim HSPs (consecutive, pair after pair, each pair's in score order) are its. */
typedef struct {
    int32_t query;
    int32_t n_hsps;
    int64_t oid;
} blz_tb_pair;

typedef struct {
    int32_t x_dropoff;   /* raw final X-drop, default 64 (25 bits; blz_search_stats.x_final) */
    int32_t gap_open, gap_extend;  /* default 11, 1 */
    int32_t cutoff;      /* HSPs scoring below it are dropped (0 keeps all; for the e-value
                            threshold E, the lowest score whose blz_evalue is within E) */
} blz_tb_params;

typedef struct {
    uint64_t pairs, prelim_hsps;
    uint64_t hsps;        /* traced HSPs returned */
    uint64_t alignments;  /* one-direction ALIGN_EX runs */
    uint64_t dp_cells;    /* cells inside the X-drop bands */
    uint64_t runs;        /* edit-script runs returned */
    uint64_t wide_pairs;  /* pairs whose band or storage outgrew the main pass */
    int32_t reruns;       /* calls repeated because the edit-script buffer filled */
    int32_t reserved;
    double ms_main, ms_wide, ms_total;  /* device time of the passes; wall time of the call */
} blz_tb_stats;

void blz_tb_default_params(blz_tb_params *p);
/* Trace n_pairs pairs against the resident database; queries[p.query] are prepared queries of
 * this device, query_cutoffs[nq] (optional) their score cutoffs in place of p->cutoff (NCBI's
 * cutoff is per query: its e-value threshold over the query's search space).  The result stays in
 * the context until its next traceback call (blz_traceback_sizes / blz_traceback_copy).
 * BLZ_EINVAL for a pair or HSP outside its query or subject; BLZ_EOVERFLOW if a pair's alignment
 * outgrows even the wide pass (2^31 traceback cells in one direction). */
blz_status blz_traceback(blz_ctx *ctx, const blz_db *db, const blz_query *const *queries, int32_t nq,
                         const int32_t *query_cutoffs, const blz_tb_pair *pairs, int64_t n_pairs,
                         const blz_hsp *hsps, const blz_tb_params *p, blz_tb_stats *stats);
blz_status blz_traceback_sizes(const blz_ctx *ctx, int64_t *n_pairs, int64_t *n_hsps, int64_t *n_runs);
/* counts[n_pairs]: traced HSPs per pair; hsps: pair after pair, score order; runs: the edit
 * scripts (op_first indexes runs).  BLZ_ECAPACITY if a capacity is short. */
blz_status blz_traceback_copy(const blz_ctx *ctx, int32_t *counts, int64_t n_pairs, blz_tb_hsp *hsps,
                              int64_t hsp_cap, uint32_t *runs, int64_t runs_cap);

/* ---- serving: a request loop over a resident database ------------------------------
 * No reference counterpart (the reference is a per-search hook, blast_engine.c:2983-2996).
 * Requests are submitted from any thread.  A preparation thread builds their lookup tables;
 * whenever one of `slots` contexts (each its own stream and buffers) is free, the oldest waiting
 * request and the same-size-class requests behind it run there as one batch
 * (blz_prelim_run_batch), admitted by footprint: up to max_batch queries, max_batch_letters query
 * letters and max_batch_work estimated work items, the estimate per request being its letters x
 * the database's subjects x a rate learned from the batches already run.  A slot starts its batch
 * only while the batches in flight are estimated to need at most overlap_ms more device time (their
 * query letters x the database's residues at a learned rate), so where one batch fills the GPU the
 * batches run in turn and requests arriving meanwhile join the next one; where batches are short,
 * those of different slots overlap on the device.  A query longer than a pack member (half of
 * 1024 - 128 - window letters) whose scan is estimated above 20 ms runs alone: it would share no
 * scan and only hold its batch-mates back.  A request's lists are exactly those blz_prelim_run
 * gives for it alone. */
typedef struct blz_server blz_server;

typedef struct {
    int32_t slots;             /* batches in flight; 0: 2 (at most 16) */
    int32_t max_batch;         /* queries per batch; 0: 64 (1: one query per launch) */
    int64_t max_batch_letters; /* query letters per batch; 0: 32768 */
    int32_t linger_us;         /* a batch with room for more requests waits up to this long after
                                  its oldest request became ready, for others to join; 0: none */
    int64_t max_batch_work;    /* device footprint of a batch, as estimated (query, subject) work
                                  items (each with its init hits, order entry and result: about
                                  100 bytes); 0: what a slot's buffers hold at start (48 Mi items),
                                  within its share of the free HBM.  A request alone always runs. */
    double overlap_ms;         /* a free slot starts a batch only while the batches in flight are
                                  estimated to need at most this much more device time, so batches
                                  that fill the GPU alone (large databases) run one after another and
                                  only overlap their ends; 0: 5 ms; negative: no limit */
} blz_server_config;

typedef struct {
    blz_status status;
    int32_t batch_queries;     /* size of the batch the request ran in */
    double latency_us;         /* submit -> lists on the host */
    double queue_us;           /* submit -> its batch started */
    int64_t n_subjects, n_hsps;
    int64_t n_overflow;        /* subjects handed back for this request (blz_server_overflow) */
} blz_request_info;

typedef struct {
    uint64_t submitted, completed, batches;
    uint64_t residues;         /* database residues scanned, summed over requests */
    double device_ms;          /* device time of the batches, summed over slots */
    uint64_t reruns;           /* batches run again with grown buffers (a footprint estimate fell short) */
    double work_rate;          /* work items per query letter and subject the admission assumes now */
    double pair_rate;          /* query letters x database residues per second it assumes now */
} blz_server_stats;

void blz_server_default_config(blz_server_config *cfg);
/* The database (uploaded with any context on `device`) must outlive the server.  tuning
 * (optional) is applied to every slot. */
blz_status blz_server_create(int32_t device, const blz_db *db, const blz_server_config *cfg,
                             const blz_tuning *tuning, blz_server **out);
void blz_server_destroy(blz_server *s);
blz_status blz_server_submit(blz_server *s, const uint8_t *query, int32_t qlen, const blz_params *params,
                             uint64_t *ticket);
/* Blocks until the request is done (timeout_ms < 0: no limit; BLZ_ETIMEOUT if it ran out). */
blz_status blz_server_wait(blz_server *s, uint64_t ticket, int32_t timeout_ms, blz_request_info *info);
blz_status blz_server_copy(blz_server *s, uint64_t ticket, blz_hsp_sink *sink);
/* Subjects of a finished request whose lists its batch handed back (status BLZ_EOVERFLOW; only
 * with tuning.overflow_mode 1 or when the overflow queue could not size its buffers). */
blz_status blz_server_overflow(blz_server *s, uint64_t ticket, int64_t *subjects, int64_t cap, int64_t *n);
blz_status blz_server_release(blz_server *s, uint64_t ticket);
blz_status blz_server_get_stats(blz_server *s, blz_server_stats *out);

/* ---- BLAST protein databases (makeblastdb format versions 4 and 5) -------------------
 * Replaces the database access of the hook (CSeqDB, seqdbimpl.cpp:716-730; gathered per chunk at
 * gpuBLAZE.cu:1587-1618).  A database name is a volume's base path (<name>.pin, .psq and .phr,
 * memory-mapped) or an alias file <name>.pal whose DBLIST line names the volumes.  The reference
 * requires the FASTA sorted by length before makeblastdb (README.md:138); here gathering can sort
 * by length instead, giving the OID of each position. */
typedef struct blz_blastdb blz_blastdb;
blz_status blz_blastdb_open(const char *path, blz_blastdb **out);
void blz_blastdb_close(blz_blastdb *db);
/* Message of the last failed blz_blastdb_* call on this thread. */
const char *blz_blastdb_last_error(void);
int64_t blz_blastdb_num_seqs(const blz_blastdb *db);
uint64_t blz_blastdb_num_residues(const blz_blastdb *db);
int32_t blz_blastdb_max_length(const blz_blastdb *db);
const char *blz_blastdb_title(const blz_blastdb *db);
/* Letters (NCBIstdaa, inside the mapped file) of one sequence. */
blz_status blz_blastdb_sequence(const blz_blastdb *db, int64_t oid, const uint8_t **letters, int32_t *len);
/* Header bytes (ASN.1 BER Blast-def-line-set, as stored) of one sequence. */
blz_status blz_blastdb_header(const blz_blastdb *db, int64_t oid, const uint8_t **bytes, int64_t *len);
/* All sequences into residues[num_residues] and offsets[num_seqs + 1]: in OID order, or with
 * sort_by_length 1 in ascending length (OID order within a length); perm[num_seqs] (optional)
 * receives the OID at each position. */
blz_status blz_blastdb_gather(const blz_blastdb *db, int32_t sort_by_length, uint8_t *residues, uint64_t *offsets,
                              int64_t *perm);
/* blz_blastdb_gather then blz_db_upload: the database resident in HBM. */
blz_status blz_db_upload_blastdb(blz_ctx *ctx, const blz_blastdb *bdb, int32_t sort_by_length, blz_db **out,
                                 int64_t *perm);

/* ---- synthetic data (NCBIstdaa, Robinson-Robinson background) ----------------------- */

typedef struct {
    uint64_t seed;
    int64_t n_subjects;
    double len_median;      /* log-normal median */
    double len_sigma;       /* log-normal sigma */
    int32_t len_min, len_max;
    double homolog_frac;    /* fraction of subjects planted as homologs */
    double ident_min, ident_max; /* identity range of planted homologs */
    double indel_rate;      /* per-residue indel start rate */
    double indel_mean;      /* geometric indel length mean */
    double repeat_frac;     /* fraction of residues in low-complexity repeat blocks */
    int32_t repeat_period_min, repeat_period_max;
    int32_t sort_by_length; /* 1: subjects sorted by length (ascending) */
} blz_synth_spec;

void blz_synth_default(blz_synth_spec *spec);
/* Random query of length qlen drawn from the background. */
void blz_synth_query(uint64_t seed, int32_t qlen, uint8_t *out);
/* Sizes of the database a spec describes (deterministic). */
blz_status blz_synth_db_size(const blz_synth_spec *spec, int64_t *n_subjects,
                             uint64_t *n_residues);
/* Fill a database; homologs are planted from the given queries (round robin).
 * homolog_of[i] (optional) receives the query index subject i derives from, or -1. */
blz_status blz_synth_db_fill(const blz_synth_spec *spec, const uint8_t *const *queries,
                             const int32_t *qlens, int32_t nq, uint8_t *residues,
                             uint64_t *offsets, int32_t *homolog_of);

#ifdef __cplusplus
}
#endif
#endif
