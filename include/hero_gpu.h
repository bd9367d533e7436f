/*
 * hero_gpu.h — C ABI of the MI355X (gfx950) implementation of HERO's read-correction hot path:
 * k-mer seed index and lookup, banded global / affine DP, alignment records, votes, consensus.
 *
 * Plain pointers and sizes only; all device memory is owned by the context. Every function
 * returns HERO_OK (0) or a negative status; hero_last_error() describes the last failure.
 * Calls on one context must not overlap (one context per host thread).
 *
 * Each entry point names the reference interface it replaces (guo-xuan/Hero, src/).
 */
#ifndef HERO_GPU_H
#define HERO_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HERO_OK 0
#define HERO_ERR_ARG (-1)        /* invalid argument / unsupported input */
#define HERO_ERR_HIP (-2)        /* HIP runtime or kernel launch failure */
#define HERO_ERR_NOMEM (-3)      /* device allocation failed / admission refused */
#define HERO_ERR_STATE (-4)      /* call out of order (e.g. align before the index exists) */
#define HERO_ERR_LIMIT (-5)      /* input beyond a documented limit */
#define HERO_ERR_REF_ABORT (-6)  /* input on which the reference aborts (see hero_last_error) */

typedef struct hero_ctx hero_ctx;

/* Run options. The first block mirrors the reference's Config fields read on the hot path
 * (src/Config.cpp:17-56 defaults, flags parsed at :118-243) and BandAligner::setCost
 * (src/BandAligner.cpp:528-542); hero_default_options() fills the reference defaults. */
typedef struct {
	int32_t k;              /* Config::iKmer (-k), 1..64 */
	int32_t aligner;        /* Config::iAlignerType (-a): 0 affine, 1 global */
	int32_t call_consensus; /* Config::bCallConsensus (-c) */
	int32_t hash_factor;    /* Config::iHashTableFactor (-f) */
	int64_t min_overlap;    /* Config::minimumOverlapLength (-l) */
	double mismatch_rate;   /* Config::dMaxMismatchRate (-m / 100) */
	double indel_rate;      /* Config::dMaxIndelRate (-i / 100) */
	double consensus_freq;  /* Config::dMinFrequencyCallConsensus (-c <f>) */
	int32_t dx, dy, ex, sx, miss, hit, unobserved; /* BandAligner::setCost */
	/* knobs the reference does not have (0 = reference behaviour) */
	int32_t band_override;  /* >0: setBand(band_override) instead of setBand(2*extra) */
	int32_t max_candidates; /* >0: align only the first N query reads per subject strand */
	/* execution */
	int32_t device;         /* HIP device ordinal */
	int32_t dp_form;        /* 0 batched (lane per pair, length buckets), 1 plain (wavefront per pair) */
	int64_t chunk_subjects; /* subject reads per device pass; 0 = automatic */
	int32_t dp_waves;       /* persistent DP wavefronts; 0 = automatic */
	int32_t records;        /* 1: keep the alignment records as .aln text (requires call_consensus = 0) */
	int32_t dp_per_lane;    /* batched form: 1 one task per lane (32-bit), 2 two tasks per lane (packed
	                           16-bit, band classes <= 12); 0 = the hero_autotune choice, else 1 */
	int32_t dp_eager;       /* 1: launch the DP stage kernel by kernel instead of as one captured graph */
	int32_t chunk_workers;  /* subject chunks in flight at once (own streams and buffers; votes only,
	                           record mode and autotuning run one); 0 = 2 */
	int32_t dp_untagged;    /* 1: one-task-per-lane affine rows in the general form; 0: the tagged form
	                           (pointer in the scores' low bits) where dx == dy and the band is <= 12 */
	int32_t mapping;        /* Config::bMappingVersion (--mapping): with call_consensus, per subject only
	                           the best record covering it votes (SubjectRead::chooseQueryWithBestAlignment);
	                           the index is sized by bases / k; read the result with
	                           hero_coverage_histogram instead of calling consensus */
	int32_t band_align_form; /* hero_band_align_batch: 0 the plain form (one wavefront per call); for tests
	                            of the pipeline's kernels on given calls (A/C/G/T, bands <= 32, < 512 rows):
	                            1 batched general rows, 2 batched tagged rows, 3 batched packed 16-bit,
	                            4 batched tagged packed 16-bit (affine) */
	int32_t long_format;      /* Config::bSpeicialVersion (-v), record mode only: seeds on a query with the
	                             subject's uniqueId are kept (src/SubjectReadAligner.cpp:464-465) and every
	                             accepted record is kept, none replaced (src/SubjectRead.cpp:26-29); read them
	                             with hero_take_records */
	int32_t merge;            /* Config::bMerging (--merge), with call_consensus: each kept record whose CIGAR has
	                             more than one op also votes the (up to) 20 subject bases after its end into its
	                             query's extension tally (SubjectRead::updateQueryCoverageConsiderMerging,
	                             src/SubjectRead.cpp:332-461); read it with hero_get_extension_tally */
	int32_t file_order;       /* 1: subjects run in file order and the tallies lie in read order; 0 (vote modes):
	                             locality order, subjects run and tallies lie in the order of each read's
	                             minimizer, so overlapping reads share cache lines; results are identical */
	int64_t chunk_seeds_max;  /* seeds of one subject chunk before it runs as two halves (each counted again,
	                             in subject order); 0 = 2^30, half of what the seed stage indexes */
} hero_options;

/* Counters and device timings (milliseconds, HIP events on the context's stream). */
typedef struct {
	uint64_t table_size, distinct_keys, total_keys, max_probe;
	uint64_t lookups, seeds, dp_tasks, dp_calls_exec, dp_cells_exec, dp_cells_all;
	uint64_t records_built, records_kept, dp_skipped, vote_oob, subjects, queries, reads_changed;
	double ms_index, ms_seed, ms_plan, ms_dp, ms_score_vote, ms_consensus, ms_total;
	double ms_seed_probe; /* the seed-count probe kernel alone (part of ms_seed) */
	uint64_t dp_graph_builds;   /* DP stage graphs captured (a buffer moved or the schedule changed) */
	uint64_t dp_graph_launches; /* DP stage graph launches (one per subject chunk) */
} hero_stats;

void hero_default_options(hero_options *opt);

/* Context; replaces the process-wide Config + per-thread BandAligner scratch. */
int hero_create(const hero_options *opt, hero_ctx **out);
void hero_destroy(hero_ctx *ctx);
const char *hero_last_error(const hero_ctx *ctx);

/* Query reads (QueryDataset::loadDatasetParallel result, src/QueryDataset.cpp:103-244): upper-case
 * ASCII `seq` with per-read offset/length and uniqueId = atoi(name). Good reads
 * (len > minOverlap && qualityFilter, src/QueryDataset.cpp:60-75,226) are indexed and corrected.
 * Resets the vote tallies. */
int hero_set_queries(hero_ctx *ctx, int64_t n, const char *seq, const int64_t *offset, const int32_t *length,
		const int64_t *unique_id);

/* HashTableMethod::createHashTables + insertQueryDataset (src/HashTableLongKey.cpp:41-68,250-342). */
int hero_build_index(hero_ctx *ctx);

/* SubjectReadAligner::start (src/SubjectReadAligner.cpp:493-645) over one batch of subject reads
 * (SubjectDataset::loadNextChunkParallel chunks, src/SubjectDataset.cpp:45-179): seeds, chains,
 * banded DP, records, dedup, and (under -c) votes into the query tallies. May be called
 * repeatedly; votes accumulate. */
int hero_align_subjects(hero_ctx *ctx, int64_t n, const char *seq, const int64_t *offset, const int32_t *length,
		const int64_t *unique_id);

/* Same, with the query set itself as the subject stream (self-correction: -q f -s f); no second
 * upload. */
int hero_align_self(hero_ctx *ctx);

/* Record mode (options.records = 1, without -c): the .aln text that SubjectReadAligner writes for
 * the subjects aligned since the last call (src/SubjectReadAligner.cpp:854-910,
 * AlignmentRecord.cpp:102-149, SubjectRead.cpp:25-79,179-193): ">sid\n", then
 * "qid,strand,pos,CIGAR,interiorInsertedBases," per record, then "\n", for every subject in input
 * order. *len receives the text's size; when out is not null and cap >= *len the text is copied
 * (without a terminating zero) and released. */
int hero_take_alignments(hero_ctx *ctx, char *out, int64_t cap, int64_t *len);

/* Record mode: the accepted alignment records of the subjects aligned since the last call, as data
 * (SubjectRead::vAlignment after addAlignment, src/SubjectRead.cpp:25-79; under long_format every
 * accepted record, in processing order). Records are grouped by subject in stream order; subject
 * counts subjects since the last call. Each record's CIGAR is ops[op_off .. op_off + op_n): (count << 8)
 * | type with type 'M', 'I', 'D' or the query base of a mismatch run (AlignmentRecord::addCigar,
 * src/AlignmentRecord.cpp:35-48). *n_recs / *n_ops receive the sizes; when recs and ops are not
 * null and the caps suffice, the records are copied and released. */
typedef struct {
	int64_t subject;  /* subject index since the last call */
	int64_t query;    /* innerId of the query read */
	int32_t strand;   /* AlignmentRecord::iOrientation: 0 forward, 1 reverse complement */
	int32_t position; /* iPositionOfQueryOnSubject */
	int64_t op_off;
	int32_t op_n, _pad;
} hero_record;
int hero_take_records(hero_ctx *ctx, hero_record *recs, int64_t rec_cap, int64_t *n_recs, uint32_t *ops, int64_t op_cap,
		int64_t *n_ops);

/* QueryDataset::callConsensus (src/QueryDataset.cpp:393-412, QueryRead.cpp:148-180): writes the
 * corrected reads in the layout of hero_set_queries' `seq` (reads that are not good are copied
 * unchanged, as the reference writes them). With corrected == NULL the corrected reads stay in
 * device memory; hero_copy_corrected fetches them. */
int hero_call_consensus(hero_ctx *ctx, char *corrected);
int hero_copy_corrected(hero_ctx *ctx, char *corrected);

/* The vote matrices (QueryRead::aCoverageMatrix, 4 u32 per base in A,C,T,G order) as they stand
 * after the subject passes, before consensus adds the self vote. */
int hero_get_tally(hero_ctx *ctx, uint32_t *tally);

/* Writes the vote matrices (QueryRead::aCoverageMatrix, src/QueryRead.h:19, the tally that
 * QueryRead::callConsensus reads at src/QueryRead.cpp:148-180): 4 u32 per base in A,C,T,G order at
 * the queries' offsets, the layout hero_get_tally returns. It replaces the votes of the subject
 * passes; hero_call_consensus then decides every base from it (self vote included). */
int hero_set_tally(hero_ctx *ctx, const uint32_t *tally);

/* --merge: QueryRead::aExtraCoverageMatrix of every query read (src/QueryRead.cpp:100-115,353-371),
 * 20 positions x 4 u32 in A,C,T,G order, position 0 = the first subject base after a record's end;
 * out holds 80 x the query count. */
int hero_get_extension_tally(hero_ctx *ctx, uint32_t *out);

/* --mapping output (QueryDataset::getCoverageAndWriteResults, src/QueryDataset.cpp:560-600): counts[v]
 * = bases of the query reads whose coverage (QueryRead::getCoverageAt: the sum of the base's four
 * vote counters) is v, for v < cap; *max_coverage = the largest coverage. The reference writes
 * "COV\tFRE" and "v\tcounts[v]" for every v <= max_coverage with counts[v] > 0. Returns
 * HERO_ERR_LIMIT when max_coverage >= cap (counts[cap-1] then also holds every larger value). */
int hero_coverage_histogram(hero_ctx *ctx, uint64_t *counts, int32_t cap, int32_t *max_coverage);

int hero_get_stats(hero_ctx *ctx, hero_stats *out);
int hero_reset_stats(hero_ctx *ctx);

/* BandAligner drop-in (src/BandAligner.h:20-92): n independent calls of
 *   setBand(band); setSequences(S[s_off, s_off+s_len), Q[q_off, q_off+q_len)); flag 1 setAlignLeft /
 *   2 setAlignRight; calculateAlignment();
 * over bytes A/C/G/T/N in `seq`. Outputs per call the widened m_band and the CIGAR of
 * getCigarType()/getCigarCount() (mismatch ops typed by the query base, unmerged) at
 * cigar_types/cigar_counts[cigar_off[i] .. cigar_off[i] + cigar_n[i]); cigar_cap bounds the total. */
typedef struct {
	int64_t s_off, q_off;
	int32_t s_len, q_len;
	int32_t band, flag;
} hero_band_pair;
int hero_band_align_batch(hero_ctx *ctx, int32_t n, const hero_band_pair *pairs, const char *seq, int32_t *band_out,
		int64_t *cigar_off, int32_t *cigar_n, int32_t *cigar_counts, char *cigar_types, int64_t cigar_cap);

/* No counterpart in the reference. For callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc): hero_band_align_batch on
 * device-resident arrays, stream-ordered on `stream` (a hipStream_t; null = the context's stream),
 * allocating nothing and never synchronizing. Every pointer is device memory: pairs, seq (bytes
 * A/C/G/T/N), band_out, cigar_n, and cigar_counts / cigar_types, where call i's ops go at
 * cigar_off[i] (an input: the caller reserves s_len + q_len ops there). max_len bounds both sides
 * of every call (<= 16383) and max_band its widened band (<= 1023). cigar_n[i] < 0 reports a call
 * the blocking entry rejects: HERO_ERR_REF_ABORT (negative start), HERO_ERR_ARG (both sides
 * empty), HERO_ERR_LIMIT (beyond max_len / max_band). Work counters and the traceback pointers of
 * calls too large for LDS live in one fixed 64 MiB device array per GPU (HERO_ERR_LIMIT when calls
 * of max_len x max_band need more); calls that overlap on one GPU share it and race. */
int hero_band_align_batch_noalloc(hero_ctx *ctx, int32_t n, const hero_band_pair *pairs, const char *seq,
		int32_t max_len, int32_t max_band, int32_t *band_out, const int64_t *cigar_off, int32_t *cigar_n,
		int32_t *cigar_counts, char *cigar_types, void *stream);

/* HashTableLongKey::getPrimeLargerThanNumber (src/HashTableLongKey.cpp:101-166): the smallest prime
 * of the reference's table-size list above n, else n + 1. hero_build_index sizes the query index with
 * it (factor x reads, or factor x bases / k under --mapping, :52-62). Host only. */
uint64_t hero_prime_larger_than(uint64_t n);

/* HashTableLongKey::hashFunction (src/HashTableLongKey.cpp:188-206) on the device for n k-mers
 * of length k packed back to back in `kmers` (A/C/G/T). */
int hero_hash_kmers(hero_ctx *ctx, int32_t n, int32_t k, const char *kmers, uint64_t table_size, uint64_t *out);

/* HashTableLongKey::getListOfReads (src/HashTableLongKey.cpp:79-99) on the device: for n windows
 * of length k, the bucket size and its entries ((pos << 40) | innerId, insertion order) packed at
 * entries[entry_off[i] ..]; entry_cap bounds the total. */
int hero_lookup(hero_ctx *ctx, int32_t n, const char *windows, int64_t *count, int64_t *entry_off, uint64_t *entries,
		int64_t entry_cap);

/* Autotuner of the batched DP form: runs the planning of the first max_subjects subjects of the
 * query set (0 = 65,536) and times, per band class, every schedule candidate (one task per lane in
 * 32-bit or two per lane in packed 16-bit; persistent wavefronts per CU from 1 to the occupancy
 * limit), then keeps the fastest for later calls with dp_per_lane = 0 and dp_waves = 0. Casts no
 * votes and leaves the statistics as they were. Every candidate computes identical results. */
typedef struct {
	int32_t per_lane, waves_per_cu;
	float ms;
} hero_tune_candidate;
typedef struct {
	int32_t band, per_lane, waves_per_cu, ncand;
	int64_t tasks, cells; /* the class's DP calls and their band cells */
	float ms, ms_default; /* chosen candidate / one task per lane at the occupancy limit */
	hero_tune_candidate cand[16];
} hero_tune_class;
#define HERO_TUNE_MAX_CLASSES 12 /* band classes of the batched form: 4, 6, 8, 10, 12, 16, 24, 32 */
typedef struct {
	int32_t classes, _pad;
	int64_t subjects;
	hero_tune_class cls[HERO_TUNE_MAX_CLASSES];
} hero_tune_report;
int hero_autotune(hero_ctx *ctx, int64_t max_subjects, hero_tune_report *report);

/* ---- Serving -------------------------------------------------------------------------------
 * The subject set stays resident with an index of its k-mers at every position; batches of query
 * reads arrive and are corrected against it (-c). The index and probe sides of
 * SubjectReadAligner::searchHashTable2 (src/SubjectReadAligner.cpp:444-491) are exchanged: each
 * query key and its reverse complement probe the subjects, which yields the reference's seed set;
 * everything after the seeds is the offline path. A query's result does not depend on the other
 * queries of its batch (max_candidates must be 0). */
int hero_set_subjects(hero_ctx *ctx, int64_t n, const char *seq, const int64_t *offset, const int32_t *length,
		const int64_t *unique_id);
int hero_build_subject_index(hero_ctx *ctx);
/* One batch corrected against the resident subjects; output in the layout of `seq` (reads that are
 * not good are copied unchanged). Uses the context's own stream. */
int hero_correct_batch(hero_ctx *ctx, int64_t n, const char *seq, const int64_t *offset, const int32_t *length,
		const int64_t *unique_id, char *corrected);

/* Request loop with continuous batching: submitted batches wait in arrival order; each worker
 * (stream) takes the longest prefix of waiting batches that fits its share of hbm_budget and
 * max_step_reads, corrects them as one step and completes them, then takes the next. Steps of
 * different workers overlap on the device. The caller's buffers must stay valid until
 * hero_server_wait returns for that ticket. Times are steady-clock seconds. */
typedef struct hero_server hero_server;
typedef struct {
	int32_t workers;        /* concurrent steps (streams); 0 = 2 */
	int32_t max_step_reads; /* reads merged into one step at most; 0 = 32768 */
	int64_t hbm_budget;     /* device bytes for in-flight steps; 0 = half of the free memory */
	int32_t reserved[8];
} hero_server_options;
typedef struct {
	int32_t status;        /* HERO_OK or the step's error */
	int32_t step_reads;    /* reads of the step that carried this batch */
	int32_t step_batches;  /* batches merged into that step */
	int32_t _pad;
	int64_t step_id;
	double t_submit, t_admit, t_done;
} hero_request_info;
int hero_server_start(hero_ctx *ctx, const hero_server_options *opt, hero_server **out);
/* HERO_ERR_NOMEM: the batch alone exceeds hbm_budget (never admissible). */
int hero_server_submit(hero_server *srv, int64_t n, const char *seq, const int64_t *offset, const int32_t *length,
		const int64_t *unique_id, char *corrected, int64_t *ticket);
/* Returns the batch's status (also in info->status); a failed batch's error text is then in
 * hero_last_error. A failed batch leaves the server and its workers serving the next ones. */
int hero_server_wait(hero_server *srv, int64_t ticket, hero_request_info *info);
/* Drains the queue, stops the workers; `merged` (optional) receives their summed device timings.
 * Batch errors are reported by hero_server_wait, not here. */
int hero_server_stop(hero_server *srv, hero_stats *merged);

#ifdef __cplusplus
}
#endif
#endif
