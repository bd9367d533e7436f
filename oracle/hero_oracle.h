/*
 * hero_oracle.h — CPU restatement of HERO's read-correction hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker for the HIP implementation in
 * hero_amd/: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 * The product path never links, loads or calls it.
 *
 * Every function restates the reference algorithm and cites the reference file:line it follows
 * (paths relative to the reference repository root, guo-xuan/Hero).
 */
#ifndef HERO_ORACLE_H
#define HERO_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Run configuration: the Config fields the hot path reads (src/Config.cpp:17-56,118-243),
 * the BandAligner cost table (src/BandAligner.cpp:528-542) and two build-only knobs. */
typedef struct {
	int32_t k;              /* Config::iKmer */
	int32_t aligner;        /* Config::iAlignerType: 0 affine, 1 global */
	int32_t call_consensus; /* Config::bCallConsensus (-c) */
	int32_t hash_factor;    /* Config::iHashTableFactor (-f) */
	int64_t min_overlap;    /* Config::minimumOverlapLength (-l) */
	double mismatch_rate;   /* Config::dMaxMismatchRate (-m / 100) */
	double indel_rate;      /* Config::dMaxIndelRate (-i / 100) */
	double consensus_freq;  /* Config::dMinFrequencyCallConsensus (-c <f>) */
	int32_t band_override;  /* build-only: >0 replaces setBand(2*extra) by setBand(band_override) */
	int32_t max_candidates; /* build-only: >0 keeps only the first N query groups per subject strand */
	int32_t dx, dy, ex, sx, miss, hit, unobserved; /* BandAligner::setCost */
	int32_t mapping;        /* Config::bMappingVersion (--mapping) */
} oracle_cfg;

/* A read set as the datasets hold it after loading: upper-case ASCII, concatenated. */
typedef struct {
	int64_t n;
	const char *seq;
	const int64_t *offset;
	const int32_t *length;
	const int64_t *unique_id; /* atoi(read name) (src/QueryDataset.cpp:215, SubjectDataset.cpp:154) */
} oracle_reads;

typedef struct {
	uint64_t table_size;        /* P */
	uint64_t distinct_keys;
	uint64_t total_keys;
	uint64_t max_probe;         /* maxSingleHashCollision */
	uint64_t lookups;           /* subject windows probed */
	uint64_t seeds;             /* PairKeyPosition objects created */
	uint64_t dp_calls;          /* DP calls the reference control flow executes */
	uint64_t dp_cells;          /* in-band cells of those calls */
	uint64_t records_built;     /* records that reached SubjectRead::addAlignment */
	uint64_t records_kept;      /* records left after dedup */
	uint64_t reads_changed;     /* corrected reads that differ from their input */
	double t_index, t_align, t_consensus; /* wall seconds */
} oracle_stats;

/* Default reference configuration (Config.cpp defaults + BandAligner::setCost). */
void oracle_default_cfg(oracle_cfg *cfg);

/* Dataset predicates of every read: query_good = len > (u16) minOverlap && qualityFilter
 * (src/QueryDataset.cpp:60-75,226); subject_good = len < (u16) minOverlap || qualityFilter
 * (src/SubjectDataset.cpp:157-162). */
void oracle_read_flags(const oracle_cfg *cfg, const oracle_reads *reads, uint8_t *query_good, uint8_t *subject_good);

/* HashTableLongKey::hashFunction (src/HashTableLongKey.cpp:188-206). */
uint64_t oracle_hash_kmer(const char *s, int64_t k, uint64_t table_size);

/* Table size: smallest prime > factor*n_reads (role of src/HashTableLongKey.cpp:41-68,101-166). */
uint64_t oracle_table_size(uint64_t n_reads, uint64_t factor);

/* One BandAligner call: setBand(*band) (unless keep_band), setSequences(S[s0,s1), Q[q0,q1)),
 * flag 1 = setAlignLeft, 2 = setAlignRight, calculateAlignment().
 * Writes the CIGAR (counts, type chars) and returns the number of ops, or -1 if cap is too small.
 * *band is updated to m_band after the call (widening, src/BandAligner.cpp:59-62,144-148). */
int oracle_band_align(const oracle_cfg *cfg, const char *subject, int64_t s0, int64_t s1, const char *query,
		int64_t q0, int64_t q1, int64_t *band, int flag, int32_t *counts, char *types, int cap, uint64_t *cells);

/* AlignmentRecord scoring (src/AlignmentRecord.cpp:35-100,302-316) over a CIGAR appended op by op
 * with addCigar merging. out[0..5] = nMis, nIndel, length, good(caps), good(rates), rateSum.
 * Returns the merged op count (merged ops written back to counts/types). */
int oracle_record_score(const oracle_cfg *cfg, int32_t *counts, char *types, int n, int64_t cap_mis,
		int64_t cap_indel, double *out);

/* QueryRead::callConsensus on one read (src/QueryRead.cpp:148-180): tally is 4*len u32 and is
 * updated in place with the self vote; seq is rewritten. Returns the number of changed bases. */
int oracle_consensus_read(const oracle_cfg *cfg, char *seq, int32_t len, uint32_t *tally);

/* Full path (src/main.cpp:40-106 with -c, src/SubjectReadAligner.cpp:493-645):
 * index the queries, stream the subjects, vote, call consensus.
 *   corrected  : out, same layout as queries->seq (may be NULL when !call_consensus)
 *   tally      : out, 4 u32 per query base at queries->offset (after consensus self-votes are
 *                NOT included: it is the vote state when the subject pass ends); may be NULL
 *   aln        : out, .aln text when !call_consensus (malloc'd, caller frees with oracle_free)
 *   records    : out, optional record dump, one line per kept record (malloc'd)
 *   threads    : OpenMP threads (<=0: library default)
 * Returns 0 on success, negative on error. */
int oracle_run(const oracle_cfg *cfg, const oracle_reads *queries, const oracle_reads *subjects, char *corrected,
		uint32_t *tally, char **aln, char **records, int threads, oracle_stats *stats);

/* Seeds of one subject strand, sorted (src/SubjectReadAligner.cpp:444-491, SubjectRead.cpp:157-177)
 * against an index built over `queries`. out is 3*cap int64 (innerId, posS, posQ); returns count
 * or -1 when cap is too small. The index is rebuilt per call (test helper). */
int64_t oracle_seeds(const oracle_cfg *cfg, const oracle_reads *queries, const char *subject, int32_t len,
		int64_t subject_unique_id, int reverse, int64_t *out, int64_t cap);

/* --mapping output (QueryDataset::getCoverageAndWriteResults, src/QueryDataset.cpp:560-600) from
 * the vote tally oracle_run returns: "COV\tFRE\n" then "<coverage>\t<bases>\n" for each coverage
 * some base has. malloc'd; free with oracle_free. */
char *oracle_coverage_text(const oracle_reads *queries, const uint32_t *tally);

void oracle_free(void *p);

#ifdef __cplusplus
}
#endif
#endif
