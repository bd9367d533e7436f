// Device data structures and kernel launchers of the correction path (internal to libhero_gpu).
#pragma once
#include <hip/hip_runtime.h>

#include "hd.h"

namespace hero {

// Open-addressing slot of the seed index. The table is keyed by canonical k-mers: the smaller
// (as the number (key, key_lo)) of a k-mer and its reverse complement, first 32 bases in key, the
// rest of a longer one in key_lo. Its bucket in `entries` is [start, start + count): the nfwd
// occurrences of the canonical k-mer itself, then the nrc occurrences of its reverse complement,
// so one probe answers a window on both strands. count == 0 marks an empty slot. 32 bytes,
// aligned, so a probe reads one cache line.
struct __align__(32) Slot {
	uint64_t key;
	uint64_t key_lo;
	uint32_t start;
	uint32_t count;
	uint32_t nfwd, nrc;
};

// The canonical form (chi, clo) of k-mer (hi, lo); true when the k-mer is its own canonical form
// (as a palindrome of even k is).
HD bool kmer_canonical(int k, uint64_t hi, uint64_t lo, uint64_t &chi, uint64_t &clo) {
	uint64_t rhi, rlo;
	kmer_revcomp(k, hi, lo, rhi, rlo);
	const bool fwd = hi < rhi || (hi == rhi && lo <= rlo);
	chi = fwd ? hi : rhi, clo = fwd ? lo : rlo;
	return fwd;
}

// Bucket entry: (tag << 54) | (pos << 40) | innerId (src/HashTableLongKey.cpp:334-336 without
// the tag). The tag hashes the read's uniqueId, so the self-hit test (same uniqueId as the
// subject) loads the uniqueId only when the tags agree.
constexpr uint64_t kIdMask = (1ull << 40) - 1; // iMask for iBitOfReadId = 40
constexpr uint64_t kEntryMask = (1ull << 54) - 1;
HD uint32_t uid_tag(int64_t uid) { return (uint32_t) (((uint64_t) uid * 0x9E3779B97F4A7C15ull) >> 54); }
HD int entry_pos(uint64_t e) { return (int) ((e >> 40) & 0x3FFFu); }
HD uint32_t entry_tag(uint64_t e) { return (uint32_t) (e >> 54); }

struct DevIndex {
	const Slot *slots;
	const uint64_t *entries;
	uint64_t P;
	uint32_t max_probe;
	int32_t k;
	FastMod fm; // P with its reciprocal
};

#if defined(__HIPCC__)
// The slot of canonical k-mer (chi, clo): from its home slot to the first empty one.
__device__ __forceinline__ bool ix_probe(const DevIndex &ix, uint64_t chi, uint64_t clo, Slot &out) {
	uint64_t h = hash_kmer(ix.k, chi, clo, ix.fm);
	for (uint32_t pr = 0; pr <= ix.max_probe; pr++) {
		const Slot s = ix.slots[h];
		if (s.count == 0) return false;
		if (s.key == chi && (ix.k <= 32 || s.key_lo == clo)) {
			out = s;
			return true;
		}
		h = (h + 1 == ix.P) ? 0 : h + 1;
	}
	return false;
}
// ix_probe with the home slot h already loaded (first): the probe continues past it only on a
// collision.
__device__ __forceinline__ bool ix_probe_from(const DevIndex &ix, uint64_t chi, uint64_t clo, uint64_t h,
		const Slot &first, Slot &out) {
	Slot s = first;
	for (uint32_t pr = 0; pr <= ix.max_probe; pr++) {
		if (pr) s = ix.slots[h];
		if (s.count == 0) return false;
		if (s.key == chi && (ix.k <= 32 || s.key_lo == clo)) {
			out = s;
			return true;
		}
		h = (h + 1 == ix.P) ? 0 : h + 1;
	}
	return false;
}
// The half of a canonical bucket holding the occurrences of the k-mer whose canonical-form test
// gave fwd (kmer_canonical).
__device__ __forceinline__ void ix_half(const Slot &s, bool fwd, uint32_t &start, uint32_t &count) {
	start = fwd ? s.start : s.start + s.nfwd;
	count = fwd ? s.nfwd : s.count - s.nfwd;
}
// Occurrences of exactly the k-mer (hi, lo) (HashTableLongKey::getListOfReads,
// src/HashTableLongKey.cpp:79-99): false when it has none.
__device__ __forceinline__ bool ix_lookup(const DevIndex &ix, uint64_t hi, uint64_t lo, uint32_t &start, uint32_t &count) {
	uint64_t chi, clo;
	const bool fwd = kmer_canonical(ix.k, hi, lo, chi, clo);
	Slot s;
	if (!ix_probe(ix, chi, clo, s)) return false;
	ix_half(s, fwd, start, count);
	return count != 0;
}
#endif

// One accepted alignment record of a chunk (record mode, without -c). Slots are laid out in the
// reference's processing order: subject, strand, query group (innerId), chain index; a slot whose
// chain was not accepted keeps qid = kNoRecord.
constexpr uint32_t kNoRecord = 0xFFFFFFFFu;
struct RecHead {
	uint32_t qid;    // innerId of the query read
	uint32_t script; // the record's script (index into the chunk's program words)
	uint32_t task0;  // first DP task of the record
	int32_t sub;     // subject index within the chunk
	int32_t strand, _pad;
	double rate_sum, len; // isGoodAlignment() rates, for SubjectRead::addAlignment
};

// The kept record of one (subject, query) pair, for its votes.
struct VoteItem {
	uint32_t q;      // innerId of the query read; kNoRecord for an unused slot
	uint32_t script; // the record's script (index into the chunk's program words)
	uint32_t task0;  // first DP task of the record
	int32_t sub;     // subject index within the chunk
	int32_t strand;
	// 1: the query's entry in the subject's record list (SubjectRead::vAlignment) came from a forward
	// record, so it precedes every entry first added by a reverse record (src/SubjectRead.cpp:25-79)
	int32_t first_fwd;
};

// Per-chunk device buffers. Seed-parallel arrays hold, at a group's first seed, that group's
// task and edit-word counts and their exclusive scans.
struct ChunkDev {
	int64_t s_begin;
	int32_t nsub;
	int32_t win_stride;        // windows per (subject, strand) slot of win_info
	uint64_t *win_info;        // per window: bucket (start << 32 | count), 0 = no bucket; null = probe again
	uint32_t *lane_n;          // per (subject, strand) slot and lane: seeds of that lane's windows
	uint32_t *seg_n, *seg_off; // 2 per subject (forward, reverse)
	uint64_t *seeds;
	int32_t *cnext, *cfirst, *clast, *clen;
	uint8_t *gflag; // bit0 group start, bit1 group kept (within max_candidates)
	int32_t *g_nch, *g_best; // chain count and best covered length of a kept group
	uint32_t *g_tasks, *g_task_off, *g_words, *g_word_off;
	uint32_t *g_prog, *g_prog_off; // script words of a kept group (record scripts) and their scan
	uint32_t *prog;                // record scripts of every kept group
	uint32_t *g_recs, *g_rec_off; // record mode: scripts of a kept group (record slots) and their scan
	RecHead *recs;                // record mode: one slot per chain of every kept group
	uint32_t *sub_pairs, *sub_pair_off; // vote mode: kept groups of both strands per subject, scan
	VoteItem *votes;                    // vote mode: the kept record of each pair (q = kNoRecord: none)
	DpTask *tasks;
	uint32_t *words, *nsteps;
	StepSummary *sums;
};

// Stats slots (u64 atomics on the device).
enum StatSlot : int {
	kStLookups = 0,
	kStSeeds,
	kStCallsExec,
	kStCellsExec,
	kStBuilt,
	kStKept,
	kStOob,
	kStDpSkipped,
	kStReadsChanged,
	kStCellsAll,
	kStCount
};
// Device counters are striped over kStatStripes cache lines (kStatStride u64 apart), chosen by
// workgroup, so that per-wave updates do not serialize on one address; readers sum the stripes.
constexpr int kStatStripes = 256;
constexpr int kStatStride = 16;
static_assert(kStCount <= kStatStride, "one stripe holds every counter");

struct DpLaunch {
	const DpTask *tasks;
	uint32_t ntasks;
	const uint32_t *order; // optional permutation (length buckets)
	Packed S, Q;
	Opts o;
	uint32_t *words, *nsteps;
	StepSummary *sums;
	uint32_t *scratch;
	int rows_max;
	int tagged;                 // one task per lane, affine: rows in the tagged form (dp_tagged_form)
	// one task per lane: the wavefront's LDS bytes (the launch's dynamic LDS) for its first pointer
	// rows; the rest go to scratch (0: every row in scratch)
	uint32_t ptr_lds_bytes;
	uint32_t *batch_counter;
	unsigned long long *stats;
	// graph launches: the task range [range[0], range[1]) of the order array, read on the device
	// (null: tasks order[0 .. ntasks))
	const uint32_t *range;
};

// launchers (kernels.hip)
hipError_t launch_index_gen(const Packed &Q, const uint64_t *koff, int k, bool all_positions, uint64_t *occ_hi,
		uint64_t *occ_lo, uint64_t *occ_entry, uint32_t *occ_rank, hipStream_t st);
hipError_t launch_index_claim(const uint64_t *occ_hi, const uint64_t *occ_lo, uint64_t nocc, uint64_t P, int k,
		uint32_t *owner, uint32_t *occ_slot, uint32_t *max_probe, int *err, hipStream_t st);
hipError_t launch_index_count(const uint32_t *occ_slot, uint64_t nocc, Slot *slots, uint32_t *occ_rank, hipStream_t st);
hipError_t launch_index_alloc(const uint32_t *owner, uint64_t P, const uint64_t *occ_hi, const uint64_t *occ_lo, int k,
		Slot *slots, unsigned long long *cursor, hipStream_t st);
hipError_t launch_index_fill(const uint32_t *occ_slot, const uint32_t *occ_rank, const uint64_t *occ_entry, uint64_t nocc,
		const Slot *slots, uint64_t *entries, hipStream_t st);
hipError_t launch_index_sort_buckets(const Slot *slots, uint64_t P, uint64_t *entries, hipStream_t st);

hipError_t launch_seed_count(const Packed &S, const Packed &Q, const DevIndex &ix, const Opts &o, const ChunkDev &c,
		unsigned long long *stats, hipStream_t st);
hipError_t launch_pass_a1(const Packed &S, const Packed &Q, const DevIndex &ix, const Opts &o, const ChunkDev &c,
		unsigned long long *stats, hipStream_t st);
hipError_t launch_pass_a2(const Packed &S, const Packed &Q, const Opts &o, const ChunkDev &c, hipStream_t st);
// Pass B scores every (subject, query) pair and writes the winner of each kept pair to c.votes
// (at c.sub_pair_off[subject] + rank); launch_votes then casts their votes.
hipError_t launch_pass_b(const Packed &S, const Packed &Q, const Opts &o, const ChunkDev &c, unsigned long long *stats,
		hipStream_t st);
// ext (--merge, else null): kExtPositions x 4 u32 per query read, the extension tally
hipError_t launch_votes(const Packed &S, const Packed &Q, const Opts &o, const ChunkDev &c, uint32_t nvotes,
		const uint64_t *q_boff, uint32_t *tally, uint32_t *diff, uint32_t *ext, unsigned long long *stats, hipStream_t st);
// --mapping: per subject of the chunk, keep the votes of the best record covering it (see the kernel)
hipError_t launch_map_select(const Packed &S, const ChunkDev &c, hipStream_t st);
// --mapping output: hist[v] += bases of the nq reads with coverage v (bins 0 .. cap-1, the last one
// taking every larger value), *max_cov = max(*max_cov, largest coverage)
hipError_t launch_coverage_hist(const int32_t *len, const uint64_t *q_boff, const uint32_t *diff, int64_t nq, uint32_t cap,
		unsigned long long *hist, uint32_t *max_cov, hipStream_t st);
// Record mode: op count (ops == nullptr) or ops of every accepted record slot; each op is
// (count << 8) | type, type 'M' / 'I' / 'D' or the query base letter of a mismatch run.
hipError_t launch_record_ops(const Packed &S, const Packed &Q, const Opts &o, const ChunkDev &c, uint32_t nslots,
		uint32_t *op_n, const uint32_t *op_off, int32_t *pos, uint32_t *ops, hipStream_t st);
// per_lane: tasks per lane, 1 (32-bit) or 2 (packed 16-bit, band classes up to kPk16MaxBand)
hipError_t launch_dp_batched(const DpLaunch &L, int bmax, int per_lane, int nwaves, hipStream_t st);
// occupancy of one batched instantiation (rows_max sizes the packed form's LDS; tagged: the
// one-task-per-lane affine rows in the tagged form)
int dp_batched_waves_per_cu(int bmax, bool affine, int per_lane, int rows_max, bool tagged);
// LDS bytes for the near pointer rows of a one-task-per-lane launch at waves_per_cu waves per CU:
// the CU's LDS shared out over those waves, no more than still lets them all be resident
uint32_t dp_batched_lds_bytes(int bmax, bool affine, bool tagged, int waves_per_cu);
constexpr int kPk16MaxBand = 12;   // widest band class of the two-tasks-per-lane form (windows of 32 cells)
constexpr int kTaggedMaxBand = 12; // widest band class of the tagged affine rows (windows of 32 cells)
// Whether one-task-per-lane rows may run in the tagged form: the global aligner, or the affine one
// with gap opens of equal cost both ways; every scaled value of a task of up to rows_max rows
// within range (dp_core.h).
bool dp_tagged_form(const Opts &o, int rows_max);
// Whether a band class may run two tasks per lane in 16-bit halves (tagged: the tagged packed
// rows of affine launches with L.tagged set; their range test is dp_pk16.h's pk16t_fits).
bool pk_form_fits(const Opts &o, int rows_max, int bmax, bool tagged);
// The packed form stages each lane's four segments (< rows_max codes each) in LDS: words per
// segment (odd, so the lanes' segments start in different banks) and bytes per wavefront.
inline __host__ __device__ int pk_lds_words(int rows_max) { return ((rows_max + 15) / 16) | 1; }
inline __host__ __device__ size_t pk_lds_bytes(int rows_max) { return (size_t) 64 * 4 * pk_lds_words(rows_max) * 4; }

// Band classes of the batched form; tasks above the last class (or with rows beyond the pointer
// scratch) go to the plain form. Buckets are (class, 1023 - min(lenS, 1023)): one class at a time,
// longest rows first, so the 64 lanes of a wavefront run tasks of similar length.
// the reference's bands are even (2 round(rate * overlap), then widened to the length gap): one
// class per even band from 4 to 12 (band 2 calls are short and few: they share class 4), then 16,
// 24, 32
constexpr int kNumBandClasses = 8;
constexpr int kBandClass[kNumBandClasses] = {4, 6, 8, 10, 12, 16, 24, 32};
constexpr int kLenBins = 1024; // per band class: 512 row counts x (aligned-left or not)
constexpr int kNumBuckets = (kNumBandClasses + 1) * kLenBins;
// ntasks_dev (optional) supplies the task count on the device; the grid then covers max_tasks
hipError_t launch_task_histogram(const DpTask *tasks, uint32_t ntasks, const uint32_t *ntasks_dev, int rows_max,
		uint32_t *hist, hipStream_t st);
hipError_t launch_task_scatter(const DpTask *tasks, uint32_t ntasks, const uint32_t *ntasks_dev, int rows_max,
		uint32_t *cursor, uint32_t *order, hipStream_t st);
// Graph form of the DP stage: class task ranges from the bucket offsets (ctl[1 + c], c = 0 ..
// kNumBandClasses, then ctl[2 + kNumBandClasses] = ntasks = ctl[0]) and zeroed class counters
// (ctl[16 + c]); the plain form's control words follow the counters.
constexpr int kCtlRange = 1, kCtlCounter = 16;
constexpr int kCtlWave = 32; // the plain form's kWaveCtlWords control words (zeroed with the counters)
constexpr int kCtlWords = 40;
static_assert(kCtlRange + kNumBandClasses + 2 <= kCtlCounter, "class ranges overlap the class counters");
static_assert(kCtlCounter + kNumBandClasses <= kCtlWave, "class counters overlap the plain form's words");
hipError_t launch_class_ranges(const uint32_t *bucket_off, uint32_t *ctl, hipStream_t st);
hipError_t launch_fill_u32(uint32_t *p, uint32_t n, uint32_t v, hipStream_t st);
// Tallies: 4 counters per base (mismatch votes by base code) at 4 * t_off[q], and the matched-run
// difference array, 1 counter per base, at t_off[q]; the read's outputs (codes_out, decoded_out)
// at its base offset q_boff[q].
// ASCII A/C/G/T bases -> 2-bit codes, 16 per u32 ((nb + 15) / 16 words), for the copy to the host.
hipError_t launch_pack_bases(const uint8_t *ascii, uint64_t nb, uint32_t *packed, hipStream_t st);
hipError_t launch_consensus(const Packed &Q, const Opts &o, const uint64_t *q_boff, const uint64_t *t_off,
		const uint32_t *tally, const uint32_t *diff, int64_t q0, int64_t nq, uint8_t *codes_out, uint32_t *decoded_out,
		unsigned long long *stats, hipStream_t st);

// Serving batch intake: packs raw bytes into 2-bit words, evaluates the query predicate into good
// and writes key references (query << 14 | position) at koff for reads longer than min_ov.
hipError_t launch_read_intake(const char *raw, const uint64_t *boff, const int32_t *len, const uint64_t *woff,
		const uint64_t *koff, int64_t n, int k, uint32_t min_ov, uint64_t *words, uint64_t nwords, uint8_t *good,
		uint64_t *kref, hipStream_t st);
// Serving orientation: the subjects' k-mers at every position are indexed; each query key (kref =
// query << 14 | position) and its reverse complement probe them. Count pass (keys_out == nullptr)
// writes per-key seed counts; emit pass writes, at off[key], keys_out = strand << 28 | posQ << 14 |
// posS and vals_out = subject << 22 | query.
hipError_t launch_query_probe(const Packed &Q, const Packed &S, const DevIndex &six, const Opts &o, const uint64_t *kref,
		uint32_t nkeys, uint32_t *cnt, const uint32_t *off, uint64_t *keys_out, uint64_t *vals_out, hipStream_t st);
// A serving batch's sorted seeds (hi = subject << 22 | query, lo = strand << 28 | posQ << 14 |
// posS), its groups (glist: first seed of each (subject, query, strand) run; gx: exclusive scan of
// group starts per seed) and pairs (plist: first seed of each (subject, query) run).
struct ServeDev {
	const uint64_t *hi, *lo;
	const uint32_t *glist, *gx, *plist;
	uint32_t nseeds, ngroups, npairs;
};
hipError_t launch_serve_mark(const uint64_t *hi, const uint64_t *lo, uint32_t n, uint32_t *gflag, uint32_t *pflag,
		hipStream_t st);
hipError_t launch_serve_lists(const uint64_t *hi, const uint64_t *lo, uint32_t n, const uint32_t *gflag,
		const uint32_t *gx, const uint32_t *pflag, const uint32_t *px, uint32_t *glist, uint32_t *plist, uint64_t *seeds,
		hipStream_t st);
hipError_t launch_serve_a1(const Packed &S, const Packed &Q, const Opts &o, const ChunkDev &c, const ServeDev &v,
		unsigned long long *stats, hipStream_t st);
hipError_t launch_serve_a2(const Packed &S, const Packed &Q, const Opts &o, const ChunkDev &c, const ServeDev &v,
		hipStream_t st);
hipError_t launch_serve_b(const Packed &S, const Packed &Q, const Opts &o, const ChunkDev &c, const ServeDev &v,
		const uint64_t *q_boff, uint32_t *tally, uint32_t *diff, unsigned long long *stats, hipStream_t st);

// Radix sort of (key, value) u64 pairs on the low key_bits bits (serve_sort.hip).
hipError_t sort_pairs_u64(const uint64_t *keys_in, uint64_t *keys_out, const uint64_t *vals_in, uint64_t *vals_out,
		uint32_t n, int key_bits, void *tmp, size_t &tmp_bytes, hipStream_t st);

// Locality order of a read set (locality.hip): reads sorted by their minimizer (the smallest mix
// of a canonical 16-base window; inactive reads last), written as an ordered copy: read r of the
// copy is read perm[r] of the set, its flag taken from `flag`. With t_off given, t_off[i] is the
// base offset of read i's place in the order (the tally layout); with perm given, the order.
struct OrderOut {
	uint64_t *words, *woff; // the copy's packed words (capacity: the set's words) and word offsets
	int32_t *len;
	int64_t *uid;
	uint8_t *good;
	uint64_t *t_off; // may be null
	uint64_t *perm;  // may be null
};
size_t order_tmp_bytes(int64_t n);
// Index key offsets of a read set on the device: koff[i] = keys of the good reads before i (the
// reference's key positions, or every position), koff[n] = their total.
size_t key_offsets_tmp_bytes(int64_t n);
// exclusive scan of n u64 (locality.hip); tmp of scan_u64_tmp_bytes(n)
size_t scan_u64_tmp_bytes(int64_t n);
hipError_t scan_u64(const uint64_t *in, uint64_t *out, int64_t n, void *tmp, size_t tmp_bytes, hipStream_t st);
// Exclusive scans of up to kScanMulti u32 arrays of one length at once (kernels.hip); their exact
// 64-bit totals go to totals_dev[0 .. K). tmp of scan_multi_tmp_bytes(K, n).
constexpr int kScanMulti = 4;
size_t scan_multi_tmp_bytes(int K, int64_t n);
hipError_t scan_u32_multi(int K, const uint32_t *const *in, uint32_t *const *out, int64_t n, uint64_t *totals_dev,
		void *tmp, size_t tmp_bytes, hipStream_t st);
hipError_t launch_key_offsets(const Packed &R, int k, bool all_positions, uint64_t *koff, void *tmp, size_t tmp_bytes,
		hipStream_t st);
hipError_t launch_order(const Packed &R, const uint8_t *active, const uint8_t *flag, OrderOut &out, void *tmp,
		size_t tmp_bytes, hipStream_t st);

#if defined(__HIPCC__)
// Item of workgroup b of n: workgroups go to the 8 XCDs in turn (b % 8), so XCD x is given the
// x-th contiguous eighth of [0, n), in order; neighbouring items (subjects in locality order) then
// share that XCD's L2.
__device__ __forceinline__ uint32_t xcd_item(uint32_t b, uint32_t n) {
	const uint32_t q = n >> 3, r = n & 7u, x = b & 7u, t = b >> 3;
	return x * q + (x < r ? x : r) + t;
}
#endif

// Exclusive scan of n u32 values (n < 2^31); *total_dev receives the u64 sum.
hipError_t scan_u32(const uint32_t *in, uint32_t *out, int64_t n, unsigned long long *total_dev, uint32_t *tmp,
		size_t tmp_words, hipStream_t st);
size_t scan_tmp_words(int64_t n);

// Plain form of the DP (dp_wave.hip): one wavefront per call, bands up to kWaveMaxBand in
// kWaveLaunches launches by band range (bands 0-63, 64-255, 256-1023). Per wave: the codes of both
// sides in LDS (seq_cap bytes) and the traceback pointers in LDS (ptr_lds_words) or, for a call
// whose pointers do not fit there, in the wave's global scratch (scratch_words per wave, wave w at
// scratch + w * scratch_words).
constexpr int kWaveLaunches = 3;
constexpr int kWaveMaxBand = 1023;
struct WaveSizing {
	int32_t seq_cap;       // LDS bytes for the codes of one call (lenS + lenQ)
	int32_t ptr_lds_words; // LDS words for one call's pointers
	uint32_t *scratch;     // global pointer words (may be null)
	size_t scratch_words;  // per wave
	int32_t nwaves;
};
int wave_band_hi(int launch);          // widest band of a launch (its narrowest is the previous one's + 1)
size_t wave_lds_bytes(const WaveSizing &z);
size_t wave_call_ptr_words(int lenS, int b);
// The pipeline's DP tasks (DpLaunch: tasks, order / range, words, nsteps, sums, stats) whose band
// lies in the launch's range. ctl_wave holds kWaveCtlWords words, zero before launch 0: the launches'
// work counters and the overflow list's range; launch 0 runs the given tasks and lists those wider
// than its range in `over` (capacity: the task count) for launches 1 and 2.
constexpr int kWaveCtlCounter = 0, kWaveCtlOver = 4, kWaveCtlWords = 8;
static_assert(kCtlWave + kWaveCtlWords <= kCtlWords, "the plain form's words lie inside the control block");
hipError_t launch_dp_wave_tasks(const DpLaunch &L, const WaveSizing &z, int launch, uint32_t *ctl_wave, uint32_t *over,
		hipStream_t st);

// BandAligner calls over bytes A/C/G/T/N (the C-ABI drop-in); pairs are hero_band_pair records
// (include/hero_gpu.h). Per call: band_out, and cigar_n ops (or a negative status) written at
// cigar_off[t], whose region holds s_len + q_len ops. Calls beyond max_len / max_band get
// HERO_ERR_LIMIT from launch 0.
struct WavePairs {
	const void *pairs; // hero_band_pair[n]
	int32_t n;
	int32_t max_len, max_band;
	const char *seq;
	int32_t *band_out;
	const int64_t *cigar_off;
	int32_t *cigar_n, *cigar_counts;
	char *cigar_types;
	uint32_t *counter; // work counter, zero at launch
};
hipError_t launch_dp_wave_pairs(const WavePairs &P, const Opts &o, const WaveSizing &z, int launch, hipStream_t st);

} // namespace hero
