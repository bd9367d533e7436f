// C ABI of libhero_gpu: context, device residency of reads / index / tallies, and the per-chunk
// orchestration of the kernels in kernels.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdarg>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/hero_gpu.h"
#include "dp_pk16.h"
#include "kernels.h"
#include "pack.h"
#include "prime_table.h"
#include "pipeline_core.h"

using namespace hero;

namespace {

// Smallest prime strictly above n (deterministic Miller-Rabin for 64-bit n). Role of
// HashTableLongKey::getPrimeLargerThanNumber (src/HashTableLongKey.cpp:101-166): the table size
// only places slots, which never reaches the output.
uint64_t mulmod(uint64_t a, uint64_t b, uint64_t m) { return (uint64_t) (((__uint128_t) a * b) % m); }
uint64_t powmod(uint64_t a, uint64_t e, uint64_t m) {
	uint64_t r = 1;
	a %= m;
	while (e) {
		if (e & 1) r = mulmod(r, a, m);
		a = mulmod(a, a, m);
		e >>= 1;
	}
	return r;
}
bool is_prime(uint64_t n) {
	static const uint64_t B[12] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
	if (n < 2) return false;
	for (uint64_t b : B) {
		if (n == b) return true;
		if (n % b == 0) return false;
	}
	uint64_t d = n - 1;
	int r = 0;
	while (!(d & 1)) d >>= 1, r++;
	for (uint64_t b : B) {
		uint64_t x = powmod(b, d, n);
		if (x == 1 || x == n - 1) continue;
		bool comp = true;
		for (int j = 1; j < r; j++) {
			x = mulmod(x, x, n);
			if (x == n - 1) {
				comp = false;
				break;
			}
		}
		if (comp) return false;
	}
	return true;
}
uint64_t next_prime_above(uint64_t n) {
	uint64_t p = n + 1;
	while (!is_prime(p)) p++;
	return p;
}

// Device and pinned allocations, frees and DP graph captures of the whole process take turns.
// Two chunk or serving workers run on their own host threads and streams: a free (which waits for
// the device) or an allocation on one of them must not fall inside the other's stream capture,
// whose work is recorded, not run, until the capture ends.
std::recursive_mutex g_dev_mu;

template <class T>
struct DevBuf {
	T *p = nullptr;
	size_t n = 0;
	void release() {
		std::lock_guard<std::recursive_mutex> lk(g_dev_mu);
		if (p) (void) hipFree(p);
		p = nullptr, n = 0;
	}
	hipError_t ensure(size_t want) {
		if (want <= n && p) return hipSuccess;
		std::lock_guard<std::recursive_mutex> lk(g_dev_mu);
		release();
		size_t alloc = want + want / 4 + 64;
		hipError_t e = hipMalloc((void **) &p, alloc * sizeof(T));
		if (e != hipSuccess) {
			p = nullptr;
			return e;
		}
		n = alloc;
		return hipSuccess;
	}
	hipError_t exact(size_t want) {
		std::lock_guard<std::recursive_mutex> lk(g_dev_mu);
		release();
		hipError_t e = hipMalloc((void **) &p, (want ? want : 1) * sizeof(T));
		if (e != hipSuccess) {
			p = nullptr;
			return e;
		}
		n = want;
		return hipSuccess;
	}
};

// Pinned host staging, grow-only like DevBuf.
template <class T>
struct HostBuf {
	T *p = nullptr;
	size_t n = 0;
	void release() {
		std::lock_guard<std::recursive_mutex> lk(g_dev_mu);
		if (p) (void) hipHostFree(p);
		p = nullptr, n = 0;
	}
	hipError_t ensure(size_t want) {
		if (want <= n && p) return hipSuccess;
		std::lock_guard<std::recursive_mutex> lk(g_dev_mu);
		release();
		const size_t alloc = want + want / 4 + 64;
		hipError_t e = hipHostMalloc((void **) &p, alloc * sizeof(T), hipHostMallocDefault);
		if (e != hipSuccess) {
			p = nullptr;
			return e;
		}
		n = alloc;
		return hipSuccess;
	}
};

struct ReadsDev {
	DevBuf<uint64_t> words, woff, boff;
	DevBuf<int32_t> len;
	DevBuf<int64_t> uid;
	DevBuf<uint8_t> good;
	int64_t n = 0;
	uint64_t gen = 0; // which upload the contents come from (0: none)
	void release() {
		words.release(), woff.release(), boff.release(), len.release(), uid.release(), good.release();
		n = 0, gen = 0;
	}
};
std::atomic<uint64_t> g_upload_gen{0};

struct Timer {
	hipEvent_t a = nullptr, b = nullptr;
};

// A device seed index: open-addressing slots over CSR buckets of (tag << 54) | (pos << 40) | read
// entries.
struct IndexBufs {
	DevBuf<Slot> slots;
	DevBuf<uint64_t> entries;
	uint64_t P = 0, total_keys = 0, distinct_keys = 0;
	uint32_t max_probe = 0;
	bool built = false;
	// build scratch, kept between builds of the query index (grow-only): key offsets of the read
	// set they were computed for (src_gen), occurrences, slot owners and counters
	DevBuf<uint64_t> dkoff, occ_hi, occ_lo, occ_entry;
	DevBuf<uint32_t> occ_slot, occ_rank, owner, maxp;
	DevBuf<int> errf;
	DevBuf<unsigned long long> cursor;
	DevBuf<uint8_t> ktmp; // key-offset scan scratch
	uint64_t src_gen = 0, src_nocc = 0;
	int64_t src_n = -1;
	DevIndex view(int k) const {
		return DevIndex{slots.p, entries.p, P, max_probe, k, make_fastmod(P ? P : 1)};
	}
	void release_scratch() {
		dkoff.release(), occ_hi.release(), occ_lo.release(), occ_entry.release();
		occ_slot.release(), occ_rank.release(), owner.release(), maxp.release(), errf.release(), cursor.release();
		ktmp.release();
		src_gen = 0, src_nocc = 0, src_n = -1;
	}
	void release() {
		slots.release(), entries.release();
		release_scratch();
		P = total_keys = distinct_keys = 0, max_probe = 0, built = false;
	}
};

// One pipeline: a stream, its events, device counters and the per-chunk work buffers. The context
// owns one; the serving loop owns one per step it keeps in flight.
struct Worker {
	hipStream_t st = nullptr;
	Timer ev[10];
	hero_stats *acc = nullptr; // host-side timings accumulate here
	DevBuf<uint32_t> seg_n, seg_off, scan_tmp, lane_n;
	DevBuf<uint64_t> seeds, win_info;
	DevBuf<int32_t> cnext, cfirst, clast, clen;
	DevBuf<uint8_t> gflag;
	DevBuf<uint32_t> g_tasks, g_task_off, g_words, g_word_off, g_prog, g_prog_off, prog;
	DevBuf<DpTask> tasks;
	DevBuf<uint32_t> words, nsteps, dp_scratch, wave_scratch, wave_over;
	DevBuf<uint32_t> bucket_hist, bucket_off, order;
	DevBuf<int32_t> g_nch, g_best;
	DevBuf<uint32_t> g_recs, g_rec_off, rec_opn, rec_opoff, rec_ops;
	DevBuf<RecHead> recs;
	DevBuf<int32_t> rec_pos;
	DevBuf<uint32_t> sub_pairs, sub_pair_off;
	DevBuf<VoteItem> votes;
	DevBuf<StepSummary> sums;
	DevBuf<unsigned long long> dstats, dtotal;
	DevBuf<uint64_t> dtotals; // the multi-array scans' totals
	DevBuf<uint32_t> counter;
	// the DP stage captured as a graph (bucketing + every band class + plain form), rebuilt when a
	// buffer it reads moves or the schedule changes; the task count reaches it through ctl[0]
	DevBuf<uint32_t> ctl;
	uint32_t *ctl_host = nullptr; // pinned
	hipGraph_t graph = nullptr;
	hipGraphExec_t graph_exec = nullptr;
	std::vector<uintptr_t> graph_key;
	uint64_t graph_builds = 0, graph_launches = 0;
	unsigned long long skip_seen = 0; // DP tasks skipped as beyond the limits, as last checked
	// serving: one batch of query reads and its seeds
	ReadsDev bq;
	DevBuf<uint8_t> b_cons;
	DevBuf<char> b_raw;
	DevBuf<uint64_t> b_koff;
	HostBuf<uint64_t> h_meta; // per-read base / word / key offsets of the batch
	HostBuf<char> h_bytes;    // gather / scatter of batches whose reads are not back to back
	DevBuf<uint32_t> b_tally, b_cnt, b_off, b_gflag, b_pflag, b_gx, b_px, b_glist, b_plist;
	DevBuf<uint64_t> b_kref, b_keys, b_vals, b_keys2, b_vals2;
	DevBuf<uint8_t> b_sort_tmp;
	// frees the grow-only work buffers (stream, events, counters and the DP graph stay)
	void release_buffers() {
		std::lock_guard<std::recursive_mutex> lk(g_dev_mu);
		if (st) (void) hipStreamSynchronize(st);
		if (graph_exec) (void) hipGraphExecDestroy(graph_exec);
		if (graph) (void) hipGraphDestroy(graph);
		graph_exec = nullptr, graph = nullptr, graph_key.clear();
		seg_n.release(), seg_off.release(), scan_tmp.release(), lane_n.release(), seeds.release(), win_info.release();
		cnext.release(), cfirst.release(), clast.release(), clen.release(), gflag.release();
		g_tasks.release(), g_task_off.release(), g_words.release(), g_word_off.release();
		g_prog.release(), g_prog_off.release(), prog.release();
		tasks.release(), words.release(), nsteps.release(), dp_scratch.release(), wave_scratch.release(), wave_over.release();
		bucket_hist.release(), bucket_off.release(), order.release(), g_nch.release(), g_best.release();
		g_recs.release(), g_rec_off.release(), rec_opn.release(), rec_opoff.release(), rec_ops.release();
		recs.release(), rec_pos.release(), sums.release(), sub_pairs.release(), sub_pair_off.release(), votes.release();
		bq.release(), b_cons.release(), b_raw.release(), b_koff.release();
		b_tally.release(), b_cnt.release(), b_off.release();
		b_gflag.release(), b_pflag.release(), b_gx.release(), b_px.release(), b_glist.release(), b_plist.release();
		b_kref.release(), b_keys.release(), b_vals.release(), b_keys2.release(), b_vals2.release(), b_sort_tmp.release();
	}
	// device bytes the worker's grow-only buffers hold
	size_t device_bytes() const {
		size_t b = 0;
		auto add = [&](const auto &buf) { b += buf.n * sizeof(*buf.p); };
		add(seg_n), add(seg_off), add(scan_tmp), add(lane_n), add(seeds), add(win_info), add(cnext), add(cfirst);
		add(clast), add(clen), add(gflag), add(g_tasks), add(g_task_off), add(g_words), add(g_word_off), add(g_prog);
		add(g_prog_off), add(prog), add(tasks), add(words), add(nsteps), add(dp_scratch), add(wave_scratch), add(wave_over);
		add(bucket_hist), add(bucket_off), add(order), add(g_nch), add(g_best), add(g_recs), add(g_rec_off);
		add(rec_opn), add(rec_opoff), add(rec_ops), add(recs), add(rec_pos), add(sub_pairs), add(sub_pair_off);
		add(votes), add(sums), add(bq.words), add(bq.woff), add(bq.boff), add(bq.len), add(bq.uid), add(bq.good);
		add(b_cons), add(b_raw), add(b_koff), add(b_tally), add(b_cnt), add(b_off), add(b_gflag), add(b_pflag);
		add(b_gx), add(b_px), add(b_glist), add(b_plist), add(b_kref), add(b_keys), add(b_vals), add(b_keys2);
		add(b_vals2), add(b_sort_tmp);
		return b;
	}
	hipError_t init() {
		std::lock_guard<std::recursive_mutex> lk(g_dev_mu);
		hipError_t e = hipStreamCreate(&st);
		for (auto &t : ev)
			if (e == hipSuccess) e = hipEventCreate(&t.a);
		if (e == hipSuccess) e = dstats.exact((size_t) kStatStripes * kStatStride);
		if (e == hipSuccess) e = dtotal.exact(1);
		if (e == hipSuccess) e = dtotals.exact(kScanMulti);
		if (e == hipSuccess) e = counter.exact(1);
		if (e == hipSuccess) e = hipMemset(dstats.p, 0, (size_t) kStatStripes * kStatStride * 8);
		if (e == hipSuccess) e = ctl.exact(kCtlWords);
		if (e == hipSuccess) e = hipHostMalloc((void **) &ctl_host, sizeof(uint32_t) * kCtlWords, hipHostMallocDefault);
		return e;
	}
	void release() {
		std::lock_guard<std::recursive_mutex> lk(g_dev_mu);
		if (st) (void) hipStreamSynchronize(st);
		seg_n.release(), seg_off.release(), scan_tmp.release(), lane_n.release(), seeds.release(), win_info.release();
		cnext.release(), cfirst.release(), clast.release(), clen.release(), gflag.release();
		g_tasks.release(), g_task_off.release(), g_words.release(), g_word_off.release();
		g_prog.release(), g_prog_off.release(), prog.release();
		tasks.release(), words.release(), nsteps.release(), dp_scratch.release(), wave_scratch.release(), wave_over.release();
		bucket_hist.release(), bucket_off.release(), order.release(), g_nch.release(), g_best.release();
		g_recs.release(), g_rec_off.release(), rec_opn.release(), rec_opoff.release(), rec_ops.release();
		recs.release(), rec_pos.release(), sums.release(), dstats.release(), dtotal.release(), dtotals.release();
		counter.release();
		sub_pairs.release(), sub_pair_off.release(), votes.release();
		bq.release(), b_cons.release(), b_raw.release(), b_koff.release(), h_meta.release(), h_bytes.release();
		b_tally.release(), b_cnt.release(), b_off.release();
		b_gflag.release(), b_pflag.release(), b_gx.release(), b_px.release(), b_glist.release(), b_plist.release();
		b_kref.release();
		b_keys.release(), b_vals.release(), b_keys2.release(), b_vals2.release(), b_sort_tmp.release();
		if (graph_exec) (void) hipGraphExecDestroy(graph_exec);
		if (graph) (void) hipGraphDestroy(graph);
		graph_exec = nullptr, graph = nullptr, graph_key.clear();
		ctl.release();
		if (ctl_host) (void) hipHostFree(ctl_host);
		ctl_host = nullptr;
		for (auto &t : ev)
			if (t.a) (void) hipEventDestroy(t.a), t.a = nullptr;
		if (st) (void) hipStreamDestroy(st);
		st = nullptr;
	}
};

} // namespace

struct hero_ctx {
	hero_options opt;
	Opts o;
	std::string err;
	int cu_count = 0;
	// queries (host copy kept for the uncorrected reads of the output)
	std::vector<char> q_seq;
	std::vector<int64_t> q_off;
	std::vector<int32_t> q_len;
	std::vector<uint8_t> q_good, q_active_as_subject;
	std::vector<int64_t> q_not_good, q_uid;
	std::vector<int64_t> q_keep_off; // without a full copy: where each not-good read's bytes sit in q_seq
	bool q_full_copy = false;        // q_seq holds every byte of the caller's layout
	bool q_contiguous = true; // caller offsets are the back-to-back prefix sums of the lengths
	int64_t q_bases = 0;
	DevBuf<uint8_t> cons_out;
	DevBuf<uint32_t> cons_packed; // cons_out as 2-bit codes for the copy to the host
	bool cons_valid = false; // cons_out holds the consensus of the current tallies
	bool votes_clear = false; // the tallies (and the extension tally) are zero or hold this query set's votes
	ReadsDev Qd;
	DevBuf<uint8_t> q_subject_good; // subject predicate over the query set (self-correction)
	// locality order (vote modes unless file_order): the query set as subjects in minimizer order
	// (flags: the subject predicate), each query's tally offset in that order, and the ordered copy
	// of a separate subject set
	bool q_ordered = false;
	ReadsDev Qo, So;
	DevBuf<uint64_t> t_off;
	DevBuf<uint8_t> order_tmp;
	const uint64_t *tally_off() const { return q_ordered ? t_off.p : Qd.boff.p; }
	bool locality_order() const { return o.call_consensus && !opt.records && !opt.file_order; }
	// subjects (separate stream)
	ReadsDev Sd;
	// seed index over the query reads (the reference's orientation)
	IndexBufs qix;
	// serving: resident subject reads and their all-positions index
	ReadsDev Rd;
	std::vector<int32_t> r_len;
	std::vector<uint8_t> r_active;
	int r_max_len = 0;
	IndexBufs six;
	// tallies: 4 u32 per query base (own-base slot difference coded, see VoteSink)
	DevBuf<uint32_t> tally;
	DevBuf<uint32_t> ext; // --merge: kExtPositions x 4 u32 per query read
	// record mode: the accepted records since the last hero_take_records
	std::vector<hero_record> rec_out;
	std::vector<uint32_t> rec_out_ops;
	int64_t rec_subjects = 0; // subjects aligned since the last hero_take_records
	// record mode: .aln text of the subjects aligned since the last hero_take_alignments
	std::string aln;
	// the context's own pipeline (stream, events, chunk buffers)
	Worker w;
	// a second pipeline for subject chunks aligned concurrently with w's (created on first use;
	// its host timings accumulate in st_w2, its device counters in its own stripes)
	Worker w2;
	bool w2_ready = false;
	hero_stats st_w2;
	std::mutex err_mu; // fail() may run on a chunk worker's thread
	// autotuned DP schedule (hero_autotune): tasks per lane and persistent waves per CU per band class
	int tuned_per_lane[kNumBandClasses] = {1, 1, 1, 1, 1, 1, 1, 1};
	int tuned_waves_per_cu[kNumBandClasses] = {0, 0, 0, 0, 0, 0, 0, 0};
	static_assert(kNumBandClasses == 8 && kNumBandClasses <= HERO_TUNE_MAX_CLASSES, "per-class tables");
	hero_tune_report *tune = nullptr; // set while hero_autotune runs
	// BandAligner drop-in (hero_band_align_batch): grow-only device buffers
	DevBuf<hero_band_pair> ba_pairs;
	DevBuf<char> ba_seq, ba_types;
	DevBuf<int64_t> ba_off;
	DevBuf<int32_t> ba_band, ba_n, ba_counts;
	DevBuf<uint32_t> ba_scratch;
	HostBuf<uint64_t> stage; // pinned staging of packed reads and their offsets (upload_reads)
	// stats
	hero_stats st_host;
};

namespace {

// Device scratch of hero_band_align_batch_noalloc: a fixed array in every device's copy of the
// code object, so that entry point allocates nothing. The first kNoAllocCtlBytes hold the plain
// form's work counters (cleared by a kernel on each call); the rest holds the traceback pointers
// of calls that do not fit in LDS. Host code uses the per-device address from
// hipGetSymbolAddress (noalloc_base), never the array's name.
constexpr size_t kNoAllocScratchBytes = (size_t) 64 << 20;
constexpr size_t kNoAllocCtlBytes = 256;
static __device__ __align__(256) uint8_t noalloc_scratch[kNoAllocScratchBytes];
constexpr int kNoAllocMaxDevices = 64;
std::mutex noalloc_mu;
uint8_t *noalloc_base[kNoAllocMaxDevices] = {};

// Looks up this device's address of noalloc_scratch once; the device must be current.
hipError_t noalloc_setup(int device) {
	if (device < 0 || device >= kNoAllocMaxDevices) return hipErrorInvalidDevice;
	std::lock_guard<std::mutex> lk(noalloc_mu);
	if (noalloc_base[device]) return hipSuccess;
	void *ptr = nullptr;
	const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(noalloc_scratch));
	if (e != hipSuccess) return e;
	noalloc_base[device] = (uint8_t *) ptr;
	return hipSuccess;
}

uint8_t *noalloc_scratch_of(int device) {
	if (device < 0 || device >= kNoAllocMaxDevices) return nullptr;
	std::lock_guard<std::mutex> lk(noalloc_mu);
	return noalloc_base[device];
}

// Largest per-chunk window-bucket buffer (8 B per subject window and strand); a chunk beyond it
// probes the index a second time in pass A1 instead.
constexpr uint64_t kWinInfoBytes = 4ull << 30;
// seeds of one chunk before it is split in two (the seed stage indexes them in 31 bits); the
// chunk_seeds_max option lowers it
constexpr uint64_t kChunkSeedsMax = 1ull << 30;
// Subjects per chunk when the caller leaves it automatic, budgeted at 32 KB per subject within 40 %
// of the free device memory (the chunk buffers of the 33 M-read repeat set take about 28 KB per
// subject): the whole set as one chunk when it fits the budget and kChunkMax; otherwise balanced
// chunks sized for two chunk workers at once, never fewer than kChunkMin subjects. Larger chunks
// cut the per-chunk launches, scans and kernel tails (1.67 M C2 subjects: 98.3 ms per step in
// chunks of 262,144, 93.9 ms as one chunk).
constexpr int64_t kChunkMax = 1 << 21, kChunkMin = 1 << 18;
int64_t auto_chunk_subjects(int64_t n) {
	size_t free_b = 0, total_b = 0;
	double budget = 0;
	if (hipMemGetInfo(&free_b, &total_b) == hipSuccess) budget = (double) free_b * 0.4 / 32768.0;
	if (n <= kChunkMax && (double) n <= budget) return std::max<int64_t>(n, 1);
	const int64_t cap = std::max<int64_t>(std::min<int64_t>(kChunkMax, (int64_t) (budget / 2)), kChunkMin);
	const int64_t nch = (n + cap - 1) / cap;
	return (n + nch - 1) / nch;
}

int fail(hero_ctx *c, int code, const char *fmt, ...) {
	char buf[512];
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(buf, sizeof buf, fmt, ap);
	va_end(ap);
	if (c) {
		std::lock_guard<std::mutex> lk(c->err_mu);
		c->err = buf;
	}
	return code;
}
#define HIPCHK(c, expr)                                                                                     \
	do {                                                                                                    \
		hipError_t e_ = (expr);                                                                             \
		if (e_ != hipSuccess) return fail((c), e_ == hipErrorOutOfMemory ? HERO_ERR_NOMEM : HERO_ERR_HIP, "%s: %s (%s:%d)", \
				#expr, hipGetErrorString(e_), __FILE__, __LINE__);                                          \
	} while (0)

Opts make_opts(const hero_options &p) {
	Opts o;
	std::memset(&o, 0, sizeof o);
	o.k = p.k, o.aligner = p.aligner, o.call_consensus = p.call_consensus;
	o.band_override = p.band_override, o.max_candidates = p.max_candidates;
	o.min_overlap = p.min_overlap;
	o.mismatch_rate = p.mismatch_rate, o.indel_rate = p.indel_rate, o.consensus_freq = p.consensus_freq;
	o.dx = p.dx, o.dy = p.dy, o.ex = p.ex, o.sx = p.sx, o.miss = p.miss, o.hit = p.hit, o.unobserved = p.unobserved;
	o.mapping = p.mapping;
	o.long_format = p.long_format, o.merge = p.merge;
	return o;
}

Packed view(const ReadsDev &r, const uint8_t *good) {
	Packed P;
	P.words = r.words.p, P.woff = r.woff.p, P.len = r.len.p, P.uid = r.uid.p, P.good = good, P.n = r.n;
	return P;
}

// Host copy by every host thread (the staging copies below are memory-bound on one thread).
void par_memcpy(char *dst, const char *src, size_t n) {
	constexpr size_t kPiece = (size_t) 256 << 10;
	const int64_t pieces = (int64_t) ((n + kPiece - 1) / kPiece);
#pragma omp parallel for schedule(static) if (pieces > 1)
	for (int64_t p = 0; p < pieces; p++) {
		const size_t o = (size_t) p * kPiece;
		std::memcpy(dst + o, src + o, std::min(kPiece, n - o));
	}
}

// Caller (pageable) memory <-> device through pinned staging in two halves of kStageChunk bytes:
// the link moves one half while the host threads fill or drain the other, which runs near the
// link's rate where the runtime's own pageable path does not. Both return once the last byte has
// landed (the staging is free again); the stream's earlier work finishes first.
constexpr size_t kStageChunk = (size_t) 8 << 20;
constexpr size_t kStageMin = (size_t) 1 << 20; // smaller copies go straight through the runtime
template <class T> int staged_d2h(hero_ctx *c, hipStream_t st, HostBuf<T> &stg, char *dst, const char *src, size_t n) {
	if (n < kStageMin) {
		if (n) HIPCHK(c, hipMemcpyAsync(dst, src, n, hipMemcpyDeviceToHost, st));
		HIPCHK(c, hipStreamSynchronize(st));
		return HERO_OK;
	}
	HIPCHK(c, stg.ensure((2 * kStageChunk + sizeof(T) - 1) / sizeof(T)));
	char *half[2] = {(char *) stg.p, (char *) stg.p + kStageChunk};
	const size_t chunks = (n + kStageChunk - 1) / kStageChunk;
	HIPCHK(c, hipMemcpyAsync(half[0], src, std::min(kStageChunk, n), hipMemcpyDeviceToHost, st));
	for (size_t i = 0; i < chunks; i++) {
		HIPCHK(c, hipStreamSynchronize(st)); // chunk i is in half[i % 2]
		if (i + 1 < chunks) {
			const size_t o = (i + 1) * kStageChunk;
			HIPCHK(c, hipMemcpyAsync(half[(i + 1) % 2], src + o, std::min(kStageChunk, n - o), hipMemcpyDeviceToHost, st));
		}
		const size_t o = i * kStageChunk;
		par_memcpy(dst + o, half[i % 2], std::min(kStageChunk, n - o));
	}
	return HERO_OK;
}
// Corrected bases packed by launch_pack_bases (2-bit codes, four per byte) -> ASCII at dst, through
// the pinned staging halves: the link moves a quarter of the bytes and the host threads unpack one
// half while the other is on the link.
template <class T> int staged_d2h_bases(hero_ctx *c, hipStream_t st, HostBuf<T> &stg, char *dst, const uint32_t *packed,
		size_t nbases) {
	static const struct Lut {
		uint32_t v[256];
		Lut() {
			const char base[4] = {'A', 'C', 'T', 'G'}; // the 2-bit codes' bases
			for (int b = 0; b < 256; b++) {
				uint32_t w = 0;
				for (int i = 0; i < 4; i++) w |= (uint32_t) (uint8_t) base[(b >> (2 * i)) & 3] << (8 * i);
				v[b] = w;
			}
		}
	} lut;
	const size_t nbytes = (nbases + 3) / 4;
	if (nbytes == 0) return HERO_OK;
	HIPCHK(c, stg.ensure((2 * kStageChunk + sizeof(T) - 1) / sizeof(T)));
	const uint8_t *src = (const uint8_t *) packed;
	uint8_t *half[2] = {(uint8_t *) stg.p, (uint8_t *) stg.p + kStageChunk};
	const size_t chunks = (nbytes + kStageChunk - 1) / kStageChunk;
	HIPCHK(c, hipMemcpyAsync(half[0], src, std::min(kStageChunk, nbytes), hipMemcpyDeviceToHost, st));
	for (size_t i = 0; i < chunks; i++) {
		HIPCHK(c, hipStreamSynchronize(st)); // chunk i is in half[i % 2]
		if (i + 1 < chunks) {
			const size_t o = (i + 1) * kStageChunk;
			HIPCHK(c, hipMemcpyAsync(half[(i + 1) % 2], src + o, std::min(kStageChunk, nbytes - o), hipMemcpyDeviceToHost, st));
		}
		const size_t o = i * kStageChunk, m = std::min(kStageChunk, nbytes - o);
		const uint8_t *h = half[i % 2];
		char *d = dst + 4 * o;
		const size_t bases = std::min(4 * m, nbases - 4 * o);
		const int64_t full = (int64_t) (bases / 4);
#pragma omp parallel for schedule(static) if (full > 65536)
		for (int64_t b = 0; b < full; b++) std::memcpy(d + 4 * b, &lut.v[h[b]], 4);
		for (size_t r = (size_t) full * 4; r < bases; r++) d[r] = (char) ((lut.v[h[r / 4]] >> (8 * (r % 4))) & 0xFF);
	}
	return HERO_OK;
}
template <class T> int staged_h2d(hero_ctx *c, hipStream_t st, HostBuf<T> &stg, char *dst, const char *src, size_t n) {
	if (n < kStageMin) {
		if (n) HIPCHK(c, hipMemcpyAsync(dst, src, n, hipMemcpyHostToDevice, st));
		HIPCHK(c, hipStreamSynchronize(st));
		return HERO_OK;
	}
	HIPCHK(c, stg.ensure((2 * kStageChunk + sizeof(T) - 1) / sizeof(T)));
	char *half[2] = {(char *) stg.p, (char *) stg.p + kStageChunk};
	const size_t chunks = (n + kStageChunk - 1) / kStageChunk;
	for (size_t i = 0; i < chunks; i++) {
		const size_t o = i * kStageChunk, m = std::min(kStageChunk, n - o);
		par_memcpy(half[i % 2], src + o, m); // chunk i - 1 is on the link meanwhile
		HIPCHK(c, hipStreamSynchronize(st)); // chunk i - 1 has landed: its half is free next
		HIPCHK(c, hipMemcpyAsync(dst + o, half[i % 2], m, hipMemcpyHostToDevice, st));
	}
	HIPCHK(c, hipStreamSynchronize(st));
	return HERO_OK;
}

// Host passes over a read set's arrays, by every host thread: blocks of reads, each block's sums
// first, then each block's own prefix sums from its offset.
constexpr int64_t kHostBlock = 1 << 16;
inline int64_t host_blocks(int64_t n) { return std::max<int64_t>(1, (n + kHostBlock - 1) / kHostBlock); }
template <class T> void par_copy(T *dst, const T *src, int64_t n) {
	const int64_t nblk = host_blocks(n);
#pragma omp parallel for schedule(static)
	for (int64_t b = 0; b < nblk; b++) {
		const int64_t i0 = b * kHostBlock, i1 = std::min<int64_t>(n, i0 + kHostBlock);
		if (i1 > i0) std::memcpy(dst + i0, src + i0, (size_t) (i1 - i0) * sizeof(T));
	}
}
// woff / boff: each read's first word and first base (exclusive prefix sums of ceil(len / 32) and
// len), entry n the totals; when the outputs are null only the totals are computed.
void read_offsets(const int32_t *len, int64_t n, uint64_t *woff, uint64_t *boff, uint64_t &nw, uint64_t &nb) {
	const int64_t nblk = host_blocks(n);
	std::vector<uint64_t> bw((size_t) nblk + 1, 0), bb((size_t) nblk + 1, 0);
#pragma omp parallel for schedule(static)
	for (int64_t b = 0; b < nblk; b++) {
		uint64_t sw = 0, sb = 0;
		for (int64_t i = b * kHostBlock, e = std::min<int64_t>(n, (b + 1) * kHostBlock); i < e; i++)
			sw += (uint64_t) (len[i] + 31) / 32, sb += (uint64_t) len[i];
		bw[(size_t) b + 1] = sw, bb[(size_t) b + 1] = sb;
	}
	for (int64_t b = 0; b < nblk; b++) bw[(size_t) b + 1] += bw[(size_t) b], bb[(size_t) b + 1] += bb[(size_t) b];
	nw = bw[(size_t) nblk], nb = bb[(size_t) nblk];
	if (!woff) return;
#pragma omp parallel for schedule(static)
	for (int64_t b = 0; b < nblk; b++) {
		uint64_t w = bw[(size_t) b], x = bb[(size_t) b];
		for (int64_t i = b * kHostBlock, e = std::min<int64_t>(n, (b + 1) * kHostBlock); i < e; i++) {
			woff[i] = w, boff[i] = x;
			w += (uint64_t) (len[i] + 31) / 32, x += (uint64_t) len[i];
		}
	}
	woff[n] = nw, boff[n] = nb;
}

// With `flags`, the reads' flags (read_flags) are computed in the packing pass into flags->good and
// flags->active, and flags->good is the good array uploaded (`good` is not read).
struct ReadFlagsOut {
	int64_t min_overlap;
	int k;
	uint8_t *good, *active;
};
int upload_reads_on(hero_ctx *c, hipStream_t st, ReadsDev &d, int64_t n, const char *seq, const int64_t *off,
		const int32_t *len, const int64_t *uid, const std::vector<uint8_t> &good, const ReadFlagsOut *flags = nullptr);
int upload_reads(hero_ctx *c, ReadsDev &d, int64_t n, const char *seq, const int64_t *off, const int32_t *len,
		const int64_t *uid, const std::vector<uint8_t> &good, const ReadFlagsOut *flags = nullptr) {
	return upload_reads_on(c, c->w.st, d, n, seq, off, len, uid, good, flags);
}
int upload_reads_on(hero_ctx *c, hipStream_t st, ReadsDev &d, int64_t n, const char *seq, const int64_t *off,
		const int32_t *len, const int64_t *uid, const std::vector<uint8_t> &good, const ReadFlagsOut *flags) {
	// everything goes through pinned staging, so the upload runs at the link's rate:
	// [words | woff | boff | uid | len | good], the 2-bit words packed by every host thread
	uint64_t nw = 0, nb = 0;
	read_offsets(len, n, nullptr, nullptr, nw, nb);
	// DP tasks locate their segments by u32 word indices (DpTask::s_word, q_word)
	if (nw + 2 > 0xFFFFFFFFull) return fail(c, HERO_ERR_LIMIT, "read set of %llu packed words exceeds 2^32 - 1", (unsigned long long) nw);
	const size_t words = (size_t) nw + 2; // two zero words of tail padding
	const size_t n1 = (size_t) n + 1, len_words = (n1 * 4 + 7) / 8, good_words = (n1 + 7) / 8;
	HIPCHK(c, c->stage.ensure(words + 3 * n1 + len_words + good_words));
	uint64_t *hw = c->stage.p, *hwoff = hw + words, *hboff = hwoff + n1, *huid = hboff + n1;
	int32_t *hlen = (int32_t *) (huid + n1);
	uint8_t *hgood = (uint8_t *) (huid + n1 + len_words);
	read_offsets(len, n, hwoff, hboff, nw, nb);
	if (flags)
		pack_reads_flags_into(seq, off, len, n, hwoff, hw, words, flags->min_overlap, flags->k, flags->good, flags->active);
	else
		pack_reads_into(seq, off, len, n, hwoff, hw, words);
	if (n) {
		par_copy(huid, (const uint64_t *) uid, n);
		par_copy(hlen, len, n);
		par_copy(hgood, flags ? flags->good : good.data(), n);
	}
	// grow-only buffers: a steady stream of batches allocates nothing after the largest one
	HIPCHK(c, d.words.ensure(words));
	HIPCHK(c, d.woff.ensure((size_t) n + 1));
	HIPCHK(c, d.boff.ensure((size_t) n + 1));
	HIPCHK(c, d.len.ensure((size_t) n + 1));
	HIPCHK(c, d.uid.ensure((size_t) n + 1));
	HIPCHK(c, d.good.ensure((size_t) n + 1));
	HIPCHK(c, hipMemcpyAsync(d.words.p, hw, words * 8, hipMemcpyHostToDevice, st));
	if (n) {
		HIPCHK(c, hipMemcpyAsync(d.woff.p, hwoff, (size_t) n * 8, hipMemcpyHostToDevice, st));
		HIPCHK(c, hipMemcpyAsync(d.boff.p, hboff, (size_t) n * 8, hipMemcpyHostToDevice, st));
		HIPCHK(c, hipMemcpyAsync(d.len.p, hlen, (size_t) n * 4, hipMemcpyHostToDevice, st));
		HIPCHK(c, hipMemcpyAsync(d.uid.p, huid, (size_t) n * 8, hipMemcpyHostToDevice, st));
		HIPCHK(c, hipMemcpyAsync(d.good.p, hgood, (size_t) n, hipMemcpyHostToDevice, st));
	}
	HIPCHK(c, hipStreamSynchronize(st));
	d.n = n;
	d.gen = ++g_upload_gen;
	return HERO_OK;
}

// The locality-ordered copy `out` of read set R (launch_order): flags from `flag`, keys over the
// reads `active` marks; t_off, when given, receives each read's tally offset in that order.
int order_reads(hero_ctx *c, hipStream_t st, const ReadsDev &R, const uint8_t *active, const uint8_t *flag,
		const std::vector<int32_t> &len, ReadsDev &out, uint64_t *t_off) {
	const int64_t n = R.n;
	uint64_t nw = 0;
	for (int64_t i = 0; i < n; i++) nw += (uint64_t) (len[(size_t) i] + 31) / 32;
	HIPCHK(c, out.words.ensure((size_t) nw + 2));
	HIPCHK(c, out.woff.ensure((size_t) n + 1));
	HIPCHK(c, out.len.ensure((size_t) n + 1));
	HIPCHK(c, out.uid.ensure((size_t) n + 1));
	HIPCHK(c, out.good.ensure((size_t) n + 1));
	HIPCHK(c, c->order_tmp.ensure(order_tmp_bytes(n)));
	OrderOut o{out.words.p, out.woff.p, out.len.p, out.uid.p, out.good.p, t_off, nullptr};
	const Packed P{R.words.p, R.woff.p, R.len.p, R.uid.p, flag, n};
	HIPCHK(c, launch_order(P, active, flag, o, c->order_tmp.p, c->order_tmp.n, st));
	HIPCHK(c, hipMemsetAsync(out.words.p + nw, 0, 16, st)); // the two zero words of tail padding
	HIPCHK(c, hipStreamSynchronize(st));
	out.n = n;
	return HERO_OK;
}

int check_reads(hero_ctx *c, int64_t n, const char *seq, const int64_t *off, const int32_t *len, const int64_t *uid) {
	if (n < 0 || (n > 0 && (!seq || !off || !len || !uid))) return fail(c, HERO_ERR_ARG, "null read arrays");
	if (n >= (int64_t) 1 << 32) return fail(c, HERO_ERR_LIMIT, "at most 2^32-1 reads per set");
	int64_t bad = n; // the first read out of the length limits
#pragma omp parallel for reduction(min : bad) schedule(static, 65536)
	for (int64_t i = 0; i < n; i++)
		if (len[i] < 0 || len[i] > kMaxReadLen) bad = std::min<int64_t>(bad, i);
	if (bad < n) return fail(c, HERO_ERR_LIMIT, "read %lld has length %d (limit %d)", (long long) bad, len[bad], kMaxReadLen);
	return HERO_OK;
}

double elapsed(Worker *w, int a, int b) {
	float ms = 0;
	(void) hipEventElapsedTime(&ms, w->ev[a].a, w->ev[b].a);
	return ms;
}

// One subject chunk [s0, s0 + nsub) of the packed subject set S.
int format_records(hero_ctx *c, Worker *w, int64_t s0, int32_t nsub, const int64_t *s_uid, uint32_t nslots);
int run_from_seeds(hero_ctx *c, Worker *w, const Packed &S, const Packed &Q, const DevIndex &ix, ChunkDev &cd,
		uint64_t nseeds, int rows_needed, const uint64_t *q_boff, uint32_t *tally, uint32_t *diff, const int64_t *s_uid,
		bool rec_mode);

// Exclusive scan of n u32 on the worker's stream; the total comes back to the host.
int wscan(hero_ctx *c, Worker *w, const uint32_t *in, uint32_t *out, int64_t n, uint64_t &total) {
	HIPCHK(c, w->scan_tmp.ensure(scan_tmp_words(n)));
	HIPCHK(c, scan_u32(in, out, n, w->dtotal.p, w->scan_tmp.p, w->scan_tmp.n, w->st));
	unsigned long long t = 0;
	HIPCHK(c, hipMemcpyAsync(&t, w->dtotal.p, 8, hipMemcpyDeviceToHost, w->st));
	HIPCHK(c, hipStreamSynchronize(w->st));
	total = t;
	return HERO_OK;
}

int worker_stat(hero_ctx *c, Worker *w, int slot, uint64_t &out);

// Exclusive scans of K arrays of n u32 in one pass on the worker's stream; their 64-bit totals
// come back to the host with one synchronization.
int wscan_multi(hero_ctx *c, Worker *w, int K, const uint32_t *const *in, uint32_t *const *out, int64_t n,
		uint64_t *totals) {
	const size_t bytes = scan_multi_tmp_bytes(K, n);
	HIPCHK(c, w->scan_tmp.ensure((bytes + 3) / 4));
	HIPCHK(c, scan_u32_multi(K, in, out, n, w->dtotals.p, w->scan_tmp.p, w->scan_tmp.n * 4, w->st));
	HIPCHK(c, hipMemcpyAsync(totals, w->dtotals.p, (size_t) K * 8, hipMemcpyDeviceToHost, w->st));
	HIPCHK(c, hipStreamSynchronize(w->st));
	return HERO_OK;
}

// Autotuner step for one band class: runs the class's tasks under every candidate schedule (one
// or two tasks per lane; persistent waves per CU from 1 to the occupancy limit), each timed with
// HIP events after an untimed run, and keeps the fastest. Every candidate computes the same
// results; only the schedule differs.
int tune_class(hero_ctx *c, Worker *w, const DpLaunch &Lc, int cls, int bm, bool aff, int rows, int nw) {
	hero_tune_report &rep = *c->tune;
	hero_tune_class &tc = rep.cls[cls];
	tc.band = bm, tc.tasks = (int64_t) Lc.ntasks;
	float best = -1.0f;
	uint64_t cells0 = 0, cells1 = 0;
	if (int r = worker_stat(c, w, kStCellsAll, cells0)) return r;
	for (int per_lane = 1; per_lane <= (pk_form_fits(c->o, rows, bm, Lc.tagged != 0) ? 2 : 1); per_lane++) {
		const int occ = dp_batched_waves_per_cu(bm, aff, per_lane, rows, Lc.tagged != 0);
		for (int wpc = 1; wpc <= occ; wpc *= 2) {
			const int wpc_eff = wpc * 2 > occ ? occ : wpc; // the last candidate is the occupancy limit
			const int waves = (int) std::min<uint64_t>((uint64_t) c->cu_count * wpc_eff,
					(Lc.ntasks + 64 * per_lane - 1) / (64 * per_lane));
			HIPCHK(c, w->dp_scratch.ensure((size_t) waves * rows * nw * 64 * per_lane));
			DpLaunch L2 = Lc;
			L2.scratch = w->dp_scratch.p;
			L2.ptr_lds_bytes = per_lane == 1 ? dp_batched_lds_bytes(bm, aff, Lc.tagged != 0, wpc_eff) : 0u;
			float ms = 0;
			for (int rep_i = 0; rep_i < 2; rep_i++) {
				HIPCHK(c, hipMemsetAsync(w->counter.p, 0, 4, w->st));
				HIPCHK(c, hipEventRecord(w->ev[9].a, w->st));
				HIPCHK(c, launch_dp_batched(L2, bm, per_lane, waves, w->st));
				HIPCHK(c, hipEventRecord(w->ev[3].a, w->st));
				HIPCHK(c, hipStreamSynchronize(w->st));
				HIPCHK(c, hipEventElapsedTime(&ms, w->ev[9].a, w->ev[3].a));
			}
			if (tc.ncand == 0) { // band cells of the class: the first candidate's two runs
				if (int r = worker_stat(c, w, kStCellsAll, cells1)) return r;
				tc.cells = (int64_t) ((cells1 - cells0) / 2);
			}
			if (per_lane == 1 && wpc_eff == dp_batched_waves_per_cu(bm, aff, 1, rows, Lc.tagged != 0)) tc.ms_default = ms;
			if (tc.ncand < (int) (sizeof tc.cand / sizeof tc.cand[0]))
				tc.cand[tc.ncand++] = hero_tune_candidate{per_lane, wpc_eff, ms};
			if (best < 0 || ms < best) best = ms, tc.per_lane = per_lane, tc.waves_per_cu = wpc_eff, tc.ms = ms;
			if (wpc_eff == occ) break;
		}
	}
	c->tuned_per_lane[cls] = tc.per_lane;
	c->tuned_waves_per_cu[cls] = tc.waves_per_cu;
	return HERO_OK;
}

// Plain-form (wave) launches for calls whose sides have at most max_len codes: both sides' codes
// and, when they fit the LDS share, the pointers of a call in LDS; otherwise its pointers go to the
// wave's slice of wave_scratch, sized for the widest band the plain form takes (kWaveMaxBand).
// Waves are limited so that the scratch stays within kWaveScratchBytes.
constexpr size_t kWaveLdsShare = 32768;
constexpr size_t kWaveScratchBytes = (size_t) 1 << 30;
WaveSizing wave_sizing(int max_len, int max_band, int want_waves) {
	WaveSizing z;
	std::memset(&z, 0, sizeof z);
	z.seq_cap = (2 * std::max(max_len, 1) + 15) & ~15;
	const int bmax = std::min(std::max(max_band, 0), kWaveMaxBand);
	const size_t share = std::max(kWaveLdsShare, (size_t) z.seq_cap + 8192) - (size_t) z.seq_cap;
	z.ptr_lds_words = (int32_t) std::min(share / 4, wave_call_ptr_words(max_len, std::min(bmax, wave_band_hi(0))));
	const size_t need = wave_call_ptr_words(max_len, bmax);
	z.scratch_words = need > (size_t) z.ptr_lds_words ? need : 0;
	z.nwaves = std::max(1, want_waves);
	if (z.scratch_words)
		z.nwaves = (int) std::max<size_t>(1, std::min<size_t>((size_t) z.nwaves, kWaveScratchBytes / (z.scratch_words * 4)));
	return z;
}

// The three band-range launches of the plain form over Lp's tasks; ctl_wave (kWaveCtlWords) must be
// zero before the first.
hipError_t launch_wave_stage(const DpLaunch &Lp, const WaveSizing &z, uint32_t *ctl_wave, uint32_t *over, hipStream_t st) {
	for (int l = 0; l < kWaveLaunches; l++)
		if (hipError_t e = launch_dp_wave_tasks(Lp, z, l, ctl_wave, over, st); e != hipSuccess) return e;
	return hipSuccess;
}

// One device counter of a worker, summed over its stripes (waits for the worker's stream).
int worker_stat(hero_ctx *c, Worker *w, int slot, uint64_t &out) {
	std::vector<unsigned long long> stripes((size_t) kStatStripes * kStatStride);
	HIPCHK(c, hipMemcpyAsync(stripes.data(), w->dstats.p, stripes.size() * 8, hipMemcpyDeviceToHost, w->st));
	HIPCHK(c, hipStreamSynchronize(w->st));
	out = 0;
	for (int i = 0; i < kStatStripes; i++) out += stripes[(size_t) i * kStatStride + slot];
	return HERO_OK;
}

// Tasks the DP skipped since the last check (beyond kWaveMaxBand or the read-length limits).
int check_dp_skipped(hero_ctx *c, Worker *w) {
	std::vector<unsigned long long> stripes((size_t) kStatStripes * kStatStride);
	HIPCHK(c, hipMemcpyAsync(stripes.data(), w->dstats.p, stripes.size() * 8, hipMemcpyDeviceToHost, w->st));
	HIPCHK(c, hipStreamSynchronize(w->st));
	unsigned long long skipped = 0;
	for (int i = 0; i < kStatStripes; i++) skipped += stripes[(size_t) i * kStatStride + kStDpSkipped];
	const unsigned long long fresh = skipped - w->skip_seen;
	w->skip_seen = skipped;
	if (fresh) return fail(c, HERO_ERR_LIMIT, "%llu DP calls have a band above %d", fresh, kWaveMaxBand);
	return HERO_OK;
}

// The DP of a chunk's tasks (cd.tasks): batched form by band class and length bucket, plain form
// for the rest (or for everything with dp_form 1). Records ev[7] after bucketing, ev[3] at the end.
int run_dp_graph(hero_ctx *c, Worker *w, const Packed &S, const Packed &Q, const ChunkDev &cd, uint32_t ntasks,
		int rows_needed);

int run_dp(hero_ctx *c, Worker *w, const Packed &S, const Packed &Q, const ChunkDev &cd, uint32_t ntasks,
		int rows_needed) {
	if (c->opt.dp_form == 0 && !c->opt.dp_eager && !c->tune && ntasks > 0)
		return run_dp_graph(c, w, S, Q, cd, ntasks, rows_needed);
	unsigned long long *stats = w->dstats.p;
	DpLaunch L;
	std::memset(&L, 0, sizeof L);
	L.tasks = cd.tasks, L.ntasks = (uint32_t) ntasks, L.order = nullptr, L.S = S, L.Q = Q, L.o = c->o;
	L.words = cd.words, L.nsteps = cd.nsteps, L.sums = cd.sums, L.stats = stats, L.batch_counter = w->counter.p;
	const bool aff = c->o.aligner == 0;
	const int rows = std::min(rows_needed, 512);
	const int plain_waves = c->opt.dp_waves > 0 ? c->opt.dp_waves : c->cu_count * 8;
	L.tagged = !c->opt.dp_untagged && dp_tagged_form(c->o, rows) ? 1 : 0;
	uint32_t plain_begin = 0, plain_n = (uint32_t) ntasks;
	if (c->opt.dp_form == 0 && ntasks > 0) {
		// batched form: bucket tasks by (band class, descending length), then one persistent launch
		// per class at the smallest band width that holds it
		HIPCHK(c, w->bucket_hist.ensure(kNumBuckets));
		HIPCHK(c, w->bucket_off.ensure(kNumBuckets));
		HIPCHK(c, w->order.ensure((size_t) ntasks));
		HIPCHK(c, hipMemsetAsync(w->bucket_hist.p, 0, kNumBuckets * 4, w->st));
		HIPCHK(c, launch_task_histogram(cd.tasks, (uint32_t) ntasks, nullptr, rows, w->bucket_hist.p, w->st));
		uint64_t total = 0;
		if (int r = wscan(c, w, w->bucket_hist.p, w->bucket_off.p, kNumBuckets, total)) return r;
		std::vector<uint32_t> off(kNumBuckets);
		HIPCHK(c, hipMemcpyAsync(off.data(), w->bucket_off.p, kNumBuckets * 4, hipMemcpyDeviceToHost, w->st));
		HIPCHK(c, hipStreamSynchronize(w->st));
		HIPCHK(c, launch_task_scatter(cd.tasks, (uint32_t) ntasks, nullptr, rows, w->bucket_off.p, w->order.p, w->st));
		HIPCHK(c, hipEventRecord(w->ev[7].a, w->st));
		for (int cls = 0; cls <= kNumBandClasses; cls++) {
			const uint32_t b0 = off[(size_t) cls * kLenBins];
			const uint32_t b1 = cls + 1 <= kNumBandClasses ? off[(size_t) (cls + 1) * kLenBins] : (uint32_t) ntasks;
			if (cls == kNumBandClasses) {
				plain_begin = b0, plain_n = b1 - b0;
				break;
			}
			if (b1 == b0) continue;
			const int bm = kBandClass[cls], nw = (2 * bm + 1 + 15) / 16;
			// tasks per lane: 2 = packed 16-bit (band classes up to kPk16MaxBand), 1 = 32-bit; the
			// option, else the tuned choice for this class, else 1
			const int want = c->opt.dp_per_lane > 0 ? c->opt.dp_per_lane : c->tuned_per_lane[cls];
			const int per_lane = (want == 2 && pk_form_fits(c->o, rows, bm, L.tagged != 0)) ? 2 : 1;
			const int wpc = c->tuned_waves_per_cu[cls] > 0 ? c->tuned_waves_per_cu[cls]
			                                               : dp_batched_waves_per_cu(bm, aff, per_lane, rows, L.tagged != 0);
			int waves = c->opt.dp_waves > 0 ? c->opt.dp_waves : c->cu_count * wpc;
			waves = (int) std::min<uint64_t>((uint64_t) waves, (b1 - b0 + 64 * per_lane - 1) / (64 * per_lane));
			HIPCHK(c, w->dp_scratch.ensure((size_t) waves * rows * nw * 64 * per_lane));
			DpLaunch Lc = L;
			Lc.order = w->order.p + b0, Lc.ntasks = b1 - b0;
			Lc.scratch = w->dp_scratch.p, Lc.rows_max = rows;
			// the LDS share of the waves per CU the launch asks for (an explicit wave count included)
			const int wpc_launch = c->opt.dp_waves > 0 ? (c->opt.dp_waves + c->cu_count - 1) / c->cu_count : wpc;
			Lc.ptr_lds_bytes = per_lane == 1 ? dp_batched_lds_bytes(bm, aff, L.tagged != 0, wpc_launch) : 0u;
			if (c->tune) {
				// autotuning: time every (tasks per lane, waves per CU) candidate on this class's tasks
				if (int r = tune_class(c, w, Lc, cls, bm, aff, rows, nw)) return r;
				continue;
			}
			HIPCHK(c, hipMemsetAsync(w->counter.p, 0, 4, w->st));
			HIPCHK(c, launch_dp_batched(Lc, bm, per_lane, waves, w->st));
		}
	} else {
		HIPCHK(c, hipEventRecord(w->ev[7].a, w->st));
	}
	if (plain_n > 0) {
		// plain form: every task (dp_form 1), or the bucket beyond the batched band classes
		const WaveSizing z = wave_sizing(rows_needed - 1, kWaveMaxBand, plain_waves);
		HIPCHK(c, w->wave_scratch.ensure(z.scratch_words * (size_t) z.nwaves + 1));
		HIPCHK(c, w->wave_over.ensure((size_t) plain_n));
		WaveSizing zz = z;
		zz.scratch = z.scratch_words ? w->wave_scratch.p : nullptr;
		DpLaunch Lp = L;
		Lp.order = c->opt.dp_form == 0 ? w->order.p + plain_begin : nullptr;
		Lp.ntasks = plain_n;
		uint32_t *ctl_wave = w->ctl.p + kCtlWave;
		HIPCHK(c, hipMemsetAsync(ctl_wave, 0, kWaveCtlWords * 4, w->st));
		HIPCHK(c, launch_wave_stage(Lp, zz, ctl_wave, w->wave_over.p, w->st));
	}
	HIPCHK(c, hipEventRecord(w->ev[3].a, w->st));
	return check_dp_skipped(c, w);
}

// The DP stage as one graph launch. Fixed grids (persistent waves per class, grid-stride
// bucketing over the task buffer's capacity); the task count and the class ranges are read on the
// device, so one captured graph serves every chunk whose buffers have not moved.
int run_dp_graph(hero_ctx *c, Worker *w, const Packed &S, const Packed &Q, const ChunkDev &cd, uint32_t ntasks,
		int rows_needed) {
	const bool aff = c->o.aligner == 0;
	// rows rounded up so that batches of slightly different read lengths share one graph (the row
	// limit only routes tasks between the two DP forms, which give identical results)
	const int rows = std::min((rows_needed + 63) / 64 * 64, 512);
	const int plain_waves = c->opt.dp_waves > 0 ? c->opt.dp_waves : c->cu_count * 8;
	const uint32_t cap = (uint32_t) std::min<size_t>(w->tasks.n, 0xFFFFFFFFu);
	if (cd.tasks != w->tasks.p || ntasks > cap) return fail(c, HERO_ERR_STATE, "DP graph: task buffer mismatch");
	HIPCHK(c, w->bucket_hist.ensure(kNumBuckets));
	HIPCHK(c, w->bucket_off.ensure(kNumBuckets));
	HIPCHK(c, w->order.ensure(cap));
	HIPCHK(c, w->scan_tmp.ensure(scan_tmp_words(kNumBuckets)));
	int per_lane[kNumBandClasses], waves[kNumBandClasses];
	uint32_t lds[kNumBandClasses];
	size_t scratch = 0;
	const bool tagged = !c->opt.dp_untagged && dp_tagged_form(c->o, rows);
	for (int cls = 0; cls < kNumBandClasses; cls++) {
		const int bm = kBandClass[cls], nw = (2 * bm + 1 + 15) / 16;
		const int want = c->opt.dp_per_lane > 0 ? c->opt.dp_per_lane : c->tuned_per_lane[cls];
		per_lane[cls] = (want == 2 && pk_form_fits(c->o, rows, bm, tagged)) ? 2 : 1;
		const int wpc = c->tuned_waves_per_cu[cls] > 0 ? c->tuned_waves_per_cu[cls]
		                                               : dp_batched_waves_per_cu(bm, aff, per_lane[cls], rows, tagged);
		waves[cls] = c->opt.dp_waves > 0 ? c->opt.dp_waves : c->cu_count * wpc;
		const int wpc_launch = c->opt.dp_waves > 0 ? (c->opt.dp_waves + c->cu_count - 1) / c->cu_count : wpc;
		lds[cls] = per_lane[cls] == 1 ? dp_batched_lds_bytes(bm, aff, tagged, wpc_launch) : 0u;
		scratch = std::max(scratch, (size_t) waves[cls] * rows * nw * 64 * per_lane[cls]);
	}
	HIPCHK(c, w->dp_scratch.ensure(scratch));
	WaveSizing z = wave_sizing((rows_needed + 63) / 64 * 64 - 1, kWaveMaxBand, plain_waves);
	HIPCHK(c, w->wave_scratch.ensure(z.scratch_words * (size_t) z.nwaves + 1));
	HIPCHK(c, w->wave_over.ensure(cap));
	z.scratch = z.scratch_words ? w->wave_scratch.p : nullptr;
	DpLaunch L;
	std::memset(&L, 0, sizeof L); // padding too: the key below compares the struct's bytes
	L.tasks = cd.tasks, L.order = w->order.p, L.S = S, L.Q = Q, L.o = c->o;
	L.words = cd.words, L.nsteps = cd.nsteps, L.sums = cd.sums, L.stats = w->dstats.p;
	L.rows_max = rows;
	L.tagged = tagged ? 1 : 0;
	// everything the captured launches take by value
	std::vector<uintptr_t> key((sizeof L + sizeof(uintptr_t) - 1) / sizeof(uintptr_t), 0);
	std::memcpy(key.data(), &L, sizeof L);
	for (uintptr_t v : {(uintptr_t) cap, (uintptr_t) w->dp_scratch.p, (uintptr_t) w->wave_scratch.p, (uintptr_t) w->wave_over.p,
				 (uintptr_t) z.seq_cap, (uintptr_t) z.ptr_lds_words, (uintptr_t) z.scratch_words, (uintptr_t) z.nwaves,
				 (uintptr_t) w->bucket_hist.p, (uintptr_t) w->bucket_off.p, (uintptr_t) w->scan_tmp.p,
				 (uintptr_t) w->scan_tmp.n, (uintptr_t) w->dtotal.p, (uintptr_t) w->ctl.p, (uintptr_t) plain_waves})
		key.push_back(v);
	for (int cls = 0; cls < kNumBandClasses; cls++) {
		key.push_back((uintptr_t) (per_lane[cls] * 100000 + waves[cls]));
		key.push_back((uintptr_t) lds[cls]);
	}
	if (!w->graph_exec || key != w->graph_key) {
		std::lock_guard<std::recursive_mutex> lk(g_dev_mu); // no allocation or free elsewhere meanwhile
		if (w->graph_exec) {
			HIPCHK(c, hipStreamSynchronize(w->st)); // the old graph has finished before it goes
			(void) hipGraphExecDestroy(w->graph_exec);
		}
		if (w->graph) (void) hipGraphDestroy(w->graph);
		w->graph_exec = nullptr, w->graph = nullptr, w->graph_key.clear();
		uint32_t *ctl = w->ctl.p;
		HIPCHK(c, hipStreamBeginCapture(w->st, hipStreamCaptureModeThreadLocal));
		hipError_t e = launch_fill_u32(w->bucket_hist.p, kNumBuckets, 0u, w->st);
		if (e == hipSuccess) e = launch_task_histogram(cd.tasks, cap, ctl, rows, w->bucket_hist.p, w->st);
		if (e == hipSuccess)
			e = scan_u32(w->bucket_hist.p, w->bucket_off.p, kNumBuckets, w->dtotal.p, w->scan_tmp.p, w->scan_tmp.n, w->st);
		if (e == hipSuccess) e = launch_class_ranges(w->bucket_off.p, ctl, w->st);
		if (e == hipSuccess) e = launch_task_scatter(cd.tasks, cap, ctl, rows, w->bucket_off.p, w->order.p, w->st);
		for (int cls = 0; cls < kNumBandClasses && e == hipSuccess; cls++) {
			DpLaunch Lc = L;
			Lc.range = ctl + kCtlRange + cls, Lc.batch_counter = ctl + kCtlCounter + cls;
			Lc.scratch = w->dp_scratch.p, Lc.ptr_lds_bytes = lds[cls];
			e = launch_dp_batched(Lc, kBandClass[cls], per_lane[cls], waves[cls], w->st);
		}
		if (e == hipSuccess) {
			DpLaunch Lp = L;
			Lp.range = ctl + kCtlRange + kNumBandClasses;
			e = launch_wave_stage(Lp, z, ctl + kCtlWave, w->wave_over.p, w->st);
		}
		hipGraph_t g = nullptr;
		const hipError_t e2 = hipStreamEndCapture(w->st, &g);
		if (e != hipSuccess || e2 != hipSuccess) {
			if (g) (void) hipGraphDestroy(g);
			return fail(c, HERO_ERR_HIP, "capturing the DP stage: %s", hipGetErrorString(e != hipSuccess ? e : e2));
		}
		w->graph = g;
		HIPCHK(c, hipGraphInstantiate(&w->graph_exec, w->graph, nullptr, nullptr, 0));
		w->graph_key = key;
		w->graph_builds++;
	}
	w->ctl_host[0] = ntasks;
	HIPCHK(c, hipMemcpyAsync(w->ctl.p, w->ctl_host, 4, hipMemcpyHostToDevice, w->st));
	HIPCHK(c, hipEventRecord(w->ev[7].a, w->st));
	HIPCHK(c, hipGraphLaunch(w->graph_exec, w->st));
	w->graph_launches++;
	HIPCHK(c, hipEventRecord(w->ev[3].a, w->st));
	return check_dp_skipped(c, w);
}

int run_chunk(hero_ctx *c, Worker *w, const Packed &S, int64_t s0, int32_t nsub, int rows_needed, const int64_t *s_uid) {
	const bool rec_mode = c->opt.records != 0;
	const Packed Q = view(c->Qd, c->Qd.good.p);
	const DevIndex ix = c->qix.view(c->o.k);
	ChunkDev cd;
	std::memset(&cd, 0, sizeof cd);
	cd.s_begin = s0, cd.nsub = nsub;
	HIPCHK(c, w->seg_n.ensure((size_t) 2 * nsub));
	HIPCHK(c, w->seg_off.ensure((size_t) 2 * nsub));
	cd.seg_n = w->seg_n.p, cd.seg_off = w->seg_off.p;
	// window buckets kept between the seed count and pass A1 when they fit kWinInfoBytes
	const int64_t win_stride = std::max(0, rows_needed - c->o.k);
	if (win_stride > 0 && (uint64_t) nsub * 2 * win_stride * 8 <= kWinInfoBytes) {
		HIPCHK(c, w->win_info.ensure((size_t) nsub * 2 * win_stride));
		HIPCHK(c, w->lane_n.ensure((size_t) nsub * 2 * 64));
		cd.win_stride = (int32_t) win_stride, cd.win_info = w->win_info.p, cd.lane_n = w->lane_n.p;
	}
	unsigned long long *stats = w->dstats.p;
	HIPCHK(c, hipEventRecord(w->ev[0].a, w->st));
	HIPCHK(c, launch_seed_count(S, Q, ix, c->o, cd, stats, w->st));
	HIPCHK(c, hipEventRecord(w->ev[8].a, w->st));
	uint64_t nseeds = 0;
	if (int r = wscan(c, w, cd.seg_n, cd.seg_off, 2 * (int64_t) nsub, nseeds)) return r;
	// a chunk whose seeds pass the seed stage's budget runs as two halves, in subject order
	// (repeat-heavy reads carry thousands of seeds each)
	const uint64_t seeds_max = c->opt.chunk_seeds_max > 0 ? std::min<uint64_t>((uint64_t) c->opt.chunk_seeds_max, kChunkSeedsMax)
	                                                      : kChunkSeedsMax;
	if (nseeds > seeds_max && nsub > 1) {
		const int32_t h = nsub / 2;
		if (int r = run_chunk(c, w, S, s0, h, rows_needed, s_uid)) return r;
		return run_chunk(c, w, S, s0 + h, nsub - h, rows_needed, s_uid);
	}
	return run_from_seeds(c, w, S, Q, ix, cd, nseeds, rows_needed, c->tally_off(), c->tally.p, c->tally.p + 4 * c->q_bases,
			s_uid, rec_mode);
}

// The rest of a chunk once its seeds are counted per (subject, strand) segment: seed emission,
// chaining, planning, DP, scoring, and votes or records.
int run_from_seeds(hero_ctx *c, Worker *w, const Packed &S, const Packed &Q, const DevIndex &ix, ChunkDev &cd,
		uint64_t nseeds, int rows_needed, const uint64_t *q_boff, uint32_t *tally, uint32_t *diff, const int64_t *s_uid,
		bool rec_mode) {
	const int32_t nsub = cd.nsub;
	const int64_t s0 = cd.s_begin;
	unsigned long long *stats = w->dstats.p;
	if (nseeds >= (1ull << 31)) return fail(c, HERO_ERR_LIMIT, "chunk of %d subjects yields %llu seeds; lower chunk_subjects", nsub, (unsigned long long) nseeds);
	const size_t ns = (size_t) nseeds;
	HIPCHK(c, w->seeds.ensure(ns));
	cd.seeds = w->seeds.p;
	HIPCHK(c, w->cnext.ensure(ns));
	HIPCHK(c, w->cfirst.ensure(ns));
	HIPCHK(c, w->clast.ensure(ns));
	HIPCHK(c, w->clen.ensure(ns));
	HIPCHK(c, w->gflag.ensure(ns));
	HIPCHK(c, w->g_nch.ensure(ns));
	HIPCHK(c, w->g_best.ensure(ns));
	HIPCHK(c, w->g_tasks.ensure(ns));
	HIPCHK(c, w->g_task_off.ensure(ns));
	HIPCHK(c, w->g_words.ensure(ns));
	HIPCHK(c, w->g_word_off.ensure(ns));
	HIPCHK(c, w->g_prog.ensure(ns));
	HIPCHK(c, w->g_prog_off.ensure(ns));
	cd.g_prog = w->g_prog.p, cd.g_prog_off = w->g_prog_off.p;
	cd.cnext = w->cnext.p, cd.cfirst = w->cfirst.p, cd.clast = w->clast.p, cd.clen = w->clen.p;
	cd.gflag = w->gflag.p, cd.g_tasks = w->g_tasks.p, cd.g_task_off = w->g_task_off.p;
	cd.g_nch = w->g_nch.p, cd.g_best = w->g_best.p;
	cd.g_words = w->g_words.p, cd.g_word_off = w->g_word_off.p;
	if (rec_mode) {
		HIPCHK(c, w->g_recs.ensure(ns));
		HIPCHK(c, w->g_rec_off.ensure(ns));
		cd.g_recs = w->g_recs.p;
	}
	const bool vote_mode = c->o.call_consensus && !rec_mode;
	if (vote_mode) {
		HIPCHK(c, w->sub_pairs.ensure((size_t) nsub + 1));
		HIPCHK(c, w->sub_pair_off.ensure((size_t) nsub + 1));
		cd.sub_pairs = w->sub_pairs.p;
	}
	HIPCHK(c, launch_pass_a1(S, Q, ix, c->o, cd, stats, w->st));
	HIPCHK(c, hipEventRecord(w->ev[1].a, w->st));
	uint64_t ntasks = 0, nwords = 0, nprog = 0;
	{
		const uint32_t *in[3] = {cd.g_tasks, cd.g_words, cd.g_prog};
		uint32_t *out[3] = {cd.g_task_off, cd.g_word_off, cd.g_prog_off};
		uint64_t tot[3] = {0, 0, 0};
		if (int r = wscan_multi(c, w, 3, in, out, (int64_t) ns, tot)) return r;
		ntasks = tot[0], nwords = tot[1], nprog = tot[2];
	}
	if (ntasks >= (1ull << 31) || nwords >= (1ull << 32) - 16)
		return fail(c, HERO_ERR_LIMIT, "chunk of %d subjects yields %llu DP tasks; lower chunk_subjects", nsub, (unsigned long long) ntasks);
	if (nprog >= (1ull << 32) - 16) return fail(c, HERO_ERR_LIMIT, "chunk of %d subjects yields %llu script words; lower chunk_subjects", nsub, (unsigned long long) nprog);
	HIPCHK(c, w->prog.ensure((size_t) nprog + 1));
	cd.prog = w->prog.p;
	uint64_t nvotes = 0;
	if (vote_mode) {
		if (int r = wscan(c, w, cd.sub_pairs, w->sub_pair_off.p, nsub, nvotes)) return r;
		if (nvotes >= (1ull << 31)) return fail(c, HERO_ERR_LIMIT, "chunk of %d subjects yields %llu pairs; lower chunk_subjects", nsub, (unsigned long long) nvotes);
		HIPCHK(c, w->votes.ensure((size_t) nvotes + 1));
		HIPCHK(c, hipMemsetAsync(w->votes.p, 0xFF, (size_t) nvotes * sizeof(VoteItem), w->st));
		cd.votes = w->votes.p, cd.sub_pair_off = w->sub_pair_off.p;
	}
	uint64_t nslots = 0;
	if (rec_mode) {
		if (int r = wscan(c, w, cd.g_recs, w->g_rec_off.p, (int64_t) ns, nslots)) return r;
		if (nslots >= (1ull << 31)) return fail(c, HERO_ERR_LIMIT, "chunk of %d subjects yields %llu record slots; lower chunk_subjects", nsub, (unsigned long long) nslots);
		HIPCHK(c, w->recs.ensure((size_t) nslots));
		HIPCHK(c, hipMemsetAsync(w->recs.p, 0xFF, (size_t) nslots * sizeof(RecHead), w->st));
		cd.recs = w->recs.p, cd.g_rec_off = w->g_rec_off.p;
	}
	HIPCHK(c, w->tasks.ensure((size_t) ntasks));
	HIPCHK(c, w->words.ensure((size_t) nwords + 1));
	HIPCHK(c, w->nsteps.ensure((size_t) ntasks));
	HIPCHK(c, w->sums.ensure((size_t) ntasks));
	cd.tasks = w->tasks.p, cd.words = w->words.p, cd.nsteps = w->nsteps.p, cd.sums = w->sums.p;
	HIPCHK(c, launch_pass_a2(S, Q, c->o, cd, w->st));
	HIPCHK(c, hipEventRecord(w->ev[2].a, w->st));
	if (int r = run_dp(c, w, S, Q, cd, (uint32_t) ntasks, rows_needed)) return r;
	if (c->tune) return HERO_OK; // autotuning times the DP only: no records, no votes
	HIPCHK(c, launch_pass_b(S, Q, c->o, cd, stats, w->st));
	if (vote_mode && c->o.mapping) HIPCHK(c, launch_map_select(S, cd, w->st)); // SubjectReadAligner.cpp:558-560
	if (vote_mode)
		HIPCHK(c, launch_votes(S, Q, c->o, cd, (uint32_t) nvotes, q_boff, tally, diff, c->o.merge ? c->ext.p : nullptr, stats, w->st));
	HIPCHK(c, hipEventRecord(w->ev[4].a, w->st));
	HIPCHK(c, hipStreamSynchronize(w->st));
	w->acc->ms_seed += elapsed(w, 0, 1);
	w->acc->ms_seed_probe += elapsed(w, 0, 8);
	w->acc->ms_plan += elapsed(w, 1, 7); // pass A2 + task bucketing
	w->acc->ms_dp += elapsed(w, 7, 3);   // DP kernels only
	w->acc->ms_score_vote += elapsed(w, 3, 4);
	w->acc->dp_tasks += ntasks;
	w->acc->subjects += (uint64_t) nsub;
	if (!rec_mode) return HERO_OK;
	// records: op counts, their scan, ops; then the host applies SubjectRead::addAlignment and
	// writes the .aln text
	const uint32_t nsl = (uint32_t) nslots;
	HIPCHK(c, w->rec_opn.ensure((size_t) nsl + 1));
	HIPCHK(c, w->rec_opoff.ensure((size_t) nsl + 1));
	HIPCHK(c, w->rec_pos.ensure((size_t) nsl + 1));
	HIPCHK(c, launch_record_ops(S, Q, c->o, cd, nsl, w->rec_opn.p, nullptr, nullptr, nullptr, w->st));
	uint64_t nops = 0;
	if (int r = wscan(c, w, w->rec_opn.p, w->rec_opoff.p, (int64_t) nsl, nops)) return r;
	if (nops >= (1ull << 32)) return fail(c, HERO_ERR_LIMIT, "chunk of %d subjects yields %llu CIGAR ops; lower chunk_subjects", nsub, (unsigned long long) nops);
	HIPCHK(c, w->rec_ops.ensure((size_t) nops + 1));
	HIPCHK(c, launch_record_ops(S, Q, c->o, cd, nsl, w->rec_opn.p, w->rec_opoff.p, w->rec_pos.p, w->rec_ops.p, w->st));
	HIPCHK(c, hipStreamSynchronize(w->st));
	return format_records(c, w, s0, nsub, s_uid, nsl);
}

// SubjectRead::addAlignment (src/SubjectRead.cpp:25-79) over a subject's accepted records in
// processing order, then the .aln text: ">sid\n" + "qid,strand,pos,CIGAR,interiorInsertedBases,"
// per kept record + "\n" (AlignmentRecord.cpp:102-149, SubjectRead.cpp:179-193,
// SubjectReadAligner.cpp:854-910). Record ids are query uniqueIds (without -c, :230-231).
int format_records(hero_ctx *c, Worker *w, int64_t s0, int32_t nsub, const int64_t *s_uid, uint32_t nslots) {
	std::vector<RecHead> h(nslots);
	std::vector<uint32_t> opn(nslots), opoff(nslots);
	std::vector<int32_t> pos(nslots);
	if (nslots) {
		HIPCHK(c, hipMemcpy(h.data(), w->recs.p, (size_t) nslots * sizeof(RecHead), hipMemcpyDeviceToHost));
		HIPCHK(c, hipMemcpy(opn.data(), w->rec_opn.p, (size_t) nslots * 4, hipMemcpyDeviceToHost));
		HIPCHK(c, hipMemcpy(opoff.data(), w->rec_opoff.p, (size_t) nslots * 4, hipMemcpyDeviceToHost));
		HIPCHK(c, hipMemcpy(pos.data(), w->rec_pos.p, (size_t) nslots * 4, hipMemcpyDeviceToHost));
	}
	const uint64_t nops = nslots ? (uint64_t) opoff[nslots - 1] + opn[nslots - 1] : 0;
	std::vector<uint32_t> ops((size_t) nops);
	if (nops) HIPCHK(c, hipMemcpy(ops.data(), w->rec_ops.p, (size_t) nops * 4, hipMemcpyDeviceToHost));
	// accepted records per subject, in slot (= processing) order
	std::vector<uint32_t> first((size_t) nsub + 1, 0), order;
	for (uint32_t i = 0; i < nslots; i++)
		if (h[i].qid != kNoRecord) first[(size_t) h[i].sub + 1]++;
	for (int32_t j = 0; j < nsub; j++) first[(size_t) j + 1] += first[(size_t) j];
	order.resize(first[(size_t) nsub]);
	{
		std::vector<uint32_t> cur(first.begin(), first.end() - 1);
		for (uint32_t i = 0; i < nslots; i++)
			if (h[i].qid != kNoRecord) order[cur[(size_t) h[i].sub]++] = i;
	}
	std::vector<std::string> text((size_t) nsub);
	std::vector<std::vector<uint32_t>> kept_of((size_t) nsub);
#pragma omp parallel for schedule(dynamic, 256)
	for (int32_t j = 0; j < nsub; j++) {
		std::string &t = text[(size_t) j];
		char buf[64];
		t += '>';
		t += std::to_string((long long) s_uid[s0 + j]);
		t += '\n';
		// addAlignment: keep one record per query id; lower rate sum wins, then the longer. Under -v
		// every accepted record is kept in processing order (src/SubjectRead.cpp:26-29).
		std::vector<uint32_t> kept;
		if (c->o.long_format) {
			for (uint32_t x = first[(size_t) j]; x < first[(size_t) j + 1]; x++) kept.push_back(order[x]);
			kept_of[(size_t) j] = kept;
			continue;
		}
		uint64_t largest = 0;
		auto id_of = [&](uint32_t i) { return (uint64_t) c->q_uid[h[i].qid]; };
		auto keep_better = [&](uint32_t &slot, uint32_t r) {
			const double a = h[slot].rate_sum, b = h[r].rate_sum;
			if (a < b) return;
			if (a == b && h[slot].len >= h[r].len) return;
			slot = r;
		};
		for (uint32_t x = first[(size_t) j]; x < first[(size_t) j + 1]; x++) {
			const uint32_t r = order[x];
			const uint64_t id = id_of(r);
			if (kept.empty()) {
				largest = id;
				kept.push_back(r);
			} else if (id_of(kept.back()) == id) {
				keep_better(kept.back(), r);
			} else if (largest >= id) {
				bool found = false;
				for (uint32_t &k : kept) {
					if (id_of(k) > id) break;
					if (id_of(k) == id) {
						keep_better(k, r);
						found = true;
						break;
					}
				}
				if (!found) kept.push_back(r);
			} else {
				largest = id;
				kept.push_back(r);
			}
		}
		kept_of[(size_t) j] = kept;
		for (uint32_t r : kept) {
			const RecHead &x = h[r];
			const int n = (int) opn[r];
			const uint32_t *op = ops.data() + opoff[r];
			snprintf(buf, sizeof buf, "%lld,%d,%d,", (long long) c->q_uid[x.qid], (int) x.strand, (int) pos[r]);
			t += buf;
			for (int i = 0; i < n; i++) {
				snprintf(buf, sizeof buf, "%u", op[i] >> 8);
				t += buf;
				t += (char) (op[i] & 0xFF);
			}
			t += ',';
			// interior inserted bases, from the query in the record's orientation
			const char *qs = c->q_seq.data() + c->q_off[x.qid];
			const int Lq = c->q_len[x.qid];
			const bool rc = x.strand == 1 && !c->o.call_consensus;
			int index = 0;
			for (int i = 0; i < n; i++) {
				const char ty = (char) (op[i] & 0xFF);
				const int cnt = (int) (op[i] >> 8);
				if (ty == 'D') continue;
				if (ty == 'I' && i != 0 && i != n - 1) {
					for (int m = 0; m < cnt; m++) {
						const int p = index + m;
						t += rc ? code_ascii(ascii_code(qs[Lq - 1 - p]) ^ 2) : qs[p];
					}
				}
				index += cnt;
			}
			t += ',';
		}
		t += '\n';
	}
	if (!c->o.long_format)
		for (const std::string &t : text) c->aln += t;
	// the same records as data (hero_take_records)
	for (int32_t j = 0; j < nsub; j++) {
		for (uint32_t r : kept_of[(size_t) j]) {
			hero_record rec;
			std::memset(&rec, 0, sizeof rec);
			rec.subject = c->rec_subjects + j, rec.query = h[r].qid, rec.strand = h[r].strand, rec.position = pos[r];
			rec.op_off = (int64_t) c->rec_out_ops.size(), rec.op_n = (int32_t) opn[r];
			c->rec_out_ops.insert(c->rec_out_ops.end(), ops.begin() + opoff[r], ops.begin() + opoff[r] + opn[r]);
			c->rec_out.push_back(rec);
		}
	}
	c->rec_subjects += nsub;
	return HERO_OK;
}

// Serving: one batch of query reads corrected against the resident subjects (-c). Seeds come
// from the subject index (every k-mer position) probed with the batch's query keys, sorted by
// (subject, strand) and handed to the offline pipeline's per-subject passes; per-query results
// do not depend on which other queries share the batch.
int correct_batch_on(hero_ctx *c, Worker *w, int64_t n, const char *seq, const int64_t *off, const int32_t *len,
		const int64_t *uid, char *corrected) {
	if (!c->six.built) return fail(c, HERO_ERR_STATE, "hero_build_subject_index must run before correcting batches");
	if (!c->o.call_consensus) return fail(c, HERO_ERR_ARG, "batch correction needs call_consensus = 1");
	if (c->o.max_candidates > 0) return fail(c, HERO_ERR_ARG, "max_candidates depends on the batch mix; not allowed in batch correction");
	if (int r = check_reads(c, n, seq, off, len, uid)) return r;
	if (n >= (1 << 22)) return fail(c, HERO_ERR_LIMIT, "at most 2^22 reads per batch");
	const int k = c->o.k;
	const uint32_t min_ov = (uint16_t) c->o.min_overlap;
	// per-read offsets of the batch: bases (boff), packed words (woff) and key references (koff);
	// reads not longer than the minimum overlap can never be good and get no keys
	HIPCHK(c, w->h_meta.ensure((size_t) 3 * (n + 1)));
	uint64_t *hb = w->h_meta.p, *hw = hb + (n + 1), *hk = hw + (n + 1);
	uint64_t bases = 0, rwords = 0, nkeys64 = 0;
	bool contiguous = true;
	int max_len = c->r_max_len;
	for (int64_t i = 0; i < n; i++) {
		hb[i] = bases, hw[i] = rwords, hk[i] = nkeys64;
		contiguous = contiguous && off[i] == off[0] + (int64_t) bases;
		bases += (uint64_t) len[i], rwords += (uint64_t) (len[i] + 31) / 32;
		if ((uint64_t) len[i] > min_ov) nkeys64 += (uint64_t) num_keys(len[i], k);
		max_len = std::max(max_len, len[i]);
	}
	hb[n] = bases, hw[n] = rwords, hk[n] = nkeys64;
	if (nkeys64 >= (1ull << 31)) return fail(c, HERO_ERR_LIMIT, "batch has too many keys");
	const uint32_t nkeys = (uint32_t) nkeys64;
	// the batch's bytes go to the device as they are (one copy when its reads are back to back);
	// packing, the quality filter and the key references are computed there
	const char *src = n ? seq + off[0] : nullptr;
	if (!contiguous) {
		HIPCHK(c, w->h_bytes.ensure((size_t) bases + 1));
		for (int64_t i = 0; i < n; i++) std::memcpy(w->h_bytes.p + hb[i], seq + off[i], (size_t) len[i]);
		src = w->h_bytes.p;
	}
	ReadsDev &d = w->bq;
	HIPCHK(c, d.words.ensure((size_t) rwords + 2));
	HIPCHK(c, d.woff.ensure((size_t) n + 1));
	HIPCHK(c, d.boff.ensure((size_t) n + 1));
	HIPCHK(c, d.len.ensure((size_t) n + 1));
	HIPCHK(c, d.uid.ensure((size_t) n + 1));
	HIPCHK(c, d.good.ensure((size_t) n + 1));
	HIPCHK(c, w->b_koff.ensure((size_t) n + 1));
	HIPCHK(c, w->b_raw.ensure((size_t) bases + 1));
	HIPCHK(c, w->b_tally.ensure((size_t) 5 * bases + 5)); // 4 mismatch counters + 1 run difference per base
	HIPCHK(c, w->b_cons.ensure((size_t) bases + 1));
	HIPCHK(c, w->b_kref.ensure((size_t) nkeys + 1));
	HIPCHK(c, w->b_cnt.ensure((size_t) nkeys + 1));
	HIPCHK(c, w->b_off.ensure((size_t) nkeys + 1));
	if (bases && contiguous) {
		if (int r = staged_h2d(c, w->st, w->h_bytes, (char *) w->b_raw.p, src, (size_t) bases)) return r;
	} else if (bases) {
		HIPCHK(c, hipMemcpyAsync(w->b_raw.p, src, (size_t) bases, hipMemcpyHostToDevice, w->st));
	}
	if (n) {
		HIPCHK(c, hipMemcpyAsync(d.boff.p, hb, (size_t) n * 8, hipMemcpyHostToDevice, w->st));
		HIPCHK(c, hipMemcpyAsync(d.woff.p, hw, (size_t) n * 8, hipMemcpyHostToDevice, w->st));
		HIPCHK(c, hipMemcpyAsync(w->b_koff.p, hk, (size_t) n * 8, hipMemcpyHostToDevice, w->st));
		HIPCHK(c, hipMemcpyAsync(d.len.p, len, (size_t) n * 4, hipMemcpyHostToDevice, w->st));
		HIPCHK(c, hipMemcpyAsync(d.uid.p, uid, (size_t) n * 8, hipMemcpyHostToDevice, w->st));
	}
	d.n = n;
	HIPCHK(c, launch_read_intake(w->b_raw.p, d.boff.p, d.len.p, d.woff.p, w->b_koff.p, n, k, min_ov, d.words.p, rwords,
			d.good.p, w->b_kref.p, w->st));
	HIPCHK(c, hipMemsetAsync(w->b_tally.p, 0, (size_t) 20 * bases, w->st));
	const Packed Q = view(w->bq, w->bq.good.p);
	const Packed S = view(c->Rd, c->Rd.good.p);
	const DevIndex six = c->six.view(k);
	HIPCHK(c, hipEventRecord(w->ev[0].a, w->st));
	HIPCHK(c, launch_query_probe(Q, S, six, c->o, w->b_kref.p, nkeys, w->b_cnt.p, nullptr, nullptr, nullptr, w->st));
	uint64_t nseeds = 0;
	if (int r = wscan(c, w, w->b_cnt.p, w->b_off.p, nkeys, nseeds)) return r;
	if (nseeds >= (1ull << 31)) return fail(c, HERO_ERR_LIMIT, "batch yields %llu seeds", (unsigned long long) nseeds);
	const uint32_t ns = (uint32_t) nseeds;
	HIPCHK(c, w->b_keys.ensure((size_t) ns + 1));
	HIPCHK(c, w->b_vals.ensure((size_t) ns + 1));
	HIPCHK(c, w->b_keys2.ensure((size_t) ns + 1));
	HIPCHK(c, w->b_vals2.ensure((size_t) ns + 1));
	HIPCHK(c, launch_query_probe(Q, S, six, c->o, w->b_kref.p, nkeys, w->b_cnt.p, w->b_off.p, w->b_keys.p, w->b_vals.p, w->st));
	// two stable radix passes: by lo = (strand, posQ, posS), then by hi = (subject, query)
	int hi_bits = 23;
	while (((uint64_t) 1 << (hi_bits - 22)) < (uint64_t) c->Rd.n + 1) hi_bits++;
	size_t tmp_bytes = 0;
	HIPCHK(c, sort_pairs_u64(w->b_keys.p, w->b_keys2.p, w->b_vals.p, w->b_vals2.p, ns, 29, nullptr, tmp_bytes, w->st));
	size_t tmp2 = 0;
	HIPCHK(c, sort_pairs_u64(w->b_vals2.p, w->b_vals.p, w->b_keys2.p, w->b_keys.p, ns, hi_bits, nullptr, tmp2, w->st));
	HIPCHK(c, w->b_sort_tmp.ensure(std::max(tmp_bytes, tmp2) + 1));
	tmp_bytes = w->b_sort_tmp.n;
	HIPCHK(c, sort_pairs_u64(w->b_keys.p, w->b_keys2.p, w->b_vals.p, w->b_vals2.p, ns, 29, w->b_sort_tmp.p, tmp_bytes, w->st));
	tmp_bytes = w->b_sort_tmp.n;
	HIPCHK(c, sort_pairs_u64(w->b_vals2.p, w->b_vals.p, w->b_keys2.p, w->b_keys.p, ns, hi_bits, w->b_sort_tmp.p, tmp_bytes, w->st));
	// now b_vals = hi, b_keys = lo, sorted by (hi, lo): groups and pairs are contiguous runs
	HIPCHK(c, w->b_gflag.ensure((size_t) ns + 1));
	HIPCHK(c, w->b_pflag.ensure((size_t) ns + 1));
	HIPCHK(c, w->b_gx.ensure((size_t) ns + 1));
	HIPCHK(c, w->b_px.ensure((size_t) ns + 1));
	HIPCHK(c, launch_serve_mark(w->b_vals.p, w->b_keys.p, ns, w->b_gflag.p, w->b_pflag.p, w->st));
	uint64_t ngroups = 0, npairs = 0;
	if (int r = wscan(c, w, w->b_gflag.p, w->b_gx.p, ns, ngroups)) return r;
	if (int r = wscan(c, w, w->b_pflag.p, w->b_px.p, ns, npairs)) return r;
	HIPCHK(c, w->b_glist.ensure((size_t) ngroups + 1));
	HIPCHK(c, w->b_plist.ensure((size_t) npairs + 1));
	HIPCHK(c, w->seeds.ensure((size_t) ns + 1));
	HIPCHK(c, launch_serve_lists(w->b_vals.p, w->b_keys.p, ns, w->b_gflag.p, w->b_gx.p, w->b_pflag.p, w->b_px.p,
			w->b_glist.p, w->b_plist.p, w->seeds.p, w->st));
	HIPCHK(c, hipEventRecord(w->ev[8].a, w->st));
	ServeDev v;
	v.hi = w->b_vals.p, v.lo = w->b_keys.p, v.glist = w->b_glist.p, v.gx = w->b_gx.p, v.plist = w->b_plist.p;
	v.nseeds = ns, v.ngroups = (uint32_t) ngroups, v.npairs = (uint32_t) npairs;
	ChunkDev cd;
	std::memset(&cd, 0, sizeof cd);
	HIPCHK(c, w->cnext.ensure((size_t) ns + 1));
	HIPCHK(c, w->cfirst.ensure((size_t) ns + 1));
	HIPCHK(c, w->clast.ensure((size_t) ns + 1));
	HIPCHK(c, w->clen.ensure((size_t) ns + 1));
	HIPCHK(c, w->g_nch.ensure((size_t) ns + 1));
	HIPCHK(c, w->g_best.ensure((size_t) ns + 1));
	HIPCHK(c, w->g_tasks.ensure((size_t) ngroups + 1));
	HIPCHK(c, w->g_task_off.ensure((size_t) ngroups + 1));
	HIPCHK(c, w->g_words.ensure((size_t) ngroups + 1));
	HIPCHK(c, w->g_word_off.ensure((size_t) ngroups + 1));
	HIPCHK(c, w->g_prog.ensure((size_t) ngroups + 1));
	HIPCHK(c, w->g_prog_off.ensure((size_t) ngroups + 1));
	cd.g_prog = w->g_prog.p, cd.g_prog_off = w->g_prog_off.p;
	cd.seeds = w->seeds.p, cd.cnext = w->cnext.p, cd.cfirst = w->cfirst.p, cd.clast = w->clast.p, cd.clen = w->clen.p;
	cd.g_nch = w->g_nch.p, cd.g_best = w->g_best.p, cd.g_tasks = w->g_tasks.p, cd.g_task_off = w->g_task_off.p;
	cd.g_words = w->g_words.p, cd.g_word_off = w->g_word_off.p;
	unsigned long long *stats = w->dstats.p;
	HIPCHK(c, launch_serve_a1(S, Q, c->o, cd, v, stats, w->st));
	HIPCHK(c, hipEventRecord(w->ev[1].a, w->st));
	uint64_t ntasks = 0, nwords = 0, nprog = 0;
	{
		const uint32_t *in[3] = {cd.g_tasks, cd.g_words, cd.g_prog};
		uint32_t *out[3] = {cd.g_task_off, cd.g_word_off, cd.g_prog_off};
		uint64_t tot[3] = {0, 0, 0};
		if (int r = wscan_multi(c, w, 3, in, out, (int64_t) ngroups, tot)) return r;
		ntasks = tot[0], nwords = tot[1], nprog = tot[2];
	}
	if (ntasks >= (1ull << 31) || nwords >= (1ull << 32) - 16)
		return fail(c, HERO_ERR_LIMIT, "batch yields %llu DP tasks", (unsigned long long) ntasks);
	if (nprog >= (1ull << 32) - 16) return fail(c, HERO_ERR_LIMIT, "batch yields %llu script words", (unsigned long long) nprog);
	HIPCHK(c, w->prog.ensure((size_t) nprog + 1));
	cd.prog = w->prog.p;
	HIPCHK(c, w->tasks.ensure((size_t) ntasks + 1));
	HIPCHK(c, w->words.ensure((size_t) nwords + 1));
	HIPCHK(c, w->nsteps.ensure((size_t) ntasks + 1));
	HIPCHK(c, w->sums.ensure((size_t) ntasks + 1));
	cd.tasks = w->tasks.p, cd.words = w->words.p, cd.nsteps = w->nsteps.p, cd.sums = w->sums.p;
	HIPCHK(c, launch_serve_a2(S, Q, c->o, cd, v, w->st));
	HIPCHK(c, hipEventRecord(w->ev[2].a, w->st));
	if (int r = run_dp(c, w, S, Q, cd, (uint32_t) ntasks, max_len + 1)) return r;
	HIPCHK(c, launch_serve_b(S, Q, c->o, cd, v, w->bq.boff.p, w->b_tally.p, w->b_tally.p + 4 * bases, stats, w->st));
	HIPCHK(c, hipEventRecord(w->ev[4].a, w->st));
	HIPCHK(c, hipStreamSynchronize(w->st));
	w->acc->ms_seed += elapsed(w, 0, 1);
	w->acc->ms_seed_probe += elapsed(w, 0, 8);
	w->acc->ms_plan += elapsed(w, 1, 7);
	w->acc->ms_dp += elapsed(w, 7, 3);
	w->acc->ms_score_vote += elapsed(w, 3, 4);
	w->acc->dp_tasks += ntasks;
	// consensus of the batch, corrected bases back to the caller's layout
	// corrected bytes: the batch's own bytes, overwritten by the consensus of its good reads
	HIPCHK(c, hipEventRecord(w->ev[5].a, w->st));
	if (bases) HIPCHK(c, hipMemcpyAsync(w->b_cons.p, w->b_raw.p, (size_t) bases, hipMemcpyDeviceToDevice, w->st));
	HIPCHK(c, launch_consensus(Q, c->o, w->bq.boff.p, w->bq.boff.p, w->b_tally.p, w->b_tally.p + 4 * bases, 0, n, w->b_cons.p, nullptr,
			w->dstats.p, w->st));
	HIPCHK(c, hipEventRecord(w->ev[6].a, w->st));
	// back in the caller's layout: one staged copy when its reads are back to back
	if (bases && contiguous) {
		if (int r = staged_d2h(c, w->st, w->h_bytes, corrected + off[0], (const char *) w->b_cons.p, (size_t) bases)) return r;
	} else if (bases) {
		HIPCHK(c, hipMemcpyAsync(w->h_bytes.p, w->b_cons.p, (size_t) bases, hipMemcpyDeviceToHost, w->st));
	}
	HIPCHK(c, hipStreamSynchronize(w->st));
	if (!contiguous)
		for (int64_t i = 0; i < n; i++) std::memcpy(corrected + off[i], w->h_bytes.p + hb[i], (size_t) len[i]);
	w->acc->ms_consensus += elapsed(w, 5, 6);
	w->acc->queries += (uint64_t) n;
	return HERO_OK;
}

int align_packed(hero_ctx *c, const Packed &S, int64_t n, int max_len, const int64_t *s_uid) {
	if (!c->qix.built) return fail(c, HERO_ERR_STATE, "hero_build_index must run before aligning subjects");
	c->cons_valid = false;
	const int64_t chunk = c->opt.chunk_subjects > 0 ? c->opt.chunk_subjects : auto_chunk_subjects(n);
	const int64_t nchunks = (n + chunk - 1) / chunk;
	// Chunks in flight at once: votes are integer atomic sums into the shared tallies, so their
	// order does not matter; records are written in subject order and autotuning times one chunk
	// at a time, so those run one chunk after the other.
	const bool concurrent = (c->opt.chunk_workers == 0 || c->opt.chunk_workers >= 2) && !c->opt.records && !c->tune &&
			nchunks > 1;
	if (concurrent && !c->w2_ready) {
		c->w2.acc = &c->st_w2;
		if (hipError_t e = c->w2.init(); e != hipSuccess) {
			c->w2.release();
			return fail(c, HERO_ERR_HIP, "second chunk worker: %s", hipGetErrorString(e));
		}
		c->w2_ready = true;
	}
	if (!concurrent) {
		for (int64_t s0 = 0; s0 < n; s0 += chunk) {
			const int32_t m = (int32_t) std::min<int64_t>(chunk, n - s0);
			if (int r = run_chunk(c, &c->w, S, s0, m, max_len + 1, s_uid)) return r;
		}
		return HERO_OK;
	}
	// two workers, each on its own stream and host thread, take the next chunk as they finish
	std::atomic<int64_t> next{0};
	std::atomic<bool> stop{false};
	auto work = [&](Worker *w) {
		(void) hipSetDevice(c->opt.device);
		for (;;) {
			const int64_t k = next.fetch_add(1);
			if (k >= nchunks || stop.load()) return HERO_OK;
			const int64_t s0 = k * chunk;
			const int32_t m = (int32_t) std::min<int64_t>(chunk, n - s0);
			if (int r = run_chunk(c, w, S, s0, m, max_len + 1, s_uid)) {
				stop.store(true);
				return r;
			}
		}
	};
	int r2 = HERO_OK;
	std::thread t([&] { r2 = work(&c->w2); });
	const int r1 = work(&c->w);
	t.join();
	return r1 != HERO_OK ? r1 : r2;
}

// Seed index over the good reads of R: keys at 0, k, 2k, ... and L - k (the reference's query
// index), or at every position (all_positions: the serving index over resident subjects). The
// device table is sized to the keys, not to the reference's size (slot placement never reaches an
// output); ref_size, when not 0, is the reference's table size, whose distinct keys must fit
// (HashTableLongKey.cpp:320-326 exits otherwise).
int build_index(hero_ctx *c, const ReadsDev &R, const std::vector<uint8_t> &good, const std::vector<int32_t> &len,
		bool all_positions, uint64_t ref_size, IndexBufs &ix) {
	const int k = c->o.k;
	const int64_t n = R.n;
	// The query index (not all_positions) keeps its table and build scratch between builds, and its
	// key offsets while the read set is the same upload: a step that rebuilds it allocates nothing.
	// The serving index (every position of the resident set) is built once and frees its scratch.
	const bool keep = !all_positions;
	const bool same_set = keep && R.gen != 0 && ix.src_gen == R.gen && ix.src_n == n;
	(void) good, (void) len; // the device copies of the flags and lengths (R.good, R.len) are used
	if (!keep) ix.release();
	uint64_t nocc = 0;
	if (same_set) {
		nocc = ix.src_nocc;
	} else {
		// each read's key offset, computed on the device (the good flags and lengths are there)
		ix.src_gen = 0; // set once the build has succeeded
		HIPCHK(c, ix.dkoff.ensure((size_t) n + 1));
		HIPCHK(c, ix.ktmp.ensure(key_offsets_tmp_bytes(n)));
		HIPCHK(c, launch_key_offsets(view(R, R.good.p), k, all_positions, ix.dkoff.p, ix.ktmp.p, ix.ktmp.n, c->w.st));
		HIPCHK(c, hipMemcpyAsync(&nocc, ix.dkoff.p + n, 8, hipMemcpyDeviceToHost, c->w.st));
		HIPCHK(c, hipStreamSynchronize(c->w.st));
	}
	if (nocc >= 0xFFFFFFFFull) return fail(c, HERO_ERR_LIMIT, "%llu index keys exceed 2^32-1", (unsigned long long) nocc);
	ix.built = false;
	// about 4 slots per key: half the windows probed miss, and a miss walks to the next empty slot.
	// With subjects in locality order, neighbouring probes share cache lines, and 4 and 8 slots per
	// key time the same on the 1.7 M and 33 M read sets; 4 halves the table and its clearing.
	ix.P = next_prime_above(std::max<uint64_t>((all_positions ? 2 : 4) * nocc + 1, 1021));
	if (ix.P >= 0xFFFFFFFFull) return fail(c, HERO_ERR_LIMIT, "table size %llu exceeds 2^32-1 slots", (unsigned long long) ix.P);
	hipEvent_t e0 = c->w.ev[5].a, e1 = c->w.ev[6].a;
	HIPCHK(c, hipEventRecord(e0, c->w.st));
	DevBuf<uint64_t> &dkoff = ix.dkoff, &occ_hi = ix.occ_hi, &occ_lo = ix.occ_lo, &occ_entry = ix.occ_entry;
	DevBuf<uint32_t> &occ_slot = ix.occ_slot, &occ_rank = ix.occ_rank, &owner = ix.owner, &maxp = ix.maxp;
	DevBuf<int> &errf = ix.errf;
	DevBuf<unsigned long long> &cursor = ix.cursor;
	auto cleanup = [&]() {
		if (!keep) ix.release_scratch();
	};
	int rc = HERO_OK;
	do {
#define IXCHK(expr)                                                                                                 \
	{                                                                                                               \
		hipError_t e_ = (expr);                                                                                     \
		if (e_ != hipSuccess) {                                                                                     \
			rc = fail(c, e_ == hipErrorOutOfMemory ? HERO_ERR_NOMEM : HERO_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
			break;                                                                                                  \
		}                                                                                                           \
	}
		IXCHK(dkoff.ensure((size_t) n + 1));
		IXCHK(occ_hi.ensure(nocc));
		IXCHK(occ_lo.ensure(nocc));
		IXCHK(occ_entry.ensure(nocc));
		IXCHK(occ_slot.ensure(nocc));
		IXCHK(occ_rank.ensure(nocc));
		IXCHK(owner.ensure(ix.P));
		IXCHK(maxp.ensure(1));
		IXCHK(errf.ensure(1));
		IXCHK(cursor.ensure(2));
		IXCHK(ix.slots.ensure(ix.P));
		IXCHK(ix.entries.ensure(nocc));
		IXCHK(hipMemsetAsync(owner.p, 0, ix.P * 4, c->w.st));
		IXCHK(hipMemsetAsync(ix.slots.p, 0, ix.P * sizeof(Slot), c->w.st));
		IXCHK(hipMemsetAsync(maxp.p, 0, 4, c->w.st));
		IXCHK(hipMemsetAsync(errf.p, 0, 4, c->w.st));
		IXCHK(hipMemsetAsync(cursor.p, 0, 16, c->w.st));
		const Packed Rp = view(R, R.good.p);
		IXCHK(launch_index_gen(Rp, dkoff.p, k, all_positions, occ_hi.p, occ_lo.p, occ_entry.p, occ_rank.p, c->w.st));
		IXCHK(launch_index_claim(occ_hi.p, occ_lo.p, nocc, ix.P, k, owner.p, occ_slot.p, maxp.p, errf.p, c->w.st));
		IXCHK(launch_index_count(occ_slot.p, nocc, ix.slots.p, occ_rank.p, c->w.st));
		IXCHK(launch_index_alloc(owner.p, ix.P, occ_hi.p, occ_lo.p, k, ix.slots.p, cursor.p, c->w.st));
		IXCHK(launch_index_fill(occ_slot.p, occ_rank.p, occ_entry.p, nocc, ix.slots.p, ix.entries.p, c->w.st));
		IXCHK(launch_index_sort_buckets(ix.slots.p, ix.P, ix.entries.p, c->w.st));
		IXCHK(hipEventRecord(e1, c->w.st));
		int herr = 0;
		unsigned long long used = 0, distinct = 0;
		IXCHK(hipMemcpyAsync(&ix.max_probe, maxp.p, 4, hipMemcpyDeviceToHost, c->w.st));
		IXCHK(hipMemcpyAsync(&herr, errf.p, 4, hipMemcpyDeviceToHost, c->w.st));
		IXCHK(hipMemcpyAsync(&used, cursor.p, 8, hipMemcpyDeviceToHost, c->w.st));
		IXCHK(hipMemcpyAsync(&distinct, cursor.p + 1, 8, hipMemcpyDeviceToHost, c->w.st));
		IXCHK(hipStreamSynchronize(c->w.st));
		if (herr) {
			rc = fail(c, HERO_ERR_REF_ABORT, "hash table full");
			break;
		}
		if (ref_size && distinct > ref_size) {
			rc = fail(c, HERO_ERR_REF_ABORT, "hash table too small (%llu distinct keys, %llu slots); the reference exits here",
					distinct, (unsigned long long) ref_size);
			break;
		}
		if (used != nocc) {
			rc = fail(c, HERO_ERR_HIP, "index build placed %llu of %llu keys", used, (unsigned long long) nocc);
			break;
		}
		// distinct keys = claimed slots
		ix.total_keys = nocc;
		ix.distinct_keys = distinct;
		float ms = 0;
		(void) hipEventElapsedTime(&ms, e0, e1);
		c->st_host.ms_index += ms;
#undef IXCHK
	} while (0);
	cleanup();
	if (rc != HERO_OK) return rc;
	if (keep) ix.src_gen = R.gen, ix.src_n = n, ix.src_nocc = nocc;
	ix.built = true;
	return HERO_OK;
}

} // namespace

namespace {

// Test route of hero_band_align_batch (hero_options.band_align_form 1-4): the calls become DP
// tasks over a packed read set (each call's two sides as two reads) and run through one batched
// kernel of the correction pipeline: 1 one task per lane in the general rows, 2 the tagged rows,
// 3 two tasks per lane in packed 16-bit, 4 two tasks per lane in the tagged packed 16-bit rows.
// Returns per call its edit words (at woff[i]) and step count.
int band_align_via_batched(hero_ctx *c, int32_t n, const hero_band_pair *pairs, const char *seq,
		std::vector<int64_t> &woff, std::vector<uint32_t> &words, std::vector<uint32_t> &nsteps) {
	const int form = c->opt.band_align_form;
	if (form < 1 || form > 4)
		return fail(c, HERO_ERR_ARG, "band_align_form %d (0 wave, 1 general, 2 tagged, 3 packed, 4 tagged packed)", form);
	std::vector<int64_t> roff((size_t) 2 * n);
	std::vector<int32_t> rlen((size_t) 2 * n);
	std::vector<DpTask> tasks((size_t) n);
	woff.assign((size_t) n, 0);
	int64_t nw = 0;
	int bmax = 0, rows = 1;
	for (int i = 0; i < n; i++) {
		const hero_band_pair &p = pairs[i];
		for (int64_t k = 0; k < p.s_len; k++)
			if (seq[p.s_off + k] == 'N') return fail(c, HERO_ERR_ARG, "pair %d: the batched forms take A/C/G/T only", i);
		for (int64_t k = 0; k < p.q_len; k++)
			if (seq[p.q_off + k] == 'N') return fail(c, HERO_ERR_ARG, "pair %d: the batched forms take A/C/G/T only", i);
		roff[(size_t) 2 * i] = p.s_off, rlen[(size_t) 2 * i] = p.s_len;
		roff[(size_t) 2 * i + 1] = p.q_off, rlen[(size_t) 2 * i + 1] = p.q_len;
		const bool sw = !(p.q_len <= p.s_len);
		const int lenS = sw ? p.q_len : p.s_len, lenQ = sw ? p.s_len : p.q_len;
		DpTask &t = tasks[(size_t) i];
		std::memset(&t, 0, sizeof t);
		t.s_len = (uint16_t) p.s_len, t.q_len = (uint16_t) p.q_len; // words and bits: after packing
		t.band = (uint16_t) std::max(p.band, lenS - lenQ);
		t.flags = (uint8_t) ((p.flag == 1 ? kTaskAlignLeft : p.flag == 2 ? kTaskAlignRight : 0) | (sw ? kTaskSwapped : 0));
		t.out_off = (uint32_t) nw;
		woff[(size_t) i] = nw;
		nw += edit_words(p.s_len, p.q_len);
		bmax = std::max(bmax, (int) t.band), rows = std::max(rows, lenS + 1);
	}
	int cls = -1;
	for (int k = kNumBandClasses - 1; k >= 0; k--)
		if (bmax <= kBandClass[k]) cls = k;
	const int limit = form == 1 ? kBandClass[kNumBandClasses - 1] : (form == 2 ? kTaggedMaxBand : kPk16MaxBand);
	if (cls < 0 || kBandClass[cls] > limit) return fail(c, HERO_ERR_ARG, "band %d beyond the batched form's classes", bmax);
	if (rows > 512) return fail(c, HERO_ERR_ARG, "the batched forms take at most 511 rows");
	if (form == 2 && !dp_tagged_form(c->o, rows)) return fail(c, HERO_ERR_ARG, "the tagged affine rows need dx == dy");
	if (form == 3 && !pk16_class_fits(c->o, rows)) return fail(c, HERO_ERR_ARG, "the calls exceed the 16-bit range");
	if (form == 4 && (c->o.aligner != 0 || !pk_form_fits(c->o, rows, kBandClass[cls], true)))
		return fail(c, HERO_ERR_ARG, "the tagged packed rows take affine calls within their 16-bit range");
	std::vector<uint64_t> pw, pwoff;
	pack_reads(seq, roff.data(), rlen.data(), 2 * (int64_t) n, pw, pwoff);
	if (pw.size() >= ((uint64_t) 1 << 32)) return fail(c, HERO_ERR_LIMIT, "the pairs' packed codes exceed 2^32 words");
	for (int i = 0; i < n; i++) { // both sides in their forward frames from their first base
		DpTask &t = tasks[(size_t) i];
		segment_start(pwoff[(size_t) 2 * i], t.s_len, 0, t.s_len, false, t.s_word, t.s_bit);
		segment_start(pwoff[(size_t) 2 * i + 1], t.q_len, 0, t.q_len, false, t.q_word, t.q_bit);
	}
	DevBuf<uint64_t> dw, dwoff;
	DevBuf<int32_t> dlen;
	DevBuf<DpTask> dt;
	DevBuf<uint32_t> dwords, dns, dscr;
	DevBuf<StepSummary> dsums;
	const int bm = kBandClass[cls], per_lane = form >= 3 ? 2 : 1, nwd = (2 * bm + 1 + 15) / 16;
	const int waves = (int) std::max<int64_t>(1, std::min<int64_t>((int64_t) c->cu_count * 2, (n + 64 * per_lane - 1) / (64 * per_lane)));
	hipStream_t st = c->w.st;
	HIPCHK(c, dw.exact(pw.size()));
	HIPCHK(c, dwoff.exact((size_t) 2 * n));
	HIPCHK(c, dlen.exact((size_t) 2 * n));
	HIPCHK(c, dt.exact((size_t) n));
	HIPCHK(c, dwords.exact((size_t) nw + 1));
	HIPCHK(c, dns.exact((size_t) n));
	HIPCHK(c, dsums.exact((size_t) n));
	HIPCHK(c, dscr.exact((size_t) waves * rows * nwd * 64 * per_lane));
	HIPCHK(c, hipMemcpyAsync(dw.p, pw.data(), pw.size() * 8, hipMemcpyHostToDevice, st));
	HIPCHK(c, hipMemcpyAsync(dwoff.p, pwoff.data(), (size_t) 2 * n * 8, hipMemcpyHostToDevice, st));
	HIPCHK(c, hipMemcpyAsync(dlen.p, rlen.data(), (size_t) 2 * n * 4, hipMemcpyHostToDevice, st));
	HIPCHK(c, hipMemcpyAsync(dt.p, tasks.data(), (size_t) n * sizeof(DpTask), hipMemcpyHostToDevice, st));
	HIPCHK(c, hipMemsetAsync(c->w.counter.p, 0, 4, st));
	Packed R;
	std::memset(&R, 0, sizeof R);
	R.words = dw.p, R.woff = dwoff.p, R.len = dlen.p, R.n = 2 * (int64_t) n;
	DpLaunch L;
	std::memset(&L, 0, sizeof L);
	L.tasks = dt.p, L.ntasks = (uint32_t) n, L.S = R, L.Q = R, L.o = c->o;
	L.words = dwords.p, L.nsteps = dns.p, L.sums = dsums.p, L.scratch = dscr.p, L.rows_max = rows;
	L.tagged = form == 2 || form == 4 ? 1 : 0, L.batch_counter = c->w.counter.p, L.stats = c->w.dstats.p;
	// the near pointer rows of the pipeline's default occupancy, so longer calls cross into scratch
	if (per_lane == 1)
		L.ptr_lds_bytes = dp_batched_lds_bytes(bm, c->o.aligner == 0, L.tagged != 0,
				dp_batched_waves_per_cu(bm, c->o.aligner == 0, 1, rows, L.tagged != 0));
	HIPCHK(c, launch_dp_batched(L, bm, per_lane, waves, st));
	words.resize((size_t) nw + 1);
	nsteps.resize((size_t) n);
	HIPCHK(c, hipMemcpyAsync(words.data(), dwords.p, words.size() * 4, hipMemcpyDeviceToHost, st));
	HIPCHK(c, hipMemcpyAsync(nsteps.data(), dns.p, (size_t) n * 4, hipMemcpyDeviceToHost, st));
	HIPCHK(c, hipStreamSynchronize(st));
	dw.release(), dwoff.release(), dlen.release(), dt.release(), dwords.release(), dns.release(), dsums.release(), dscr.release();
	return check_dp_skipped(c, &c->w);
}

// BandAligner::calculateCigarAndPosition[Reversed] (src/BandAligner.cpp:258-419) from edit words:
// M runs merge, every mismatch is an op of its own typed by the original query's base, I / D
// swap when the aligner's sides were swapped.
int steps_to_cigars(hero_ctx *c, int32_t n, const hero_band_pair *pairs, const char *seq, const std::vector<int64_t> &woff,
		const std::vector<uint32_t> &hw, const std::vector<uint32_t> &hn, int32_t *band_out, int64_t *cigar_off,
		int32_t *cigar_n, int32_t *cigar_counts, char *cigar_types, int64_t cigar_cap) {
	int64_t pos = 0;
	for (int i = 0; i < n; i++) {
		const hero_band_pair &p = pairs[i];
		const bool sw = !(p.q_len <= p.s_len);
		const char *A = seq + (sw ? p.q_off : p.s_off), *B = seq + (sw ? p.s_off : p.q_off);
		band_out[i] = std::max(p.band, std::abs(p.s_len - p.q_len));
		const int cap = p.s_len + p.q_len, ns = (int) hn[(size_t) i];
		cigar_off[i] = pos;
		int nops = 0, cnt = 0, ii = 0, jj = 0;
		char cur = 0;
		bool cur_mis = false;
		auto put = [&]() -> bool {
			if (pos + nops >= cigar_cap) return false;
			if (cigar_counts) cigar_counts[pos + nops] = cnt;
			if (cigar_types) cigar_types[pos + nops] = cur;
			nops++;
			return true;
		};
		bool ok = true;
		for (int e = cap - ns; e < cap && ok; e++) {
			const uint32_t stp = (hw[(size_t) woff[(size_t) i] + (e >> 4)] >> (2 * (e & 15))) & 3u;
			char t;
			bool mis = false;
			if (stp == kStepLeft) t = sw ? 'D' : 'I', jj++;
			else if (stp == kStepUp) t = sw ? 'I' : 'D', ii++;
			else {
				const char a = A[ii], q = B[jj];
				mis = a != q;
				t = mis ? (sw ? a : q) : 'M';
				ii++, jj++;
			}
			if (cnt > 0 && t == cur && !mis && !cur_mis) {
				cnt++;
				continue;
			}
			if (cnt > 0) ok = put();
			cur = t, cur_mis = mis, cnt = 1;
		}
		if (ok && cnt > 0) ok = put();
		if (!ok) return fail(c, HERO_ERR_LIMIT, "cigar_cap %lld too small", (long long) cigar_cap);
		cigar_n[i] = nops;
		pos += nops;
	}
	return HERO_OK;
}

} // namespace

extern "C" {

uint64_t hero_prime_larger_than(uint64_t n) { return ref_prime_larger_than(n); }

void hero_default_options(hero_options *p) {
	std::memset(p, 0, sizeof *p);
	p->k = 39, p->aligner = 1, p->call_consensus = 0, p->hash_factor = 32, p->min_overlap = 40;
	p->mismatch_rate = 1.0, p->indel_rate = 1.0, p->consensus_freq = 0.6;
	p->dx = -4, p->dy = -4, p->ex = -1, p->sx = 0, p->miss = -4, p->hit = 1, p->unobserved = -1;
	p->device = 0, p->dp_form = 0, p->chunk_subjects = 0, p->dp_waves = 0;
}

int hero_create(const hero_options *opt, hero_ctx **out) {
	if (!opt || !out) return HERO_ERR_ARG;
	*out = nullptr;
	if (opt->k < 1 || opt->k > 64) return HERO_ERR_ARG;
	if (opt->aligner != 0 && opt->aligner != 1) return HERO_ERR_ARG;
	if (opt->records && opt->call_consensus) return HERO_ERR_ARG; // the reference writes no .aln under -c
	if (opt->min_overlap < 0 || opt->min_overlap > 65535) return HERO_ERR_ARG;
	if (opt->chunk_seeds_max < 0) return HERO_ERR_ARG;
	if (opt->hash_factor < 1) return HERO_ERR_ARG;
	if (opt->long_format && opt->call_consensus) return HERO_ERR_ARG; // -v is built for record mode (.aln)
	if (opt->merge && (!opt->call_consensus || opt->mapping)) return HERO_ERR_ARG; // --merge acts on -c votes
	int ndev = 0;
	if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= opt->device || opt->device < 0) return HERO_ERR_HIP;
	if (hipSetDevice(opt->device) != hipSuccess) return HERO_ERR_HIP;
	hero_ctx *c = new hero_ctx();
	c->opt = *opt;
	c->o = make_opts(*opt);
	std::memset(&c->st_host, 0, sizeof c->st_host);
	std::memset(&c->st_w2, 0, sizeof c->st_w2);
	hipDeviceProp_t prop;
	if (hipGetDeviceProperties(&prop, opt->device) != hipSuccess) {
		delete c;
		return HERO_ERR_HIP;
	}
	c->cu_count = prop.multiProcessorCount;
	if (hipError_t e = noalloc_setup(opt->device); e != hipSuccess) {
		delete c;
		return HERO_ERR_HIP;
	}
	c->w.acc = &c->st_host;
	if (hipError_t e = c->w.init(); e != hipSuccess) {
		c->w.release();
		delete c;
		return e == hipErrorOutOfMemory ? HERO_ERR_NOMEM : HERO_ERR_HIP;
	}
	*out = c;
	return HERO_OK;
}

void hero_destroy(hero_ctx *c) {
	if (!c) return;
	(void) hipSetDevice(c->opt.device);
	(void) hipStreamSynchronize(c->w.st);
	c->Qd.release(), c->Sd.release(), c->q_subject_good.release();
	c->qix.release(), c->tally.release(), c->ext.release(), c->cons_out.release(), c->cons_packed.release();
	c->Rd.release(), c->six.release();
	c->ba_pairs.release(), c->ba_seq.release(), c->ba_types.release(), c->ba_off.release();
	c->ba_band.release(), c->ba_n.release(), c->ba_counts.release(), c->ba_scratch.release();
	c->stage.release();
	c->w.release();
	if (c->w2_ready) c->w2.release();
	delete c;
}

const char *hero_last_error(const hero_ctx *c) { return c ? c->err.c_str() : "null context"; }

int hero_set_queries(hero_ctx *c, int64_t n, const char *seq, const int64_t *off, const int32_t *len, const int64_t *uid) {
	if (!c) return HERO_ERR_ARG;
	if (int r = check_reads(c, n, seq, off, len, uid)) return r;
	HIPCHK(c, hipSetDevice(c->opt.device));
	c->q_off.resize((size_t) n), c->q_len.resize((size_t) n), c->q_uid.resize((size_t) n);
	par_copy(c->q_off.data(), off, n);
	par_copy(c->q_len.data(), len, n);
	par_copy(c->q_uid.data(), uid, n);
	c->q_good.resize((size_t) n);
	c->q_active_as_subject.resize((size_t) n);
	c->q_not_good.clear();
	// contiguous: each read starts where the one before it ends, the first at 0
	int64_t qb = 0;
	int contiguous = 1;
#pragma omp parallel for reduction(+ : qb) reduction(&& : contiguous) schedule(static, 65536)
	for (int64_t i = 0; i < n; i++) {
		qb += len[i];
		contiguous = contiguous && off[i] == (i ? off[i - 1] + len[i - 1] : 0);
	}
	c->q_contiguous = contiguous != 0;
	// the reads' flags (read_flags) are computed while they are packed for the upload
	const ReadFlagsOut flags{c->o.min_overlap, c->o.k, c->q_good.data(), c->q_active_as_subject.data()};
	if (int r = upload_reads(c, c->Qd, n, seq, off, len, uid, c->q_good, &flags)) return r;
	// reads that are not corrected go out as they came in: keep their bytes (all bytes when the
	// layout has gaps, which the output keeps too)
	c->q_seq.clear();
	c->q_keep_off.clear();
	// record mode writes the queries' inserted bases: it keeps every byte, like a gapped layout
	c->q_full_copy = !c->q_contiguous || c->opt.records;
	if (c->q_full_copy) {
		int64_t bases = 0;
		for (int64_t i = 0; i < n; i++) bases = std::max<int64_t>(bases, off[i] + len[i]);
		c->q_seq.assign(seq, seq + bases);
	}
	for (int64_t i = 0; i < n; i++)
		if (!c->q_good[(size_t) i]) {
			c->q_not_good.push_back(i);
			if (!c->q_full_copy) {
				c->q_keep_off.push_back((int64_t) c->q_seq.size());
				c->q_seq.insert(c->q_seq.end(), seq + off[i], seq + off[i] + len[i]);
			}
		}
	HIPCHK(c, c->q_subject_good.ensure((size_t) n + 1));
	if (n) HIPCHK(c, hipMemcpyAsync(c->q_subject_good.p, c->q_active_as_subject.data(), (size_t) n, hipMemcpyHostToDevice, c->w.st));
	c->q_ordered = false;
	if (c->locality_order() && n > 0) {
		HIPCHK(c, c->t_off.ensure((size_t) n + 1));
		if (int r = order_reads(c, c->w.st, c->Qd, c->q_subject_good.p, c->q_subject_good.p, c->q_len, c->Qo, c->t_off.p))
			return r;
		c->q_ordered = true;
	}
	c->q_bases = qb;
	// 4 mismatch counters per base, then the matched-run difference array (1 per base)
	if (c->tally.n < (size_t) 5 * qb + 1 || !c->tally.p) HIPCHK(c, c->tally.exact((size_t) 5 * qb + 1)); // no slack: 20 B per base
	if (c->opt.merge) HIPCHK(c, c->ext.ensure((size_t) kExtPositions * 4 * n + 1));
	c->votes_clear = false; // cleared by hero_build_index (every run builds one), or by the first reader
	HIPCHK(c, hipStreamSynchronize(c->w.st));
	c->qix.built = false;
	c->cons_valid = false;
	c->st_host.queries = (uint64_t) n;
	return HERO_OK;
}

// Zeroes the vote tallies and the --merge extension tally of the current query set.
static int clear_votes(hero_ctx *c) {
	if (c->q_bases) HIPCHK(c, hipMemsetAsync(c->tally.p, 0, (size_t) 20 * c->q_bases, c->w.st));
	if (c->opt.merge && c->Qd.n) HIPCHK(c, hipMemsetAsync(c->ext.p, 0, (size_t) kExtPositions * 16 * c->Qd.n, c->w.st));
	c->votes_clear = true;
	return HERO_OK;
}
// Readers of the tallies after hero_set_queries without a build since: they read zeros.
static int clear_votes_once(hero_ctx *c) { return c->votes_clear ? HERO_OK : clear_votes(c); }

int hero_build_index(hero_ctx *c) {
	if (!c) return HERO_ERR_ARG;
	HIPCHK(c, hipSetDevice(c->opt.device));
	// a new index starts a new correction run: clear the vote tallies
	if (int r = clear_votes(c)) return r;
	// --mapping sizes the table by the bases' k-mers (src/HashTableLongKey.cpp:52-60)
	const uint64_t units = c->opt.mapping ? (uint64_t) c->q_bases / (uint64_t) c->opt.k : (uint64_t) c->Qd.n;
	const uint64_t P = ref_prime_larger_than(units * (uint64_t) c->opt.hash_factor); // :61
	c->cons_valid = false;
	if (int r = build_index(c, c->Qd, c->q_good, c->q_len, false, P, c->qix)) return r;
	c->st_host.table_size = P; // the reference's size; the device table is sized to the keys
	c->st_host.distinct_keys = c->qix.distinct_keys;
	c->st_host.total_keys = c->qix.total_keys;
	c->st_host.max_probe = c->qix.max_probe;
	return HERO_OK;
}

int hero_align_subjects(hero_ctx *c, int64_t n, const char *seq, const int64_t *off, const int32_t *len, const int64_t *uid) {
	if (!c) return HERO_ERR_ARG;
	if (int r = check_reads(c, n, seq, off, len, uid)) return r;
	HIPCHK(c, hipSetDevice(c->opt.device));
	std::vector<uint8_t> act((size_t) n);
	for (int64_t i = 0; i < n; i++) act[(size_t) i] = subject_active(seq + off[i], len[i], c->o.min_overlap, c->o.k);
	if (int r = upload_reads(c, c->Sd, n, seq, off, len, uid, act)) return r;
	int ml = 0;
	for (int64_t i = 0; i < n; i++) ml = std::max(ml, len[i]);
	for (int32_t l : c->q_len) ml = std::max(ml, l);
	if (c->locality_order() && n > 0) {
		if (int r = order_reads(c, c->w.st, c->Sd, c->Sd.good.p, c->Sd.good.p, std::vector<int32_t>(len, len + n), c->So,
				nullptr))
			return r;
		return align_packed(c, view(c->So, c->So.good.p), n, ml, nullptr);
	}
	return align_packed(c, view(c->Sd, c->Sd.good.p), n, ml, uid);
}

int hero_align_self(hero_ctx *c) {
	if (!c) return HERO_ERR_ARG;
	HIPCHK(c, hipSetDevice(c->opt.device));
	int ml = 0;
	for (int32_t l : c->q_len) ml = std::max(ml, l);
	if (c->q_ordered) return align_packed(c, view(c->Qo, c->Qo.good.p), c->Qo.n, ml, nullptr);
	return align_packed(c, view(c->Qd, c->q_subject_good.p), c->Qd.n, ml, c->q_uid.data());
}

int hero_take_alignments(hero_ctx *c, char *out, int64_t cap, int64_t *len) {
	if (!c || !len) return HERO_ERR_ARG;
	*len = (int64_t) c->aln.size();
	if (!out) return HERO_OK;
	if (cap < *len) return fail(c, HERO_ERR_LIMIT, "alignment text is %lld bytes, buffer %lld", (long long) *len, (long long) cap);
	std::memcpy(out, c->aln.data(), c->aln.size());
	std::string().swap(c->aln);
	return HERO_OK;
}

int hero_take_records(hero_ctx *c, hero_record *recs, int64_t rec_cap, int64_t *n_recs, uint32_t *ops, int64_t op_cap,
		int64_t *n_ops) {
	if (!c || !n_recs || !n_ops) return HERO_ERR_ARG;
	*n_recs = (int64_t) c->rec_out.size(), *n_ops = (int64_t) c->rec_out_ops.size();
	if (!recs || !ops) return HERO_OK;
	if (rec_cap < *n_recs || op_cap < *n_ops) return fail(c, HERO_ERR_LIMIT, "record buffers too small");
	if (*n_recs) std::memcpy(recs, c->rec_out.data(), (size_t) *n_recs * sizeof(hero_record));
	if (*n_ops) std::memcpy(ops, c->rec_out_ops.data(), (size_t) *n_ops * 4);
	c->rec_out.clear(), c->rec_out_ops.clear();
	c->rec_subjects = 0;
	return HERO_OK;
}

int hero_get_extension_tally(hero_ctx *c, uint32_t *out) {
	if (!c || !out) return HERO_ERR_ARG;
	if (!c->opt.merge) return fail(c, HERO_ERR_STATE, "the extension tally exists with merge = 1");
	HIPCHK(c, hipSetDevice(c->opt.device));
	if (int r = clear_votes_once(c)) return r;
	if (c->Qd.n)
		HIPCHK(c, hipMemcpy(out, c->ext.p, (size_t) kExtPositions * 16 * c->Qd.n, hipMemcpyDeviceToHost));
	return HERO_OK;
}

int hero_call_consensus(hero_ctx *c, char *corrected) {
	if (!c) return HERO_ERR_ARG;
	HIPCHK(c, hipSetDevice(c->opt.device));
	if (int r = clear_votes_once(c)) return r;
	const int64_t n = c->Qd.n;
	HIPCHK(c, c->cons_out.ensure((size_t) c->q_bases + 1));
	HIPCHK(c, hipEventRecord(c->w.ev[5].a, c->w.st));
	HIPCHK(c, launch_consensus(view(c->Qd, c->Qd.good.p), c->o, c->Qd.boff.p, c->tally_off(), c->tally.p,
			c->tally.p + 4 * c->q_bases, 0, n, c->cons_out.p, nullptr,
			c->w.dstats.p, c->w.st));
	HIPCHK(c, hipEventRecord(c->w.ev[6].a, c->w.st));
	HIPCHK(c, hipStreamSynchronize(c->w.st));
	float ms = 0;
	(void) hipEventElapsedTime(&ms, c->w.ev[5].a, c->w.ev[6].a);
	c->st_host.ms_consensus += ms;
	c->cons_valid = true;
	return corrected ? hero_copy_corrected(c, corrected) : HERO_OK;
}

int hero_copy_corrected(hero_ctx *c, char *corrected) {
	if (!c || !corrected) return HERO_ERR_ARG;
	if (!c->cons_valid) return fail(c, HERO_ERR_STATE, "hero_call_consensus has not run on the current tallies");
	HIPCHK(c, hipSetDevice(c->opt.device));
	const int64_t n = c->Qd.n;
	// The device holds corrected bases at packed base offsets; the caller's layout usually is that
	// same back-to-back layout, which takes one copy. Reads that are not good keep their bytes.
	// Corrected bases of good reads are A/C/G/T: they cross the link as 2-bit codes.
	HIPCHK(c, c->cons_packed.ensure((size_t) (c->q_bases + 15) / 16 + 1));
	HIPCHK(c, launch_pack_bases(c->cons_out.p, (uint64_t) c->q_bases, c->cons_packed.p, c->w.st));
	if (c->q_contiguous) {
		if (int r = staged_d2h_bases(c, c->w.st, c->stage, corrected, c->cons_packed.p, (size_t) c->q_bases)) return r;
		for (size_t t = 0; t < c->q_not_good.size(); t++) {
			const int64_t i = c->q_not_good[t];
			const char *src = c->q_seq.data() + (c->q_full_copy ? c->q_off[(size_t) i] : c->q_keep_off[t]);
			std::memcpy(corrected + c->q_off[(size_t) i], src, (size_t) c->q_len[(size_t) i]);
		}
	} else {
		std::vector<char> h((size_t) c->q_bases);
		if (int r = staged_d2h_bases(c, c->w.st, c->stage, h.data(), c->cons_packed.p, h.size())) return r;
		std::memcpy(corrected, c->q_seq.data(), c->q_seq.size());
		uint64_t b = 0;
		for (int64_t i = 0; i < n; i++) {
			const int L = c->q_len[(size_t) i];
			if (c->q_good[(size_t) i]) std::memcpy(corrected + c->q_off[(size_t) i], h.data() + b, (size_t) L);
			b += (uint64_t) L;
		}
	}
	return HERO_OK;
}

int hero_get_tally(hero_ctx *c, uint32_t *tally) {
	if (!c || !tally) return HERO_ERR_ARG;
	HIPCHK(c, hipSetDevice(c->opt.device));
	if (int r = clear_votes_once(c)) return r;
	DevBuf<uint8_t> codes;
	DevBuf<uint32_t> dec;
	HIPCHK(c, codes.exact((size_t) c->q_bases));
	HIPCHK(c, dec.exact((size_t) 4 * c->q_bases));
	HIPCHK(c, launch_consensus(view(c->Qd, c->Qd.good.p), c->o, c->Qd.boff.p, c->tally_off(), c->tally.p,
			c->tally.p + 4 * c->q_bases, 0, c->Qd.n, codes.p, dec.p,
			nullptr, c->w.st));
	std::vector<uint32_t> h((size_t) 4 * c->q_bases);
	HIPCHK(c, hipMemcpyAsync(h.data(), dec.p, h.size() * 4, hipMemcpyDeviceToHost, c->w.st));
	HIPCHK(c, hipStreamSynchronize(c->w.st));
	// device layout is by base offset (reads packed back to back); map to the caller's offsets
	uint64_t b = 0;
	for (int64_t i = 0; i < c->Qd.n; i++) {
		const int L = c->q_len[(size_t) i];
		std::memcpy(tally + 4 * c->q_off[(size_t) i], h.data() + 4 * b, (size_t) L * 16);
		b += (uint64_t) L;
	}
	return HERO_OK;
}

int hero_set_tally(hero_ctx *c, const uint32_t *tally) {
	if (!c || !tally) return HERO_ERR_ARG;
	HIPCHK(c, hipSetDevice(c->opt.device));
	if (int r = clear_votes_once(c)) return r; // the extension tally; the tallies are written below
	if (!c->q_bases) return HERO_OK;
	// The device form: the four counts as given, and a difference array whose running sum is each
	// base's total count, so that consensus decodes the own base's count to exactly its given value.
	std::vector<uint32_t> h((size_t) 5 * c->q_bases);
	uint32_t *h4 = h.data(), *hd = h.data() + 4 * c->q_bases;
	// each read's counters sit at its tally offset (its place in the locality order, or its base offset)
	std::vector<uint64_t> toff((size_t) c->Qd.n);
	if (c->Qd.n) HIPCHK(c, hipMemcpy(toff.data(), c->tally_off(), toff.size() * 8, hipMemcpyDeviceToHost));
	for (int64_t i = 0; i < c->Qd.n; i++) {
		const int L = c->q_len[(size_t) i];
		const uint32_t *src = tally + 4 * c->q_off[(size_t) i];
		uint32_t prev = 0; // each read's running sum starts at its first base
		uint64_t b = toff[(size_t) i];
		for (int p = 0; p < L; p++, b++) {
			uint32_t tot = 0;
			for (int j = 0; j < 4; j++) h4[4 * b + j] = src[4 * p + j], tot += src[4 * p + j];
			hd[b] = tot - prev; // mod 2^32, as the votes themselves
			prev = tot;
		}
	}
	HIPCHK(c, hipMemcpyAsync(c->tally.p, h.data(), h.size() * 4, hipMemcpyHostToDevice, c->w.st));
	HIPCHK(c, hipStreamSynchronize(c->w.st));
	c->cons_valid = false;
	return HERO_OK;
}

extern "C" int hero_coverage_histogram(hero_ctx *c, uint64_t *counts, int32_t cap, int32_t *max_coverage) {
	if (!c || !counts || cap < 1 || !max_coverage) return HERO_ERR_ARG;
	HIPCHK(c, hipSetDevice(c->opt.device));
	if (int r = clear_votes_once(c)) return r;
	DevBuf<unsigned long long> hist;
	DevBuf<uint32_t> mx;
	HIPCHK(c, hist.exact((size_t) cap));
	HIPCHK(c, mx.exact(1));
	HIPCHK(c, hipMemsetAsync(hist.p, 0, (size_t) cap * 8, c->w.st));
	HIPCHK(c, hipMemsetAsync(mx.p, 0, 4, c->w.st));
	HIPCHK(c, launch_coverage_hist(c->Qd.len.p, c->tally_off(), c->tally.p + 4 * c->q_bases, c->Qd.n, (uint32_t) cap,
			hist.p, mx.p, c->w.st));
	uint32_t m = 0;
	HIPCHK(c, hipMemcpyAsync(counts, hist.p, (size_t) cap * 8, hipMemcpyDeviceToHost, c->w.st));
	HIPCHK(c, hipMemcpyAsync(&m, mx.p, 4, hipMemcpyDeviceToHost, c->w.st));
	HIPCHK(c, hipStreamSynchronize(c->w.st));
	*max_coverage = (int32_t) (m < 0x7FFFFFFFu ? m : 0x7FFFFFFFu);
	if ((int64_t) m >= (int64_t) cap)
		return fail(c, HERO_ERR_LIMIT, "coverage %u needs a histogram of %u bins (cap %d)", m, m + 1, cap);
	return HERO_OK;
}

int hero_get_stats(hero_ctx *c, hero_stats *out) {
	if (!c || !out) return HERO_ERR_ARG;
	HIPCHK(c, hipSetDevice(c->opt.device));
	std::vector<unsigned long long> stripes((size_t) kStatStripes * kStatStride);
	unsigned long long s[kStCount] = {};
	for (Worker *w : {&c->w, &c->w2}) {
		if (w == &c->w2 && !c->w2_ready) continue;
		HIPCHK(c, hipMemcpy(stripes.data(), w->dstats.p, stripes.size() * 8, hipMemcpyDeviceToHost));
		for (int i = 0; i < kStatStripes; i++)
			for (int j = 0; j < kStCount; j++) s[j] += stripes[(size_t) i * kStatStride + j];
	}
	*out = c->st_host;
	// the second chunk worker's device time adds to the first's (the workers overlap, so stage
	// times are GPU time summed over both streams)
	const hero_stats &b = c->st_w2;
	out->ms_seed += b.ms_seed, out->ms_seed_probe += b.ms_seed_probe, out->ms_plan += b.ms_plan;
	out->ms_dp += b.ms_dp, out->ms_score_vote += b.ms_score_vote;
	out->dp_tasks += b.dp_tasks, out->subjects += b.subjects;
	out->lookups = s[kStLookups];
	out->seeds = s[kStSeeds];
	out->dp_calls_exec = s[kStCallsExec];
	out->dp_cells_exec = s[kStCellsExec];
	out->records_built = s[kStBuilt];
	out->records_kept = s[kStKept];
	out->vote_oob = s[kStOob];
	out->dp_skipped = s[kStDpSkipped];
	out->reads_changed = s[kStReadsChanged];
	out->dp_cells_all = s[kStCellsAll];
	out->dp_graph_builds = c->w.graph_builds + c->w2.graph_builds;
	out->dp_graph_launches = c->w.graph_launches + c->w2.graph_launches;
	out->ms_total = out->ms_index + out->ms_seed + out->ms_plan + out->ms_dp + out->ms_score_vote + out->ms_consensus;
	return HERO_OK;
}

int hero_reset_stats(hero_ctx *c) {
	if (!c) return HERO_ERR_ARG;
	HIPCHK(c, hipSetDevice(c->opt.device));
	HIPCHK(c, hipMemset(c->w.dstats.p, 0, (size_t) kStatStripes * kStatStride * 8));
	if (c->w2_ready) HIPCHK(c, hipMemset(c->w2.dstats.p, 0, (size_t) kStatStripes * kStatStride * 8));
	c->w.graph_builds = 0, c->w.graph_launches = 0, c->w2.graph_builds = 0, c->w2.graph_launches = 0;
	std::memset(&c->st_w2, 0, sizeof c->st_w2);
	const uint64_t P = c->st_host.table_size, tk = c->st_host.total_keys, mp = c->st_host.max_probe, q = c->st_host.queries;
	const uint64_t dk = c->st_host.distinct_keys;
	std::memset(&c->st_host, 0, sizeof c->st_host);
	c->st_host.table_size = P, c->st_host.total_keys = tk, c->st_host.max_probe = mp, c->st_host.queries = q;
	c->st_host.distinct_keys = dk;
	return HERO_OK;
}

int hero_band_align_batch(hero_ctx *c, int32_t n, const hero_band_pair *pairs, const char *seq, int32_t *band_out,
		int64_t *cigar_off, int32_t *cigar_n, int32_t *cigar_counts, char *cigar_types, int64_t cigar_cap) {
	if (!c || n < 0 || (n > 0 && (!pairs || !seq || !band_out || !cigar_off || !cigar_n))) return HERO_ERR_ARG;
	HIPCHK(c, hipSetDevice(c->opt.device));
	if (n == 0) return HERO_OK;
	// setSequences: the longer side is the aligner's subject (the query side keeps it on ties);
	// m_band = max(band, lenS - lenQ)
	int64_t seq_end = 0, total = 0;
	int max_len = 1, max_band = 0;
	std::vector<int64_t> off((size_t) n);
	for (int i = 0; i < n; i++) {
		const hero_band_pair &p = pairs[i];
		if (p.s_len < 0 || p.q_len < 0 || p.s_off < 0 || p.q_off < 0) return fail(c, HERO_ERR_REF_ABORT, "pair %d: negative start (the reference exits)", i);
		if (p.s_len == 0 && p.q_len == 0) return fail(c, HERO_ERR_ARG, "pair %d: both sequences empty", i);
		seq_end = std::max<int64_t>(seq_end, std::max(p.s_off + p.s_len, p.q_off + p.q_len));
		const int lenS = std::max(p.s_len, p.q_len), lenQ = std::min(p.s_len, p.q_len);
		const int b = std::max(p.band, lenS - lenQ);
		if (b > kWaveMaxBand) return fail(c, HERO_ERR_LIMIT, "pair %d: band %d above %d", i, b, kWaveMaxBand);
		if (lenS > kMaxReadLen) return fail(c, HERO_ERR_LIMIT, "pair %d: %d bases (limit %d)", i, lenS, kMaxReadLen);
		max_len = std::max(max_len, lenS), max_band = std::max(max_band, b);
		off[(size_t) i] = total;
		total += (int64_t) p.s_len + p.q_len;
	}
	for (int64_t i = 0; i < seq_end; i++) {
		const char ch = seq[i];
		if (ch != 'A' && ch != 'C' && ch != 'G' && ch != 'T' && ch != 'N')
			return fail(c, HERO_ERR_ARG, "sequence byte %lld is '%c'; only A/C/G/T/N are accepted", (long long) i, ch);
	}
	if (c->opt.band_align_form != 0) {
		// test route through the pipeline's batched DP kernels
		std::vector<uint32_t> words, nsteps;
		std::vector<int64_t> woff;
		if (int r = band_align_via_batched(c, n, pairs, seq, woff, words, nsteps)) return r;
		return steps_to_cigars(c, n, pairs, seq, woff, words, nsteps, band_out, cigar_off, cigar_n, cigar_counts, cigar_types,
				cigar_cap);
	}
	hipStream_t st = c->w.st;
	HIPCHK(c, c->ba_pairs.ensure((size_t) n));
	HIPCHK(c, c->ba_seq.ensure((size_t) seq_end + 1));
	HIPCHK(c, c->ba_off.ensure((size_t) n));
	HIPCHK(c, c->ba_band.ensure((size_t) n));
	HIPCHK(c, c->ba_n.ensure((size_t) n));
	HIPCHK(c, c->ba_counts.ensure((size_t) total));
	HIPCHK(c, c->ba_types.ensure((size_t) total));
	WaveSizing z = wave_sizing(max_len, max_band, (int) std::min<int64_t>(n, (int64_t) c->cu_count * 16));
	HIPCHK(c, c->ba_scratch.ensure(z.scratch_words * (size_t) z.nwaves + 1));
	z.scratch = z.scratch_words ? c->ba_scratch.p : nullptr;
	HIPCHK(c, hipMemcpyAsync(c->ba_pairs.p, pairs, (size_t) n * sizeof(hero_band_pair), hipMemcpyHostToDevice, st));
	HIPCHK(c, hipMemcpyAsync(c->ba_seq.p, seq, (size_t) seq_end, hipMemcpyHostToDevice, st));
	HIPCHK(c, hipMemcpyAsync(c->ba_off.p, off.data(), (size_t) n * 8, hipMemcpyHostToDevice, st));
	uint32_t *ctl_wave = c->w.ctl.p + kCtlWave;
	HIPCHK(c, hipMemsetAsync(ctl_wave, 0, kWaveCtlWords * 4, st));
	WavePairs P;
	std::memset(&P, 0, sizeof P);
	P.pairs = c->ba_pairs.p, P.n = n, P.max_len = max_len, P.max_band = max_band, P.seq = c->ba_seq.p;
	P.band_out = c->ba_band.p, P.cigar_off = c->ba_off.p, P.cigar_n = c->ba_n.p, P.cigar_counts = c->ba_counts.p;
	P.cigar_types = c->ba_types.p;
	for (int l = 0; l < kWaveLaunches; l++) {
		if (l > 0 && max_band <= wave_band_hi(l - 1)) break;
		P.counter = ctl_wave + kWaveCtlCounter + l;
		HIPCHK(c, launch_dp_wave_pairs(P, c->o, z, l, st));
	}
	std::vector<int32_t> hn((size_t) n), hcnt((size_t) total);
	std::vector<char> htyp((size_t) total);
	HIPCHK(c, hipMemcpyAsync(band_out, c->ba_band.p, (size_t) n * 4, hipMemcpyDeviceToHost, st));
	HIPCHK(c, hipMemcpyAsync(hn.data(), c->ba_n.p, (size_t) n * 4, hipMemcpyDeviceToHost, st));
	HIPCHK(c, hipMemcpyAsync(hcnt.data(), c->ba_counts.p, (size_t) total * 4, hipMemcpyDeviceToHost, st));
	HIPCHK(c, hipMemcpyAsync(htyp.data(), c->ba_types.p, (size_t) total, hipMemcpyDeviceToHost, st));
	HIPCHK(c, hipStreamSynchronize(st));
	// the device wrote each call's ops at the front of its region; pack them back to back
	int64_t pos = 0;
	for (int i = 0; i < n; i++) {
		const int k = hn[(size_t) i];
		if (k < 0) return fail(c, k, "pair %d: the device aligner returned %d", i, k);
		if (pos + k > cigar_cap) return fail(c, HERO_ERR_LIMIT, "cigar_cap %lld too small", (long long) cigar_cap);
		cigar_off[i] = pos, cigar_n[i] = k;
		if (cigar_counts) std::memcpy(cigar_counts + pos, hcnt.data() + off[(size_t) i], (size_t) k * 4);
		if (cigar_types) std::memcpy(cigar_types + pos, htyp.data() + off[(size_t) i], (size_t) k);
		pos += k;
	}
	return HERO_OK;
}

int hero_band_align_batch_noalloc(hero_ctx *c, int32_t n, const hero_band_pair *pairs, const char *seq, int32_t max_len,
		int32_t max_band, int32_t *band_out, const int64_t *cigar_off, int32_t *cigar_n, int32_t *cigar_counts,
		char *cigar_types, void *stream) {
	if (!c || n < 0) return HERO_ERR_ARG;
	if (n > 0 && (!pairs || !seq || !band_out || !cigar_off || !cigar_n || !cigar_counts || !cigar_types))
		return fail(c, HERO_ERR_ARG, "null device array");
	if (max_len < 1 || max_len > kMaxReadLen) return fail(c, HERO_ERR_LIMIT, "max_len %d outside 1..%d", max_len, kMaxReadLen);
	if (max_band < 0 || max_band > kWaveMaxBand)
		return fail(c, HERO_ERR_LIMIT, "max_band %d outside 0..%d", max_band, kWaveMaxBand);
	HIPCHK(c, hipSetDevice(c->opt.device));
	uint8_t *base = noalloc_scratch_of(c->opt.device);
	if (!base) return fail(c, HERO_ERR_STATE, "no scratch address for device %d", c->opt.device);
	if (n == 0) return HERO_OK;
	WaveSizing z = wave_sizing(max_len, max_band, (int) std::min<int64_t>(n, (int64_t) c->cu_count * 16));
	const size_t avail_words = (kNoAllocScratchBytes - kNoAllocCtlBytes) / 4;
	if (z.scratch_words) {
		if (z.scratch_words > avail_words)
			return fail(c, HERO_ERR_LIMIT,
					"calls of up to %d bases with bands up to %d need %zu bytes of scratch per wavefront; the fixed scratch holds %zu",
					max_len, max_band, z.scratch_words * 4, avail_words * 4);
		z.nwaves = (int) std::min<size_t>((size_t) z.nwaves, avail_words / z.scratch_words);
		z.scratch = (uint32_t *) (base + kNoAllocCtlBytes);
	}
	hipStream_t st = stream ? (hipStream_t) stream : c->w.st;
	uint32_t *ctl = (uint32_t *) base;
	HIPCHK(c, launch_fill_u32(ctl, kWaveCtlWords, 0u, st)); // a kernel, not a memset: capturable
	WavePairs P;
	std::memset(&P, 0, sizeof P);
	P.pairs = pairs, P.n = n, P.max_len = max_len, P.max_band = max_band, P.seq = seq;
	P.band_out = band_out, P.cigar_off = cigar_off, P.cigar_n = cigar_n, P.cigar_counts = cigar_counts;
	P.cigar_types = cigar_types;
	for (int l = 0; l < kWaveLaunches; l++) {
		if (l > 0 && max_band <= wave_band_hi(l - 1)) break;
		P.counter = ctl + kWaveCtlCounter + l;
		HIPCHK(c, launch_dp_wave_pairs(P, c->o, z, l, st));
	}
	return HERO_OK;
}

} // extern "C"

// ---- small probes over the index ----------------------------------------------------------------
namespace {
__global__ void k_hash_kmers(const char *kmers, int n, int k, uint64_t P, uint64_t *out) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	uint64_t hi = 0, lo = 0;
	for (int t = 0; t < k; t++) {
		const uint64_t c = (uint64_t) ascii_code(kmers[(size_t) i * k + t]);
		if (t < 32) hi = (hi << 2) | c;
		else lo = (lo << 2) | c;
	}
	out[i] = hash_kmer(k, hi, lo, make_fastmod(P));
}
__global__ void k_lookup(DevIndex ix, const char *win, int n, int64_t *count, uint64_t *first_entry_idx) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	const int k = ix.k;
	uint64_t hi = 0, lo = 0;
	for (int t = 0; t < k; t++) {
		const uint64_t c = (uint64_t) ascii_code(win[(size_t) i * k + t]);
		if (t < 32) hi = (hi << 2) | c;
		else lo = (lo << 2) | c;
	}
	uint32_t start = 0, cnt = 0;
	const bool hit = ix_lookup(ix, hi, lo, start, cnt);
	count[i] = hit ? cnt : 0;
	first_entry_idx[i] = hit ? start : 0;
}
} // namespace

extern "C" int hero_hash_kmers(hero_ctx *c, int32_t n, int32_t k, const char *kmers, uint64_t P, uint64_t *out) {
	if (!c || n < 0 || k < 1 || k > 64 || P == 0 || (n > 0 && (!kmers || !out))) return HERO_ERR_ARG;
	if (n == 0) return HERO_OK;
	HIPCHK(c, hipSetDevice(c->opt.device));
	DevBuf<char> dk;
	DevBuf<uint64_t> dout;
	HIPCHK(c, dk.exact((size_t) n * k));
	HIPCHK(c, dout.exact((size_t) n));
	HIPCHK(c, hipMemcpyAsync(dk.p, kmers, (size_t) n * k, hipMemcpyHostToDevice, c->w.st));
	k_hash_kmers<<<(n + 255) / 256, 256, 0, c->w.st>>>(dk.p, n, k, P, dout.p);
	HIPCHK(c, hipGetLastError());
	HIPCHK(c, hipMemcpyAsync(out, dout.p, (size_t) n * 8, hipMemcpyDeviceToHost, c->w.st));
	HIPCHK(c, hipStreamSynchronize(c->w.st));
	return HERO_OK;
}

extern "C" int hero_lookup(hero_ctx *c, int32_t n, const char *windows, int64_t *count, int64_t *entry_off, uint64_t *entries,
		int64_t entry_cap) {
	if (!c || n < 0 || (n > 0 && (!windows || !count || !entry_off))) return HERO_ERR_ARG;
	if (!c->qix.built) return fail(c, HERO_ERR_STATE, "index not built");
	if (n == 0) return HERO_OK;
	HIPCHK(c, hipSetDevice(c->opt.device));
	const int k = c->o.k;
	DevBuf<char> dw;
	DevBuf<int64_t> dcount;
	DevBuf<uint64_t> dfirst;
	HIPCHK(c, dw.exact((size_t) n * k));
	HIPCHK(c, dcount.exact((size_t) n));
	HIPCHK(c, dfirst.exact((size_t) n));
	HIPCHK(c, hipMemcpyAsync(dw.p, windows, (size_t) n * k, hipMemcpyHostToDevice, c->w.st));
	const DevIndex ix = c->qix.view(k);
	k_lookup<<<(n + 255) / 256, 256, 0, c->w.st>>>(ix, dw.p, n, dcount.p, dfirst.p);
	HIPCHK(c, hipGetLastError());
	std::vector<uint64_t> first((size_t) n);
	HIPCHK(c, hipMemcpyAsync(count, dcount.p, (size_t) n * 8, hipMemcpyDeviceToHost, c->w.st));
	HIPCHK(c, hipMemcpyAsync(first.data(), dfirst.p, (size_t) n * 8, hipMemcpyDeviceToHost, c->w.st));
	HIPCHK(c, hipStreamSynchronize(c->w.st));
	int64_t pos = 0;
	for (int i = 0; i < n; i++) {
		entry_off[i] = pos;
		if (count[i] > 0 && entries) {
			if (pos + count[i] > entry_cap) return fail(c, HERO_ERR_LIMIT, "entry_cap too small");
			HIPCHK(c, hipMemcpy(entries + pos, c->qix.entries.p + first[(size_t) i], (size_t) count[i] * 8, hipMemcpyDeviceToHost));
			for (int64_t j = pos; j < pos + count[i]; j++) entries[j] &= kEntryMask; // (pos << 40) | innerId
		}
		pos += count[i];
	}
	return HERO_OK;
}

// ---------------------------------------------------------------------------------------------
// Serving: resident subjects, batch correction, and the request loop (continuous batching)
// ---------------------------------------------------------------------------------------------
extern "C" int hero_set_subjects(hero_ctx *c, int64_t n, const char *seq, const int64_t *off, const int32_t *len,
		const int64_t *uid) {
	if (!c) return HERO_ERR_ARG;
	if (int r = check_reads(c, n, seq, off, len, uid)) return r;
	if (n >= ((int64_t) 1 << 31)) return fail(c, HERO_ERR_LIMIT, "at most 2^31-1 resident subjects");
	HIPCHK(c, hipSetDevice(c->opt.device));
	c->r_len.assign(len, len + n);
	c->r_active.resize((size_t) n);
	c->r_max_len = 0;
	for (int64_t i = 0; i < n; i++) {
		c->r_active[(size_t) i] = subject_active(seq + off[i], len[i], c->o.min_overlap, c->o.k);
		c->r_max_len = std::max(c->r_max_len, len[i]);
	}
	c->six.release();
	c->Rd.release();
	return upload_reads(c, c->Rd, n, seq, off, len, uid, c->r_active);
}

extern "C" int hero_build_subject_index(hero_ctx *c) {
	if (!c) return HERO_ERR_ARG;
	HIPCHK(c, hipSetDevice(c->opt.device));
	if (int r = build_index(c, c->Rd, c->r_active, c->r_len, true, 0, c->six)) return r;
	HIPCHK(c, hipStreamSynchronize(c->w.st));
	return HERO_OK;
}

extern "C" int hero_correct_batch(hero_ctx *c, int64_t n, const char *seq, const int64_t *off, const int32_t *len,
		const int64_t *uid, char *corrected) {
	if (!c || (n > 0 && !corrected)) return HERO_ERR_ARG;
	if (c->o.merge) return fail(c, HERO_ERR_ARG, "serving does not take --merge");
	HIPCHK(c, hipSetDevice(c->opt.device));
	return correct_batch_on(c, &c->w, n, seq, off, len, uid, corrected);
}

namespace {

double now_s() {
	return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Request {
	int64_t n = 0;
	const char *seq = nullptr;
	const int64_t *off = nullptr;
	const int32_t *len = nullptr;
	const int64_t *uid = nullptr;
	char *out = nullptr;
	int64_t bases = 0, footprint = 0;
	hero_request_info info{};
	std::string err; // the step's error text when status != HERO_OK
	bool done = false;
};

} // namespace

// The request loop: submitted batches wait in arrival order; each worker (its own stream and
// buffers) takes the longest prefix of waiting batches that fits its share of the HBM budget and
// the step read limit, runs them as one step, and completes them. A worker starts its next step
// as soon as its previous one ends, so batches that arrive while others are on the device are
// admitted as soon as a worker and memory free up, and steps of different workers overlap.
struct hero_server {
	hero_ctx *c = nullptr;
	hero_server_options opt{};
	std::vector<Worker> workers;
	std::vector<hero_stats> acc;
	std::vector<std::thread> threads;
	std::mutex mu;
	std::condition_variable cv_work, cv_done;
	std::deque<int64_t> pending;
	// by ticket until hero_server_wait returns it (std::map keeps references valid across inserts)
	std::map<int64_t, Request> reqs;
	int64_t next_ticket = 0;
	std::atomic<int64_t> steps{0};
	int64_t inflight_bytes = 0;
	// device bytes a step needs per query base: kServeBytesPerBase until a step has run, then the
	// largest worker footprint over the bases of the largest step it ran (buffers are grow-only)
	int64_t bytes_per_base = 0;
	std::vector<int64_t> max_step_bases;
	bool learned = false;
	bool stop = false;
};

namespace {

// Device bytes one step needs per query base: seeds (keys and values, both sort buffers, chain
// arrays), tallies, DP tasks and step summaries, at 60x coverage with 150 bp reads (measured in
// the serving workload); the admission budget is charged this much per admitted base.
constexpr int64_t kServeBytesPerBase = 2048;

// One step over the requests `ids` on worker w: their reads merged, corrected, copied back.
// The requests are the worker's: taken under the lock as stable map nodes, erased only after they
// are done, so no lookup in the shared map happens here.
int serve_step(hero_server *sv, Worker *w, const std::vector<Request *> &ids, std::vector<char> &seq,
		std::vector<char> &out, std::vector<int64_t> &off, std::vector<int64_t> &uid, std::vector<int32_t> &len) {
	hero_ctx *c = sv->c;
	seq.clear(), off.clear(), uid.clear(), len.clear();
	for (const Request *rp : ids) {
		const Request &r = *rp;
		for (int64_t i = 0; i < r.n; i++) {
			off.push_back((int64_t) seq.size());
			len.push_back(r.len[i]);
			uid.push_back(r.uid[i]);
			seq.insert(seq.end(), r.seq + r.off[i], r.seq + r.off[i] + r.len[i]);
		}
	}
	out.resize(seq.size());
	const int rc = correct_batch_on(c, w, (int64_t) len.size(), seq.data(), off.data(), len.data(), uid.data(), out.data());
	if (rc != HERO_OK) return rc;
	size_t i0 = 0;
	for (const Request *rp : ids) {
		const Request &r = *rp;
		for (int64_t i = 0; i < r.n; i++) std::memcpy(r.out + r.off[i], out.data() + off[i0 + (size_t) i], (size_t) r.len[i]);
		i0 += (size_t) r.n;
	}
	return HERO_OK;
}

void serve_loop(hero_server *sv, int wi) {
	hero_ctx *c = sv->c;
	Worker *w = &sv->workers[(size_t) wi];
	(void) hipSetDevice(c->opt.device);
	std::vector<char> seq, out;
	std::vector<int64_t> off, uid;
	std::vector<int32_t> len;
	const int64_t share = sv->opt.hbm_budget / std::max(1, sv->opt.workers);
	for (;;) {
		std::vector<Request *> take;
		int64_t fp = 0, bases = 0;
		{
			std::unique_lock<std::mutex> lk(sv->mu);
			sv->cv_work.wait(lk, [&] { return sv->stop || !sv->pending.empty(); });
			if (sv->pending.empty()) return; // stopping with nothing left
			int64_t reads = 0;
			while (!sv->pending.empty()) {
				Request &r = sv->reqs.at(sv->pending.front());
				const int64_t rfp = r.bases * sv->bytes_per_base;
				const bool first = take.empty();
				if (!first && (reads + r.n > sv->opt.max_step_reads || fp + rfp > share)) break;
				if (!first && sv->inflight_bytes + fp + rfp > sv->opt.hbm_budget) break;
				if (first && sv->inflight_bytes + rfp > sv->opt.hbm_budget && sv->inflight_bytes > 0) break;
				take.push_back(&r);
				sv->pending.pop_front();
				reads += r.n, fp += rfp, bases += r.bases;
			}
			if (take.empty()) {
				// memory is held by other steps: wait for one to finish
				sv->cv_done.wait(lk);
				continue;
			}
			sv->inflight_bytes += fp;
			const double t = now_s();
			for (Request *rp : take) rp->info.t_admit = t;
		}
		const int64_t step_id = sv->steps.fetch_add(1);
		std::vector<int> status(take.size(), HERO_OK);
		int rc = serve_step(sv, w, take, seq, out, off, uid, len);
		if (rc == HERO_ERR_NOMEM && take.size() > 1) {
			// the merged step did not fit: free this worker's buffers, then its batches one by one
			w->release_buffers();
			for (size_t t = 0; t < take.size(); t++) status[t] = serve_step(sv, w, {take[t]}, seq, out, off, uid, len);
			rc = HERO_OK;
		} else {
			for (int &x : status) x = rc;
		}
		if (rc == HERO_OK) {
			// learn the footprint: the worker's buffers over the bases of the largest step it ran
			std::lock_guard<std::mutex> lk(sv->mu);
			int64_t &mb = sv->max_step_bases[(size_t) wi];
			mb = std::max(mb, bases);
			if (mb > 0) {
				const int64_t per = (int64_t) w->device_bytes() / mb + 1;
				sv->bytes_per_base = sv->learned ? std::max(sv->bytes_per_base, per) : per;
				sv->learned = true;
			}
		}
		{
			std::lock_guard<std::mutex> lk(sv->mu);
			std::string msg;
			for (size_t t = 0; t < take.size() && msg.empty(); t++)
				if (status[t] != HERO_OK) {
					std::lock_guard<std::mutex> ek(c->err_mu);
					msg = c->err;
				}
			const double t = now_s();
			for (size_t x = 0; x < take.size(); x++) {
				Request &r = *take[x];
				if (status[x] != HERO_OK) r.err = msg;
				r.info.status = status[x], r.info.t_done = t, r.info.step_id = step_id;
				r.info.step_reads = (int32_t) len.size(), r.info.step_batches = (int32_t) take.size();
				r.done = true;
			}
			sv->inflight_bytes -= fp;
		}
		sv->cv_done.notify_all();
	}
}

} // namespace

extern "C" int hero_server_start(hero_ctx *c, const hero_server_options *opt, hero_server **out) {
	if (!c || !out) return HERO_ERR_ARG;
	*out = nullptr;
	if (!c->six.built) return fail(c, HERO_ERR_STATE, "hero_build_subject_index must run before serving");
	if (!c->o.call_consensus) return fail(c, HERO_ERR_ARG, "serving corrects reads: call_consensus must be 1");
	if (c->o.merge) return fail(c, HERO_ERR_ARG, "serving does not take --merge");
	HIPCHK(c, hipSetDevice(c->opt.device));
	hero_server *sv = new hero_server();
	sv->c = c;
	if (opt) sv->opt = *opt;
	if (sv->opt.workers <= 0) sv->opt.workers = 2;
	if (sv->opt.max_step_reads <= 0) sv->opt.max_step_reads = 32768;
	if (sv->opt.hbm_budget <= 0) {
		size_t fr = 0, tot = 0;
		(void) hipMemGetInfo(&fr, &tot);
		sv->opt.hbm_budget = (int64_t) (fr / 2);
	}
	sv->bytes_per_base = kServeBytesPerBase;
	sv->max_step_bases.assign((size_t) sv->opt.workers, 0);
	sv->workers.resize((size_t) sv->opt.workers);
	sv->acc.assign((size_t) sv->opt.workers, hero_stats{});
	for (int i = 0; i < sv->opt.workers; i++) {
		sv->workers[(size_t) i].acc = &sv->acc[(size_t) i];
		if (hipError_t e = sv->workers[(size_t) i].init(); e != hipSuccess) {
			for (auto &w : sv->workers) w.release();
			delete sv;
			return fail(c, HERO_ERR_HIP, "server worker: %s", hipGetErrorString(e));
		}
	}
	for (int i = 0; i < sv->opt.workers; i++) sv->threads.emplace_back(serve_loop, sv, i);
	*out = sv;
	return HERO_OK;
}

extern "C" int hero_server_submit(hero_server *sv, int64_t n, const char *seq, const int64_t *off, const int32_t *len,
		const int64_t *uid, char *corrected, int64_t *ticket) {
	if (!sv || !ticket || n < 0 || (n > 0 && (!seq || !off || !len || !uid || !corrected))) return HERO_ERR_ARG;
	Request r;
	r.n = n, r.seq = seq, r.off = off, r.len = len, r.uid = uid, r.out = corrected;
	for (int64_t i = 0; i < n; i++) r.bases += len[i];
	r.info.t_submit = now_s();
	{
		std::lock_guard<std::mutex> lk(sv->mu);
		if (sv->stop) return HERO_ERR_STATE;
		r.footprint = r.bases * sv->bytes_per_base;
		if (r.footprint > sv->opt.hbm_budget) return HERO_ERR_NOMEM; // would never be admitted
		*ticket = sv->next_ticket++;
		sv->reqs.emplace(*ticket, r);
		sv->pending.push_back(*ticket);
	}
	sv->cv_work.notify_one();
	return HERO_OK;
}

extern "C" int hero_server_wait(hero_server *sv, int64_t ticket, hero_request_info *info) {
	if (!sv) return HERO_ERR_ARG;
	std::unique_lock<std::mutex> lk(sv->mu);
	if (sv->reqs.find(ticket) == sv->reqs.end()) return HERO_ERR_ARG; // unknown, or already waited for
	sv->cv_done.wait(lk, [&] { return sv->reqs.at(ticket).done; });
	const hero_request_info ri = sv->reqs.at(ticket).info;
	if (ri.status != HERO_OK) { // the batch's own error: its status here, its text in hero_last_error
		std::lock_guard<std::mutex> ek(sv->c->err_mu);
		sv->c->err = sv->reqs.at(ticket).err;
	}
	sv->reqs.erase(ticket); // a ticket is returned once; its slot is freed
	if (info) *info = ri;
	return ri.status;
}

extern "C" int hero_server_stop(hero_server *sv, hero_stats *merged) {
	if (!sv) return HERO_ERR_ARG;
	{
		std::lock_guard<std::mutex> lk(sv->mu);
		sv->stop = true;
	}
	sv->cv_work.notify_all();
	for (auto &t : sv->threads) t.join();
	if (merged) {
		std::memset(merged, 0, sizeof *merged);
		for (const hero_stats &a : sv->acc) {
			merged->ms_seed += a.ms_seed, merged->ms_plan += a.ms_plan, merged->ms_dp += a.ms_dp;
			merged->ms_score_vote += a.ms_score_vote, merged->ms_consensus += a.ms_consensus;
			merged->queries += a.queries, merged->dp_tasks += a.dp_tasks, merged->subjects += a.subjects;
		}
		merged->ms_total = merged->ms_seed + merged->ms_plan + merged->ms_dp + merged->ms_score_vote + merged->ms_consensus;
	}
	for (auto &w : sv->workers) w.release();
	delete sv; // a failed batch reported its error through hero_server_wait
	return HERO_OK;
}

extern "C" int hero_autotune(hero_ctx *c, int64_t max_subjects, hero_tune_report *report) {
	if (!c || max_subjects < 0) return HERO_ERR_ARG;
	if (!c->qix.built) return fail(c, HERO_ERR_STATE, "hero_build_index must run before hero_autotune");
	HIPCHK(c, hipSetDevice(c->opt.device));
	hero_tune_report rep;
	std::memset(&rep, 0, sizeof rep);
	rep.classes = kNumBandClasses;
	for (int i = 0; i < kNumBandClasses; i++) c->tuned_per_lane[i] = 1, c->tuned_waves_per_cu[i] = 0;
	int ml = 0;
	for (int32_t l : c->q_len) ml = std::max(ml, l);
	const int64_t n = std::min<int64_t>(max_subjects > 0 ? max_subjects : 65536, c->Qd.n);
	// tuning work is not part of the run's statistics: keep the counters and restore them after
	const hero_stats saved = c->st_host;
	std::vector<unsigned long long> dsaved((size_t) kStatStripes * kStatStride);
	HIPCHK(c, hipMemcpy(dsaved.data(), c->w.dstats.p, dsaved.size() * 8, hipMemcpyDeviceToHost));
	c->tune = &rep;
	const int rc = n > 0 ? run_chunk(c, &c->w, view(c->Qd, c->q_subject_good.p), 0, (int32_t) n, ml + 1, c->q_uid.data())
	                     : HERO_OK;
	c->tune = nullptr;
	c->st_host = saved;
	HIPCHK(c, hipMemcpy(c->w.dstats.p, dsaved.data(), dsaved.size() * 8, hipMemcpyHostToDevice));
	if (rc != HERO_OK) return rc;
	rep.subjects = n;
	if (report) *report = rep;
	return HERO_OK;
}
