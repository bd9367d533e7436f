// Per-group steps of the correction path shared by the kernels and the CPU harness:
//   plan    (pass A)  chain one (subject, strand, query) seed group and emit the DP tasks of
//                     every surviving chain (processOneSubjectRead2 + getAlignment2 geometry);
//   score   (pass B)  build the query's records over both strands from the DP results, apply
//                     early exits and acceptance, keep the first best record
//                     (SubjectRead::addAlignment under -c);
//   vote    (pass B)  cast the kept record's votes (SubjectRead::updateQueryCoverage).
#pragma once
#include "dp_core.h"
#include "record_core.h"

namespace hero {

// Seeds and chain state of one group: seeds[0..gn) sorted; chain arrays of group size.
struct Group {
	const uint64_t *seeds;
	int32_t *cnext, *cfirst, *clast, *clen;
	int gn;
	int nch;  // chains after chain_group
	int best; // best covered length
};

struct SeedAt {
	const uint64_t *s;
	HD uint64_t operator()(int t) const { return s[t]; }
};
struct I32Arr {
	int32_t *p;
	HD int32_t &operator[](int i) const { return p[i]; }
};

HD void chain(const Opts &o, Group &g) {
	SeedAt sa{g.seeds};
	I32Arr nx{g.cnext}, cf{g.cfirst}, cl{g.clast}, ln{g.clen};
	g.nch = chain_group(o, g.gn, sa, nx, cf, cl, ln, &g.best);
}

// Reads of one (subject, strand, query) pair and their frames (word pointers cached).
struct PairReads {
	int64_t s_read, q_read;
	int Ls, Lq;
	bool subj_rc; // record subject frame is the subject's reverse complement (-c, reverse strand)
	bool qry_rc;  // record query frame is the query's reverse complement (no -c, reverse strand)
	const uint64_t *sw, *qw;
	HD int sbase(int p) const { return subj_rc ? (base_at(sw, Ls - 1 - p) ^ 2) : base_at(sw, p); }
	HD int qbase(int p) const { return qry_rc ? (base_at(qw, Lq - 1 - p) ^ 2) : base_at(qw, p); }
};

HD PairReads make_pair_reads(const Packed &S, const Packed &Q, bool call_consensus, int64_t s, int Ls, int64_t q,
		int strand) {
	PairReads pr;
	pr.s_read = s, pr.q_read = q, pr.Ls = Ls, pr.Lq = Q.len[q];
	pr.subj_rc = call_consensus && strand == 1;
	pr.qry_rc = !call_consensus && strand == 1;
	pr.sw = S.words + S.woff[s];
	pr.qw = Q.words + Q.woff[q];
	return pr;
}

// ---- plan ---------------------------------------------------------------------------------
// s_woff / q_woff: the subject and query reads' first words in their sets (a task locates its
// segments by word indices in them).
HD DpTask make_task(const PairReads &pr, const DpCall &c, uint32_t out_off, uint32_t s_woff, uint32_t q_woff) {
	DpTask t;
	t.s_len = (uint16_t) (c.s1 - c.s0);
	t.q_len = (uint16_t) (c.q1 - c.q0);
	segment_start(s_woff, pr.Ls, c.s0, t.s_len, pr.subj_rc, t.s_word, t.s_bit);
	segment_start(q_woff, pr.Lq, c.q0, t.q_len, pr.qry_rc, t.q_word, t.q_bit);
	t.band = (uint16_t) c.band;
	t.flags = (uint8_t) ((c.flag == 1 ? kTaskAlignLeft : 0) | (c.flag == 2 ? kTaskAlignRight : 0) |
			(c.swapped ? kTaskSwapped : 0) | (pr.subj_rc ? kTaskSubjRC : 0) | (pr.qry_rc ? kTaskQryRC : 0));
	t._r = 0;
	t.out_off = out_off;
	t._r2 = 0;
	t._r3 = 0;
	return t;
}

// Record scripts. Planning walks each surviving chain's getAlignment2 geometry once (it never
// depends on DP results) and keeps it as a script: a header and one word per step of the walk.
// Scoring, votes and records then replay scripts against the DP results without revisiting
// seeds or chains. A group's scripts follow one count word: [nscripts, script 0, script 1, ...];
// script = [nops | ndp << 16, max_mis, max_indel, pos, op * nops]; op = kind << 30 | payload:
//   add     type << 27 | count        a fixed CIGAR op
//   dp      swapped << 29 | cells     the next DP call of the record (its task follows the last)
//   check                            isGoodAlignment(max_mis, max_indel) (early exit)
//   finish                           isGoodAlignment() on rates
// nops == 0: the walk stopped before building anything (no keys, or overlap < minOverlap).
enum : uint32_t { kScAdd = 0u, kScDp = 1u, kScCheck = 2u, kScFinish = 3u };
constexpr int kScriptHead = 4;

struct ScriptWriter {
	uint32_t *out;       // script start; null: count only
	const PairReads *pr; // with tasks: also write the DP tasks
	DpTask *tasks;
	uint32_t s_woff, q_woff; // with tasks: the reads' first words in their sets
	uint32_t next_task, next_word;
	int nops, ndp;
	uint32_t max_mis, max_indel, pos;
	HD void put(uint32_t w) {
		if (out) out[kScriptHead + nops] = w;
		nops++;
	}
	HD void start(int p) { pos = (uint32_t) p; }
	HD void add(int c, int t) { put((kScAdd << 30) | ((uint32_t) t << 27) | ((uint32_t) c & 0x7FFFFFFu)); }
	HD void dp(const DpCall &c) {
		const uint64_t cells = band_cells(c.lenS, c.lenQ, c.band);
		put((kScDp << 30) | (c.swapped ? (1u << 29) : 0u) | (uint32_t) (cells < (1u << 29) ? cells : (1u << 29) - 1));
		if (tasks) tasks[next_task] = make_task(*pr, c, next_word, s_woff, q_woff);
		next_task++;
		next_word += (uint32_t) edit_words(c.lenS, c.lenQ);
		ndp++;
	}
	HD bool check(int64_t m, int64_t i) {
		max_mis = (uint32_t) m, max_indel = (uint32_t) i;
		put(kScCheck << 30);
		return true;
	}
	HD bool finish() {
		put(kScFinish << 30);
		return true;
	}
	HD void head() const {
		if (out) out[0] = (uint32_t) nops | ((uint32_t) ndp << 16), out[1] = max_mis, out[2] = max_indel, out[3] = pos;
	}
};

HD int script_nops(const uint32_t *sc) { return (int) (sc[0] & 0xFFFFu); }
HD int script_ndp(const uint32_t *sc) { return (int) (sc[0] >> 16); }
HD int script_words(const uint32_t *sc) { return kScriptHead + script_nops(sc); }

// DP tasks, edit words and script words of all surviving chains of a chained group.
HD void plan_count(const Opts &o, const Group &g, const PairReads &pr, int &ntasks, int &nwords, int &nprog,
		int &nscripts) {
	ntasks = 0, nwords = 0, nprog = 1, nscripts = 0;
	SeedAt sa{g.seeds};
	I32Arr nx{g.cnext};
	for (int c = 0; c < g.nch; c++) {
		if (g.clen[c] != g.best) continue;
		ChainKeys<SeedAt, I32Arr> keys(sa, nx, g.cfirst[c], g.clast[c]);
		ScriptWriter w{nullptr, nullptr, nullptr, 0, 0, 0, 0, 0, 0, 0, 0, 0};
		walk_record(o, pr.Ls, pr.Lq, keys, w);
		ntasks += w.ndp, nwords += (int) w.next_word, nprog += kScriptHead + w.nops, nscripts++;
	}
}

// Writes the tasks (at tasks[task0...], edit words from word0) and the scripts (at prog) of all
// surviving chains, in chain order. s_woff / q_woff: the reads' first words in their sets.
HD void plan_emit(const Opts &o, const Group &g, const PairReads &pr, DpTask *tasks, uint32_t task0, uint32_t word0,
		uint32_t *prog, uint32_t s_woff, uint32_t q_woff) {
	SeedAt sa{g.seeds};
	I32Arr nx{g.cnext};
	uint32_t next_task = task0, next_word = word0, p = 1, n = 0;
	for (int c = 0; c < g.nch; c++) {
		if (g.clen[c] != g.best) continue;
		ChainKeys<SeedAt, I32Arr> keys(sa, nx, g.cfirst[c], g.clast[c]);
		ScriptWriter w{prog + p, &pr, tasks, s_woff, q_woff, next_task, next_word, 0, 0, 0, 0, 0};
		walk_record(o, pr.Ls, pr.Lq, keys, w);
		w.head();
		next_task = w.next_task, next_word = w.next_word;
		p += (uint32_t) (kScriptHead + w.nops), n++;
	}
	prog[0] = n;
}

// ---- DP results ---------------------------------------------------------------------------
// Task t's steps sit at words[t.out_off ...], capacity cap = s_len + q_len steps; the n steps of
// the traceback occupy [cap - n, cap) in forward (first-to-last) order, 16 per word. sums[t] is
// their run summary.
struct DpResults {
	const DpTask *tasks;
	const uint32_t *words;
	const uint32_t *nsteps;
	const StepSummary *sums;
};

// Record op type of each step code for one call: a left step consumes the aligner's query, an
// up step its subject, which are the record's subject / query when the pair was swapped
// (calculateCigarAndPositionReversed, src/BandAligner.cpp:339-419).
HD int step_type(uint32_t code, bool swapped) {
	if (code < 2) return code == kStepMatch ? kOpM : kOpX;
	return ((code == kStepLeft) != swapped) ? kOpI : kOpD;
}

// Feeds one DP call's steps in order as record ops (addAlignment, src/AlignmentRecord.cpp:50-54),
// as runs: op(count, type) for each maximal run of equal op types (a mismatch run is one call
// with its length). Within a 16-step word the loop visits only the non-match steps (match = 0, so
// the lowest set bit pair is the next one); the matches between them extend the current run.
// A call of up to kFeedWords words (128 steps) has all its words loaded before any is used, so
// their loads overlap instead of each waiting for the previous one.
constexpr int kFeedWords = 8;
template <class F>
HD void feed_runs(const DpResults &R, uint32_t task, bool swapped, F &&op) {
	const DpTask &t = R.tasks[task];
	const int cap = t.s_len + t.q_len;
	const int n = (int) R.nsteps[task] <= cap ? (int) R.nsteps[task] : 0;
	const uint32_t *w = R.words + t.out_off;
	int e = cap - n;
	int run_t = -1, run_c = 0;
	auto push = [&](int c, int ty) {
		if (ty == run_t) {
			run_c += c;
		} else {
			if (run_c) op(run_c, run_t);
			run_t = ty, run_c = c;
		}
	};
	static_assert(kStepMatch == 0, "match steps are zero bits");
	auto steps_of = [&](uint32_t raw) { // the steps e .. up to the end of e's word
		const int stop = (e | 15) + 1 < cap ? (e | 15) + 1 : cap;
		const int m = stop - e;
		uint32_t word = raw >> (2 * (e & 15));
		if (m < 16) word &= (1u << (2 * m)) - 1u;
		int done = 0;
		while (word) {
			const int j = __builtin_ctz(word) >> 1; // matches before the next non-match step
			if (j) push(j, kOpM);
			word >>= 2 * j;
			push(1, step_type(word & 3u, swapped));
			word >>= 2;
			done += j + 1;
		}
		if (m > done) push(m - done, kOpM);
		e = stop;
	};
	const int w0 = e >> 4, w1 = (cap - 1) >> 4; // words holding steps e .. cap-1
	if (w1 - w0 < kFeedWords) {
		uint32_t pre[kFeedWords];
#pragma unroll
		for (int k = 0; k < kFeedWords; k++) pre[k] = w0 + k <= w1 ? w[w0 + k] : 0u;
#pragma unroll
		for (int k = 0; k < kFeedWords; k++)
			if (e < cap) steps_of(pre[k]);
	} else {
		while (e < cap) steps_of(w[e >> 4]);
	}
	if (run_c) op(run_c, run_t);
}

// ---- score --------------------------------------------------------------------------------
// One script replayed against the DP summaries (getAlignment2's checks in order, then the rates):
// true when the record is accepted. cells / calls count the DP calls the reference executes.
HD bool score_script(const uint32_t *sc, const DpResults &R, uint32_t task, double mismatch_rate, double indel_rate,
		double &rate_sum, double &len, uint64_t &cells, uint64_t &calls) {
	const int nops = script_nops(sc);
	if (nops == 0) return false;
	const int64_t max_mis = sc[1], max_indel = sc[2];
	Score s;
	s.reset();
	for (int i = 0; i < nops; i++) {
		const uint32_t w = sc[kScriptHead + i];
		const uint32_t kind = w >> 30;
		if (kind == kScAdd) {
			s.add((int) (w & 0x7FFFFFFu), (int) ((w >> 27) & 7u));
		} else if (kind == kScDp) {
			s.add_piece(R.sums[task++], ((w >> 29) & 1u) != 0);
			cells += w & 0x1FFFFFFFu;
			calls++;
		} else if (kind == kScCheck) {
			if (!s.good_caps(max_mis, max_indel)) return false;
		} else {
			return s.good_rates(mismatch_rate, indel_rate, rate_sum, len);
		}
	}
	return false;
}

// Best record of one (subject, query) pair so far: the first record with the lowest rate sum,
// then the longest alignment (SubjectRead::addAlignment, src/SubjectRead.cpp:25-79).
struct Best {
	bool have;
	double rate_sum, len;
	int strand;
	uint32_t script; // winning record's script (index into the program words)
	uint32_t task0;  // its first DP task
};

// Scores the scripts of one group (every best-length chain, in chain order:
// processOneSubjectRead2, src/SubjectReadAligner.cpp:838-849) starting at prog[gp];
// accept(script index in the group, script offset, first task, rate sum, length) sees every
// accepted record.
struct NoAccept {
	HD void operator()(int, uint32_t, uint32_t, double, double) const {}
};
template <class F = NoAccept>
HD void score_group(const Opts &o, const uint32_t *prog, uint32_t gp, const DpResults &R, uint32_t task0, int strand,
		Best &best, uint64_t &cells_exec, uint64_t &calls_exec, uint64_t &built, F &&accept = F()) {
	const uint32_t n = prog[gp];
	uint32_t p = gp + 1, task = task0;
	for (uint32_t k = 0; k < n; k++) {
		const uint32_t *sc = prog + p;
		double rate_sum = 0, len = 0;
		if (score_script(sc, R, task, o.mismatch_rate, o.indel_rate, rate_sum, len, cells_exec, calls_exec)) {
			built++;
			accept((int) k, p, task, rate_sum, len);
			if (!best.have || record_better(rate_sum, len, best.rate_sum, best.len)) {
				best.have = true, best.rate_sum = rate_sum, best.len = len;
				best.strand = strand, best.script = p, best.task0 = task;
			}
		}
		task += (uint32_t) script_ndp(sc);
		p += (uint32_t) script_words(sc);
	}
}

// Replays an accepted script through a visitor: start(pos), add(count, type) for fixed ops,
// run(task, swapped) for each DP call, finish().
template <class V>
HD void replay_script(const uint32_t *sc, uint32_t task, V &v) {
	const int nops = script_nops(sc);
	v.start((int) sc[3]);
	for (int i = 0; i < nops; i++) {
		const uint32_t w = sc[kScriptHead + i];
		const uint32_t kind = w >> 30;
		if (kind == kScAdd) v.add((int) (w & 0x7FFFFFFu), (int) ((w >> 27) & 7u));
		else if (kind == kScDp) v.run(task++, ((w >> 29) & 1u) != 0);
	}
	v.finish();
}

// ---- votes --------------------------------------------------------------------------------
// A query's votes: `tally` holds 4 counters per base (votes for each base code cast by
// mismatches), `diff` one counter per base, a difference array whose prefix sum at p counts the
// records that align a subject base to query base p (by a match or a mismatch). Every query base
// of a record's span [0, end) is a match, a mismatch or inserted (aligned to nothing), so a record
// adds +1 at 0 and -1 at end, and an insertion run -1 at its first base and +1 after its last:
// two updates per record and per insertion, none per mismatch, and no lookup of the query's own
// bases. The match votes at p (the reference's votes for the own base) are then the prefix sum
// minus the mismatch votes at p; consensus_read and k_consensus decode it.
struct VoteSink {
	uint32_t *tally; // 4 per base
	uint32_t *diff;  // 1 per base
	PairReads pr;
	unsigned long long *oob; // counts votes that would leave a read (the reference throws out_of_range)
	HD void add_u32(uint32_t *p, uint32_t v) const {
#if defined(__HIP_DEVICE_COMPILE__)
		atomicAdd(p, v);
#else
		__atomic_fetch_add(p, v, __ATOMIC_RELAXED);
#endif
	}
	HD void flag_oob() const {
#if defined(__HIP_DEVICE_COMPILE__)
		atomicAdd(oob, 1ull);
#else
		__atomic_fetch_add(oob, 1ull, __ATOMIC_RELAXED);
#endif
	}
	// a record's span: query bases [0, end)
	HD void span(int end) const {
		if (end <= 0) return;
		if (end > pr.Lq) flag_oob(), end = pr.Lq;
		add_u32(&diff[0], 1u);
		if (end < pr.Lq) add_u32(&diff[end], 0xFFFFFFFFu);
	}
	// query bases [p, p + c) inserted (aligned to no subject base)
	HD void hole(int p, int c) const {
		if (c <= 0) return;
		if (p < 0 || p + c > pr.Lq) {
			flag_oob();
			return;
		}
		add_u32(&diff[p], 0xFFFFFFFFu);
		if (p + c < pr.Lq) add_u32(&diff[p + c], 1u);
	}
	HD void mismatch(int pq, int ps) const {
		if (pq < 0 || pq >= pr.Lq || ps < 0 || ps >= pr.Ls) {
			flag_oob();
			return;
		}
		add_u32(&tally[4 * pq + pr.sbase(ps)], 1u);
	}
};

// One DP call's steps cast straight into a record's votes: the steps of feed_runs, without merging
// them into runs first. A mismatch votes per base and a deletion only moves the subject position,
// so runs of those need no merging; an insertion run stays one hole (its first base and the base
// after it), as the merged op makes it.
template <class V>
HD void vote_steps(const DpResults &R, uint32_t task, bool swapped, V &v) {
	const DpTask &t = R.tasks[task];
	const int cap = t.s_len + t.q_len;
	const int n = (int) R.nsteps[task] <= cap ? (int) R.nsteps[task] : 0;
	const uint32_t *w = R.words + t.out_off;
	int e = cap - n;
	int ins = 0; // bases of the open insertion run, ending at v.pq
	auto close_ins = [&]() {
		if (ins) v.sink.hole(v.pq - ins, ins), ins = 0;
	};
	auto steps_of = [&](uint32_t raw) { // the steps e .. up to the end of e's word
		const int stop = (e | 15) + 1 < cap ? (e | 15) + 1 : cap;
		const int m = stop - e;
		uint32_t word = raw >> (2 * (e & 15));
		if (m < 16) word &= (1u << (2 * m)) - 1u;
		int done = 0;
		while (word) {
			const int j = __builtin_ctz(word) >> 1; // matches before the next non-match step
			if (j) {
				close_ins();
				v.pq += j, v.ps += j;
			}
			word >>= 2 * j;
			const int ty = step_type(word & 3u, swapped);
			if (ty == kOpI) {
				ins++, v.pq++;
			} else {
				close_ins();
				if (ty == kOpX) v.sink.mismatch(v.pq, v.ps), v.pq++;
				v.ps++;
			}
			word >>= 2;
			done += j + 1;
		}
		if (m > done) {
			close_ins();
			v.pq += m - done, v.ps += m - done;
		}
		e = stop;
	};
	const int w0 = e >> 4, w1 = (cap - 1) >> 4; // words holding steps e .. cap-1
	if (w1 - w0 < kFeedWords) {
		uint32_t pre[kFeedWords];
#pragma unroll
		for (int k = 0; k < kFeedWords; k++) pre[k] = w0 + k <= w1 ? w[w0 + k] : 0u;
#pragma unroll
		for (int k = 0; k < kFeedWords; k++)
			if (e < cap) steps_of(pre[k]);
	} else {
		while (e < cap) steps_of(w[e >> 4]);
	}
	close_ins();
}

struct VoteV {
	const DpResults &R;
	const VoteSink &sink;
	int ps, pq; // record frames (iPosOnSubjectRead, iPosOnQueryRead)
	HD void op(int c, int t) {
		if (t == kOpM) {
			pq += c, ps += c;
		} else if (t == kOpI) {
			sink.hole(pq, c);
			pq += c;
		} else if (t == kOpD) {
			ps += c;
		} else {
			for (int x = 0; x < c; x++) sink.mismatch(pq + x, ps + x);
			pq += c, ps += c;
		}
	}
	HD void start(int pos) { ps = pos, pq = 0; }
	HD void add(int c, int t) { op(c, t); }
	HD void run(uint32_t task, bool swapped) { vote_steps(R, task, swapped, *this); }
	HD void finish() { sink.span(pq); }
};

HD void vote_record(const uint32_t *script, const DpResults &R, uint32_t task0, const VoteSink &sink) {
	VoteV vv{R, sink, 0, 0};
	replay_script(script, task0, vv);
}

// --merge (SubjectRead::updateQueryCoverageConsiderMerging, src/SubjectRead.cpp:332-461): the
// record's votes, and when its CIGAR has more than one op (AlignmentRecord::addCigar merging, as the
// record writer counts them) the subject bases from the record's end, up to kExtPositions of them and
// in the record's subject frame, each vote into row (index, base) of the query's extension tally
// (QueryRead::updateExtraCoverageMatrix, src/QueryRead.cpp:353-371).
constexpr int kExtPositions = 20; // QueryRead::iLength4ExtraMatrix
struct MergeVoteV {
	VoteV vote;
	int nops;      // CIGAR ops after merging equal adjacent types
	int cur_t;     // type of the open op ('M', 'I', 'D' or the query base of a mismatch)
	bool pending;
	HD void push(int t) {
		if (!pending || cur_t != t) nops++;
		pending = true, cur_t = t;
	}
	HD void op(int c, int t) {
		if (t == kOpM) push('M');
		else if (t == kOpI) push('I');
		else if (t == kOpD) push('D');
		else
			for (int x = 0; x < c; x++) push(code_ascii(vote.sink.pr.qbase(vote.pq + x)));
		vote.op(c, t);
	}
	HD void start(int pos) { vote.start(pos); }
	HD void add(int c, int t) { op(c, t); }
	HD void run(uint32_t task, bool swapped) {
		feed_runs(vote.R, task, swapped, [&](int cnt, int t) { op(cnt, t); });
	}
	HD void finish() { vote.finish(); }
};

HD void vote_record_merging(const uint32_t *script, const DpResults &R, uint32_t task0, const VoteSink &sink,
		uint32_t *ext) {
	MergeVoteV mv{VoteV{R, sink, 0, 0}, 0, 0, false};
	replay_script(script, task0, mv);
	if (mv.nops <= 1) return; // a one-op record stops before the extension (src/SubjectRead.cpp:416-418)
	const PairReads &pr = sink.pr;
	for (int x = 0, ps = mv.vote.ps; x < kExtPositions && ps >= 0 && ps < pr.Ls; x++, ps++)
		sink.add_u32(&ext[4 * x + pr.sbase(ps)], 1u);
}

// ---- records (without -c) -------------------------------------------------------------------
// The record's CIGAR as AlignmentRecord::addCigar builds it (src/AlignmentRecord.cpp:35-54):
// adjacent ops of equal type merge (zero counts included); a mismatch step is typed by the
// record query's base (calculateCigarAndPosition[Reversed], src/BandAligner.cpp:272-278,353-360).
struct RecordV {
	const DpResults &R;
	PairReads pr;
	uint32_t *out; // null: count only
	int n;         // ops emitted
	int ps, pq, pos;
	bool pending;
	int cur_t, cur_c;
	HD void emit() {
		if (out) out[n] = ((uint32_t) cur_c << 8) | (uint32_t) cur_t;
		n++;
	}
	HD void push(int t, int c) {
		if (pending && cur_t == t) {
			cur_c += c;
		} else {
			if (pending) emit();
			pending = true, cur_t = t, cur_c = c;
		}
	}
	HD void op(int c, int t) {
		if (t == kOpM) {
			push('M', c), pq += c, ps += c;
		} else if (t == kOpI) {
			push('I', c), pq += c;
		} else if (t == kOpD) {
			push('D', c), ps += c;
		} else {
			for (int x = 0; x < c; x++, pq++, ps++) push(code_ascii(pr.qbase(pq)), 1);
		}
	}
	HD void start(int p) { ps = p, pq = 0, pos = p; }
	HD void add(int c, int t) { op(c, t); }
	HD void run(uint32_t task, bool swapped) { feed_runs(R, task, swapped, [&](int cnt, int t) { op(cnt, t); }); }
	HD void finish() {
		if (pending) emit();
		pending = false;
	}
};

// Ops of one accepted record (its script, DP results from task0): returns the op count and
// writes them to out unless it is null; *pos receives iPositionOfQueryOnSubject.
HD int record_ops(const uint32_t *script, const PairReads &pr, const DpResults &R, uint32_t task0, uint32_t *out,
		int *pos) {
	RecordV rv{R, pr, out, 0, 0, 0, 0, false, 0, 0};
	replay_script(script, task0, rv);
	*pos = rv.pos;
	return rv.n;
}

// ---- consensus ----------------------------------------------------------------------------
// QueryRead::callConsensus (src/QueryRead.cpp:148-180) over a difference-coded tally: the own
// base's count is the running sum of its slot plus the self vote; aDNA = {A,C,T,G} (:11) is
// exactly the code order. Writes the corrected read as codes; returns changed bases.
// One position: counts c (own base already decoded and self-voted) -> the consensus code.
HD int consensus_base(const Opts &o, const uint32_t c[4], int own) {
	double sum = 0;
	for (int j = 0; j < 4; j++) sum += c[j];
	int id = -1;
	double largest = 0, larger = 0;
	for (int j = 0; j < 4; j++) {
		const double f = ((double) c[j]) / sum;
		if (largest < f) larger = largest, largest = f, id = j;
	}
	return (largest - larger >= o.consensus_freq) ? id : own;
}

HD int consensus_read(const Opts &o, const uint32_t *tally, const uint32_t *diff, int L, const Packed &Q, int64_t r,
		uint8_t *out_codes, uint32_t *decoded /* optional 4*L, pre-self-vote counts */) {
	uint32_t run = 0;
	int changed = 0;
	const uint64_t *w = Q.words + Q.woff[r];
	for (int p = 0; p < L; p++) {
		const int own = base_at(w, p);
		uint32_t c[4];
		for (int j = 0; j < 4; j++) c[j] = tally[4 * p + j];
		run += diff[p];
		c[own] += run - (c[0] + c[1] + c[2] + c[3]); // aligned bases less mismatches: the matches
		if (decoded)
			for (int j = 0; j < 4; j++) decoded[4 * p + j] = c[j];
		c[own] += 1; // self vote (:149)
		const int outc = consensus_base(o, c, own);
		if (outc != own) changed++;
		out_codes[p] = (uint8_t) outc;
	}
	return changed;
}

} // namespace hero
