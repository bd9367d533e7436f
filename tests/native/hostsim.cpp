// CPU harness for the device logic: runs hero_amd/csrc/{dp_core,record_core,pipeline_core}.h
// through the same dataflow as the kernels (pass A plan -> DP -> pass B score/dedup/votes ->
// consensus) with plain loops, so tests can compare that logic with the oracle without a GPU.
// Test infrastructure only; the library never calls it.
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <map>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "../../hero_amd/csrc/dp_pk16.h"
#include "../../hero_amd/csrc/pack.h"
#include "../../hero_amd/csrc/pipeline_core.h"

using namespace hero;

extern "C" {
typedef struct {
	int32_t k, aligner, call_consensus, hash_factor;
	int64_t min_overlap;
	double mismatch_rate, indel_rate, consensus_freq;
	int32_t band_override, max_candidates;
	int32_t dx, dy, ex, sx, miss, hit, unobserved;
} hs_cfg;
typedef struct {
	int64_t n;
	const char *seq;
	const int64_t *offset;
	const int32_t *length;
	const int64_t *unique_id;
} hs_reads;
}

static Opts to_opts(const hs_cfg *c) {
	Opts o;
	std::memset(&o, 0, sizeof o);
	o.k = c->k, o.aligner = c->aligner, o.call_consensus = c->call_consensus;
	o.band_override = c->band_override, o.max_candidates = c->max_candidates;
	o.min_overlap = c->min_overlap;
	o.mismatch_rate = c->mismatch_rate, o.indel_rate = c->indel_rate, o.consensus_freq = c->consensus_freq;
	o.dx = c->dx, o.dy = c->dy, o.ex = c->ex, o.sx = c->sx, o.miss = c->miss, o.hit = c->hit, o.unobserved = c->unobserved;
	return o;
}

struct PackedSet {
	std::vector<uint64_t> words, woff;
	std::vector<int32_t> len;
	std::vector<int64_t> uid;
	std::vector<uint8_t> good;
	Packed P;
	void build(const hs_reads *r, const std::vector<uint8_t> &g) {
		pack_reads(r->seq, r->offset, r->length, r->n, words, woff);
		len.assign(r->length, r->length + r->n);
		uid.assign(r->unique_id, r->unique_id + r->n);
		good = g;
		P.words = words.data(), P.woff = woff.data(), P.len = len.data(), P.uid = uid.data(), P.good = good.data();
		P.n = r->n;
	}
};

// A DP task side's codes located as the device reads them (segment_code: word, bit, frame).
struct SegStream {
	static constexpr bool kTwoBit = true;
	const uint64_t *words;
	uint32_t word;
	int bit, len, i;
	bool rc;
	int next() { return i < len ? segment_code(words, word, bit, i++, rc) : kCodeNone; }
};
struct RowStore {
	std::vector<uint32_t> v;
	int nw;
	void put_row(int i, const uint32_t *w) {
		if ((size_t) (i + 1) * nw > v.size()) v.resize((size_t) (i + 1) * nw);
		for (int t = 0; t < nw; t++) v[(size_t) i * nw + t] = w[t];
	}
	uint32_t get(int i, int w) const { return v[(size_t) i * nw + w]; }
};

template <int BMAX, bool AFF, class SS, class QS>
static int run_dp(const Opts &o, int lenS, int lenQ, int b, bool la, bool ra, SS &S, QS &Q, std::vector<uint32_t> &steps) {
	RowStore ps;
	ps.nw = (2 * BMAX + 1 + 15) / 16;
	dp_forward<BMAX, AFF>(o, lenS, lenQ, b, la, ra, S, Q, ps);
	steps.clear();
	auto em = [&](uint32_t s) { steps.push_back(s); };
	dp_traceback(lenS, lenQ, b, ps, em);
	std::reverse(steps.begin(), steps.end());
	return 0;
}

template <class SS, class QS>
static int dispatch_dp(const Opts &o, int lenS, int lenQ, int b, bool la, bool ra, SS &S, QS &Q, std::vector<uint32_t> &steps) {
	const bool aff = o.aligner == 0;
#define HS_CASE(BM)                                                                                   \
	if (b <= BM) return aff ? run_dp<BM, true>(o, lenS, lenQ, b, la, ra, S, Q, steps)                  \
	                        : run_dp<BM, false>(o, lenS, lenQ, b, la, ra, S, Q, steps);
	HS_CASE(8) HS_CASE(16) HS_CASE(32) HS_CASE(64) HS_CASE(128) HS_CASE(256)
#undef HS_CASE
	return -1;
}

extern "C" int hostsim_run(const hs_cfg *cfg, const hs_reads *qr, const hs_reads *sr, char *corrected, uint32_t *tally_out,
		uint64_t *stats) {
	const Opts o = to_opts(cfg);
	const int k = o.k;
	if (k < 1 || k > 64) return -1;
	std::vector<uint8_t> qg(qr->n), sg(sr->n);
	for (int64_t i = 0; i < qr->n; i++) qg[i] = query_good(qr->seq + qr->offset[i], qr->length[i], o.min_overlap);
	for (int64_t i = 0; i < sr->n; i++) sg[i] = subject_active(sr->seq + sr->offset[i], sr->length[i], o.min_overlap, k);
	PackedSet Q, S;
	Q.build(qr, qg);
	S.build(sr, sg);
	// index: (hi, lo) -> entries (innerId, pos) in insertion order
	std::map<std::pair<uint64_t, uint64_t>, std::vector<std::pair<int64_t, int>>> index;
	for (int64_t q = 0; q < qr->n; q++) {
		if (!qg[q]) continue;
		const int L = qr->length[q];
		for (int t = 0; t < num_keys(L, k); t++) {
			uint64_t hi, lo;
			const int p = key_pos(L, k, t);
			kmer_fwd(Q.P.words + Q.P.woff[q], p, k, hi, lo);
			index[{hi, lo}].push_back({q, p});
		}
	}
	// tally per query (4 u32 per base at the query's base offset)
	int64_t qbases = 0;
	for (int64_t q = 0; q < qr->n; q++) qbases = std::max<int64_t>(qbases, qr->offset[q] + qr->length[q]);
	std::vector<uint32_t> tally((size_t) 4 * qbases, 0), diff((size_t) qbases, 0);
	uint64_t st_seeds = 0, st_tasks = 0, st_cells_exec = 0, st_calls_exec = 0, st_built = 0, st_kept = 0, st_cells_all = 0;
	unsigned long long oob = 0;
	std::vector<uint32_t> steps;
	for (int64_t s = 0; s < sr->n; s++) {
		if (!sg[s]) continue;
		const int Ls = sr->length[s];
		struct StrandData {
			std::vector<uint64_t> seeds;
			std::vector<int> gstart; // group starts (+ end)
			std::vector<int32_t> cnext, cf, cl, clen;
			std::vector<Group> groups;
			std::vector<uint32_t> gtask0, gprog;
		} sd[2];
		std::vector<uint32_t> prog;
		std::vector<DpTask> tasks;
		std::vector<uint32_t> words_nsteps;
		uint32_t words_total = 0;
		for (int strand = 0; strand < 2; strand++) {
			StrandData &D = sd[strand];
			for (int p = 0; p + k <= Ls; p++) {
				uint64_t hi, lo;
				if (strand == 0) kmer_fwd(S.P.words + S.P.woff[s], p, k, hi, lo);
				else {
					uint64_t fh, fl;
					kmer_fwd(S.P.words + S.P.woff[s], Ls - p - k, k, fh, fl);
					kmer_revcomp(k, fh, fl, hi, lo);
				}
				auto it = index.find({hi, lo});
				if (it == index.end()) continue;
				for (auto &e : it->second) {
					if (Q.P.uid[e.first] == S.P.uid[s]) continue;
					int pq = e.second, ps = p;
					if (!o.call_consensus && strand == 1) ps = Ls - (ps + k), pq = qr->length[e.first] - (pq + k);
					D.seeds.push_back(seed_pack((uint64_t) e.first, pq, ps));
				}
			}
			std::sort(D.seeds.begin(), D.seeds.end());
			st_seeds += D.seeds.size();
			const int n = (int) D.seeds.size();
			D.cnext.assign(n, 0), D.cf.assign(n, 0), D.cl.assign(n, 0), D.clen.assign(n, 0);
			for (int t = 0; t < n; t++)
				if (t == 0 || seed_qid(D.seeds[t]) != seed_qid(D.seeds[t - 1])) D.gstart.push_back(t);
			if (o.max_candidates > 0 && (int) D.gstart.size() > o.max_candidates) D.gstart.resize(o.max_candidates);
			const int ng = (int) D.gstart.size();
			for (int gi = 0; gi < ng; gi++) {
				const int g0 = D.gstart[gi];
				int g1 = g0 + 1;
				while (g1 < n && seed_qid(D.seeds[g1]) == seed_qid(D.seeds[g0])) g1++;
				Group g;
				g.seeds = D.seeds.data() + g0;
				g.cnext = D.cnext.data() + g0, g.cfirst = D.cf.data() + g0, g.clast = D.cl.data() + g0, g.clen = D.clen.data() + g0;
				g.gn = g1 - g0;
				chain(o, g);
				const int64_t q = seed_qid(D.seeds[g0]);
				PairReads pr = make_pair_reads(S.P, Q.P, o.call_consensus, s, Ls, q, strand);
				int nt, nw, np, nsc;
				plan_count(o, g, pr, nt, nw, np, nsc);
				const uint32_t t0 = (uint32_t) tasks.size(), p0 = (uint32_t) prog.size();
				tasks.resize(tasks.size() + nt);
				prog.resize(prog.size() + np);
				plan_emit(o, g, pr, tasks.data(), t0, words_total, prog.data() + p0, (uint32_t) S.P.woff[s], (uint32_t) Q.P.woff[q]);
				words_total += (uint32_t) nw;
				D.groups.push_back(g);
				D.gtask0.push_back(t0);
				D.gprog.push_back(p0);
			}
		}
		// DP
		std::vector<uint32_t> words(words_total + 1, 0), nsteps(tasks.size(), 0);
		std::vector<StepSummary> sums(tasks.size());
		st_tasks += tasks.size();
		for (size_t t = 0; t < tasks.size(); t++) {
			const DpTask &T = tasks[t];
			const bool sw = T.flags & kTaskSwapped;
			SegStream subj{S.P.words, T.s_word, T.s_bit, T.s_len, 0, (T.flags & kTaskSubjRC) != 0};
			SegStream qry{Q.P.words, T.q_word, T.q_bit, T.q_len, 0, (T.flags & kTaskQryRC) != 0};
			SegStream &AS = sw ? qry : subj;
			SegStream &AQ = sw ? subj : qry;
			const int lenS = sw ? T.q_len : T.s_len, lenQ = sw ? T.s_len : T.q_len;
			st_cells_all += band_cells(lenS, lenQ, T.band);
			if (dispatch_dp(o, lenS, lenQ, T.band, T.flags & kTaskAlignLeft, T.flags & kTaskAlignRight, AS, AQ, steps) != 0)
				return -5;
			const int cap = T.s_len + T.q_len;
			const int ns = (int) steps.size();
			for (int e = 0; e < ns; e++) {
				const int pos = cap - ns + e;
				words[T.out_off + (pos >> 4)] |= steps[e] << (2 * (pos & 15));
			}
			nsteps[t] = (uint32_t) ns;
			SummaryBuilder sb;
			for (int e = ns - 1; e >= 0; e--) sb.add(steps[e]);
			sums[t] = sb.finish();
		}
		DpResults R{tasks.data(), words.data(), nsteps.data(), sums.data()};
		// pass B: pair groups of both strands by query
		std::map<int64_t, std::pair<int, int>> pairs;
		for (int strand = 0; strand < 2; strand++)
			for (size_t gi = 0; gi < sd[strand].groups.size(); gi++) {
				const int64_t q = seed_qid(sd[strand].groups[gi].seeds[0]);
				auto &pp = pairs.emplace(q, std::make_pair(-1, -1)).first->second;
				(strand == 0 ? pp.first : pp.second) = (int) gi;
			}
		for (auto &kv : pairs) {
			const int64_t q = kv.first;
			Best best;
			best.have = false;
			for (int strand = 0; strand < 2; strand++) {
				const int gi = strand == 0 ? kv.second.first : kv.second.second;
				if (gi < 0) continue;
				score_group(o, prog.data(), sd[strand].gprog[gi], R, sd[strand].gtask0[gi], strand, best, st_cells_exec,
						st_calls_exec, st_built);
			}
			if (!best.have) continue;
			st_kept++;
			if (!o.call_consensus) continue;
			PairReads pr = make_pair_reads(S.P, Q.P, o.call_consensus, s, Ls, q, best.strand);
			VoteSink sink{tally.data() + 4 * qr->offset[q], diff.data() + qr->offset[q], pr, &oob};
			vote_record(prog.data() + best.script, R, best.task0, sink);
		}
	}
	// consensus
	std::vector<uint32_t> decoded((size_t) 4 * qbases, 0);
	std::vector<uint8_t> codes;
	if (corrected) std::memcpy(corrected, qr->seq, (size_t) qbases);
	for (int64_t q = 0; q < qr->n; q++) {
		const int L = qr->length[q];
		codes.resize(L);
		consensus_read(o, tally.data() + 4 * qr->offset[q], diff.data() + qr->offset[q], L, Q.P, q, codes.data(),
				decoded.data() + 4 * qr->offset[q]);
		if (qg[q] && corrected)
			for (int p = 0; p < L; p++) corrected[qr->offset[q] + p] = code_ascii(codes[p]);
	}
	if (tally_out) std::memcpy(tally_out, decoded.data(), decoded.size() * 4);
	if (stats) {
		stats[0] = st_seeds, stats[1] = st_tasks, stats[2] = st_calls_exec, stats[3] = st_cells_exec;
		stats[4] = st_built, stats[5] = st_kept, stats[6] = st_cells_all, stats[7] = (uint64_t) oob;
	}
	return oob ? 1 : 0;
}

// One BandAligner call on ASCII strings (A/C/G/T/N), BandAligner CIGAR as text.
struct ByteStream {
	static constexpr bool kTwoBit = false; // 'N' is code 4
	const char *s;
	int len, i;
	int next() {
		if (i >= len) return kCodeNone;
		const char c = s[i++];
		return c == 'N' ? kCodeN : ascii_code(c);
	}
};
extern "C" int hostsim_band_align(const hs_cfg *cfg, const char *subj, int ls, const char *qry, int lq, int band, int flag,
		char *cigar_out, int cap, int *band_out) {
	const Opts o = to_opts(cfg);
	const bool sw = !(lq <= ls);
	const char *A = sw ? qry : subj, *B = sw ? subj : qry;
	const int lenS = sw ? lq : ls, lenQ = sw ? ls : lq;
	if (lenS - lenQ > band) band = lenS - lenQ;
	*band_out = band;
	ByteStream SS{A, lenS, 0}, QS{B, lenQ, 0};
	std::vector<uint32_t> steps;
	if (dispatch_dp(o, lenS, lenQ, band, flag == 1, flag == 2, SS, QS, steps) != 0) return -1;
	std::string out;
	int i = 0, j = 0;
	char cur = 0;
	int cnt = 0;
	auto push = [&](char t, bool mismatch) {
		if (cnt > 0 && t == cur && !mismatch) {
			cnt++;
			return;
		}
		if (cnt > 0) out += std::to_string(cnt) + cur;
		cur = t, cnt = 1;
	};
	for (uint32_t st : steps) {
		if (st == kStepLeft) push(sw ? 'D' : 'I', false), j++;
		else if (st == kStepUp) push(sw ? 'I' : 'D', false), i++;
		else {
			const char a = A[i], q = B[j];
			if (a == q) push('M', false);
			else push(sw ? a : q, true);
			i++, j++;
		}
	}
	if (cnt > 0) out += std::to_string(cnt) + cur;
	if ((int) out.size() + 1 > cap) return -2;
	std::memcpy(cigar_out, out.c_str(), out.size() + 1);
	return 0;
}

// The product's host-side read preparation: both predicates per read (read_flags, one scan) and
// the 2-bit packing (pack_reads, 2 zero words of tail padding), so tests can hold them against the
// oracle's predicates and a plain packing.
extern "C" void hostsim_read_flags(const hs_reads *r, int64_t min_overlap, int k, uint8_t *good, uint8_t *active) {
	for (int64_t i = 0; i < r->n; i++) read_flags(r->seq + r->offset[i], r->length[i], min_overlap, k, good[i], active[i]);
}
extern "C" int64_t hostsim_pack(const hs_reads *r, uint64_t *out, int64_t cap, uint64_t *woff_out) {
	std::vector<uint64_t> words, woff;
	pack_reads(r->seq, r->offset, r->length, r->n, words, woff);
	if ((int64_t) words.size() > cap) return -(int64_t) words.size();
	std::memcpy(out, words.data(), words.size() * 8);
	std::memcpy(woff_out, woff.data(), woff.size() * 8);
	return (int64_t) words.size();
}

extern "C" uint64_t hostsim_band_cells(int lenS, int lenQ, int b) { return band_cells(lenS, lenQ, b); }

// Packed two-task DP against the one-task DP on random task pairs: the tracebacks of both tasks
// must be equal. Returns the number of differing tasks (sequences with codes 0..3 as in the
// pipeline; lengths, bands and flags random; options from cfg).
struct CodeStream {
	static constexpr bool kTwoBit = true;
	const std::vector<int> *s;
	int i;
	int next() { return i < (int) s->size() ? (*s)[(size_t) i++] : kCodeNone; }
};
template <int BMAX, bool AFF>
static int pk16_check_one(const Opts &o, uint64_t &rng, int max_len) {
	auto rnd = [&](uint64_t n) {
		rng ^= rng << 13, rng ^= rng >> 7, rng ^= rng << 17;
		return (int) (rng % n);
	};
	std::vector<int> sa, qa, sb, qb;
	Pk16Task A, B;
	auto mk = [&](Pk16Task &T, std::vector<int> &S, std::vector<int> &Q) {
		T.lenS = rnd((uint64_t) max_len + 1);
		T.lenQ = T.lenS - rnd((uint64_t) (T.lenS < BMAX ? T.lenS + 1 : BMAX + 1));
		if (rnd(5) == 0) T.lenQ = rnd((uint64_t) T.lenS + 1);
		T.b = T.lenS - T.lenQ + rnd(3);
		if (T.b > BMAX) T.b = BMAX, T.lenQ = T.lenS - BMAX > 0 ? T.lenS - BMAX : 0;
		if (T.b < T.lenS - T.lenQ) T.lenQ = T.lenS - T.b;
		const int f = rnd(3);
		T.left_al = f == 1, T.right_al = f == 2;
		S.resize((size_t) T.lenS), Q.resize((size_t) T.lenQ);
		for (auto &c : S) c = rnd(4);
		for (size_t j = 0; j < Q.size(); j++) Q[j] = (j < S.size() && rnd(10)) ? S[j] : rnd(4);
	};
	mk(A, sa, qa);
	mk(B, sb, qb);
	if (!pk16_fits(o, A.lenS, A.lenQ) || !pk16_fits(o, B.lenS, B.lenQ)) return 0;
	const int nw = (2 * BMAX + 1 + 15) / 16;
	RowStore pa, pb, ra, rb;
	pa.nw = pb.nw = ra.nw = rb.nw = nw;
	CodeStream SA{&sa, 0}, QA{&qa, 0}, SB{&sb, 0}, QB{&qb, 0};
	dp_forward_pk16<BMAX, AFF>(o, A, B, SA, QA, SB, QB, pa, pb);
	CodeStream S1{&sa, 0}, Q1{&qa, 0}, S2{&sb, 0}, Q2{&qb, 0};
	dp_forward<BMAX, AFF>(o, A.lenS, A.lenQ, A.b, A.left_al, A.right_al, S1, Q1, ra);
	dp_forward<BMAX, AFF>(o, B.lenS, B.lenQ, B.b, B.left_al, B.right_al, S2, Q2, rb);
	// the tracebacks (what every consumer reads) must agree step for step; pointer bits of cells
	// outside the matrix are never read and may differ
	auto trace = [](const RowStore &r, const Pk16Task &T, std::vector<int> &steps) {
		steps.clear();
		auto emit = [&](int st) { steps.push_back(st); };
		return dp_traceback(T.lenS, T.lenQ, T.b, r, emit);
	};
	std::vector<int> x, y;
	int bad = 0;
	if (trace(pa, A, x) != trace(ra, A, y) || x != y) bad++;
	if (trace(pb, B, x) != trace(rb, B, y) || x != y) bad++;
	return bad;
}

// Tagged packed two-task DP against dp_forward on random task pairs: each task's traceback must
// agree. mode_or forces row variants on every row. Pairs outside the 16-bit range are not drawn
// (the caller's options and max_len decide); returns the number of differing tasks.
template <int BMAX>
static int pk16t_check_one(const Opts &o, uint64_t &rng, int max_len, int mode_or) {
	auto rnd = [&](uint64_t n) {
		rng ^= rng << 13, rng ^= rng >> 7, rng ^= rng << 17;
		return (int) (rng % n);
	};
	std::vector<int> sa, qa, sb, qb;
	Pk16Task A, B;
	auto mk = [&](Pk16Task &T, std::vector<int> &S, std::vector<int> &Q) {
		T.lenS = rnd((uint64_t) max_len + 1);
		T.lenQ = T.lenS - rnd((uint64_t) (T.lenS < BMAX ? T.lenS + 1 : BMAX + 1));
		if (rnd(5) == 0) T.lenQ = rnd((uint64_t) T.lenS + 1);
		T.b = T.lenS - T.lenQ + rnd(3);
		if (rnd(2)) T.b = BMAX;
		if (T.b > BMAX) T.b = BMAX, T.lenQ = T.lenS - BMAX > 0 ? T.lenS - BMAX : 0;
		if (T.b < T.lenS - T.lenQ) T.lenQ = T.lenS - T.b;
		const int f = rnd(3);
		T.left_al = f == 1, T.right_al = f == 2;
		S.resize((size_t) T.lenS), Q.resize((size_t) T.lenQ);
		for (auto &c : S) c = rnd(4);
		for (size_t j = 0; j < Q.size(); j++) Q[j] = (j < S.size() && rnd(10)) ? S[j] : rnd(4);
	};
	mk(A, sa, qa);
	mk(B, sb, qb);
	const int nw = (2 * BMAX + 1 + 15) / 16;
	RowStore pa, pb, ra, rb;
	pa.nw = pb.nw = ra.nw = rb.nw = nw;
	CodeStream SA{&sa, 0}, QA{&qa, 0}, SB{&sb, 0}, QB{&qb, 0};
	dp_forward_pk16_tagged<BMAX>(o, A, B, SA, QA, SB, QB, pa, pb, mode_or);
	CodeStream S1{&sa, 0}, Q1{&qa, 0}, S2{&sb, 0}, Q2{&qb, 0};
	dp_forward<BMAX, true>(o, A.lenS, A.lenQ, A.b, A.left_al, A.right_al, S1, Q1, ra);
	dp_forward<BMAX, true>(o, B.lenS, B.lenQ, B.b, B.left_al, B.right_al, S2, Q2, rb);
	std::vector<int> x, y;
	auto trace = [&](const RowStore &r, const Pk16Task &T, std::vector<int> &steps, auto map) {
		steps.clear();
		auto emit = [&](int st) { steps.push_back(st); };
		return dp_traceback<decltype(map)::value>(T.lenS, T.lenQ, T.b, r, emit);
	};
	using TG = std::integral_constant<uint32_t, kPtrTaggedAffine>;
	using PL = std::integral_constant<uint32_t, kPtrPlain>;
	int bad = 0;
	if (trace(pa, A, x, TG()) != trace(ra, A, y, PL()) || x != y) bad++;
	if (trace(pb, B, x, TG()) != trace(rb, B, y, PL()) || x != y) bad++;
	return bad;
}

extern "C" int hostsim_pk16t_check(const hs_cfg *cfg, int n, uint64_t seed, int bmax, int max_len, int mode_or) {
	const Opts o = to_opts(cfg);
	if (o.aligner != 0 || !pk16t_fits(o, max_len, bmax)) return -1;
	uint64_t rng = seed | 1;
	int bad = 0;
	for (int t = 0; t < n; t++) {
#define PT_CASE(BM) \
	if (bmax == BM) bad += pk16t_check_one<BM>(o, rng, max_len, mode_or);
		PT_CASE(4) PT_CASE(6) PT_CASE(8) PT_CASE(10) PT_CASE(12)
#undef PT_CASE
	}
	return bad;
}

// Tagged affine DP against dp_forward on random tasks (as pk16_check_one): the tracebacks must
// agree. mode_or forces row variants on every row (a lane among other tasks). Returns the number
// of differing tasks; -1 when the options do not admit the tagged form.
template <int BMAX, bool AFF>
static int tagged_check_one(const Opts &o, uint64_t &rng, int max_len, int mode_or) {
	auto rnd = [&](uint64_t n) {
		rng ^= rng << 13, rng ^= rng >> 7, rng ^= rng << 17;
		return (int) (rng % n);
	};
	Pk16Task T;
	std::vector<int> s, q;
	T.lenS = rnd((uint64_t) max_len + 1);
	T.lenQ = T.lenS - rnd((uint64_t) (T.lenS < BMAX ? T.lenS + 1 : BMAX + 1));
	if (rnd(5) == 0) T.lenQ = rnd((uint64_t) T.lenS + 1);
	T.b = T.lenS - T.lenQ + rnd(3);
	if (rnd(2)) T.b = BMAX; // the full band class: the head rows of the wavefront path
	if (T.b > BMAX) T.b = BMAX, T.lenQ = T.lenS - BMAX > 0 ? T.lenS - BMAX : 0;
	if (T.b < T.lenS - T.lenQ) T.lenQ = T.lenS - T.b;
	const int f = rnd(3);
	T.left_al = f == 1, T.right_al = f == 2;
	s.resize((size_t) T.lenS), q.resize((size_t) T.lenQ);
	for (auto &c : s) c = rnd(4);
	for (size_t j = 0; j < q.size(); j++) q[j] = (j < s.size() && rnd(10)) ? s[j] : rnd(4);
	const int nw = (2 * BMAX + 1 + 15) / 16;
	RowStore rt, rr;
	rt.nw = rr.nw = nw;
	CodeStream S1{&s, 0}, Q1{&q, 0}, S2{&s, 0}, Q2{&q, 0};
	if constexpr (AFF) dp_forward_tagged<BMAX>(o, T.lenS, T.lenQ, T.b, T.left_al, T.right_al, S1, Q1, rt, mode_or);
	else dp_forward_tagged_global<BMAX>(o, T.lenS, T.lenQ, T.b, T.left_al, T.right_al, S1, Q1, rt, mode_or);
	dp_forward<BMAX, AFF>(o, T.lenS, T.lenQ, T.b, T.left_al, T.right_al, S2, Q2, rr);
	std::vector<int> x, y;
	auto emx = [&](int st) { x.push_back(st); };
	auto emy = [&](int st) { y.push_back(st); };
	const int nx = dp_traceback<AFF ? kPtrTaggedAffine : kPtrTaggedGlobal>(T.lenS, T.lenQ, T.b, rt, emx);
	const int ny = dp_traceback(T.lenS, T.lenQ, T.b, rr, emy);
	return (nx != ny || x != y) ? 1 : 0;
}

extern "C" int hostsim_tagged_check(const hs_cfg *cfg, int n, uint64_t seed, int bmax, int max_len, int mode_or) {
	const Opts o = to_opts(cfg);
	const bool aff = o.aligner == 0;
	if ((aff && o.dx != o.dy) || !dp_tagged_fits(o, max_len + 1)) return -1;
	uint64_t rng = seed | 1;
	int bad = 0;
	for (int t = 0; t < n; t++) {
#define TG_CASE(BM) \
	if (bmax == BM) bad += aff ? tagged_check_one<BM, true>(o, rng, max_len, mode_or) : tagged_check_one<BM, false>(o, rng, max_len, mode_or);
		TG_CASE(2) TG_CASE(4) TG_CASE(6) TG_CASE(8) TG_CASE(10) TG_CASE(12)
#undef TG_CASE
	}
	return bad;
}

// The split pointer store (near rows with side words, far rows beyond) against RowStore: LANES
// lanes each run a random task's forward pass into one shared near and far area, then every lane
// traces back; each traceback must equal the one over RowStore. near_words: near-store words per
// lane (so near_rows = rows_for(near_words), from 0 rows up to every row). Returns the number of
// differing tasks.
template <int BMAX, bool AFF>
static int split_check_one(const Opts &o, uint64_t &rng, int max_len, int near_words) {
	constexpr int LANES = 5;
	constexpr int NW = (2 * BMAX + 1 + 15) / 16;
	constexpr bool SIDE = NW > 1 && (2 * BMAX + 1) % 16 == 1;
	using SP = SplitRowPtrs<NW, SIDE, LANES>;
	auto rnd = [&](uint64_t n) {
		rng ^= rng << 13, rng ^= rng >> 7, rng ^= rng << 17;
		return (int) (rng % n);
	};
	// the kernels' forms: 1 tagged affine rows (equal gap opens), 2 tagged global rows, 0 the general rows
	const int form = BMAX > 12 ? 0 : AFF ? (o.dx == o.dy ? 1 : 0) : 2;
	const int near_rows = SP::rows_for(near_words);
	std::vector<uint32_t> near((size_t) (near_words > 0 ? near_words : 1) * LANES, 0xA5A5A5A5u);
	std::vector<uint32_t> far((size_t) (max_len + 1) * NW * LANES, 0x5A5A5A5Au);
	Pk16Task T[LANES];
	std::vector<int> s[LANES], q[LANES];
	RowStore ref[LANES];
	SP sp[LANES];
	for (int l = 0; l < LANES; l++) {
		Pk16Task &t = T[l];
		t.lenS = rnd((uint64_t) max_len + 1);
		t.lenQ = t.lenS - rnd((uint64_t) (t.lenS < BMAX ? t.lenS + 1 : BMAX + 1));
		if (rnd(5) == 0) t.lenQ = rnd((uint64_t) t.lenS + 1);
		t.b = t.lenS - t.lenQ + rnd(3);
		if (rnd(2)) t.b = BMAX;
		if (t.b > BMAX) t.b = BMAX, t.lenQ = t.lenS - BMAX > 0 ? t.lenS - BMAX : 0;
		if (t.b < t.lenS - t.lenQ) t.lenQ = t.lenS - t.b;
		const int f = rnd(3);
		t.left_al = f == 1, t.right_al = f == 2;
		s[l].resize((size_t) t.lenS), q[l].resize((size_t) t.lenQ);
		for (auto &c : s[l]) c = rnd(4);
		for (size_t j = 0; j < q[l].size(); j++) q[l][j] = (j < s[l].size() && rnd(10)) ? s[l][j] : rnd(4);
		ref[l].nw = NW;
		sp[l] = SP{near.data(), far.data(), l, near_rows, 0u};
		CodeStream S1{&s[l], 0}, Q1{&q[l], 0}, S2{&s[l], 0}, Q2{&q[l], 0};
		if constexpr (BMAX <= 12) {
			if (form == 1) {
				dp_forward_tagged<BMAX>(o, t.lenS, t.lenQ, t.b, t.left_al, t.right_al, S1, Q1, sp[l]);
				dp_forward_tagged<BMAX>(o, t.lenS, t.lenQ, t.b, t.left_al, t.right_al, S2, Q2, ref[l]);
			} else if (form == 2) {
				dp_forward_tagged_global<BMAX>(o, t.lenS, t.lenQ, t.b, t.left_al, t.right_al, S1, Q1, sp[l]);
				dp_forward_tagged_global<BMAX>(o, t.lenS, t.lenQ, t.b, t.left_al, t.right_al, S2, Q2, ref[l]);
			}
		}
		if (form == 0) {
			dp_forward<BMAX, AFF>(o, t.lenS, t.lenQ, t.b, t.left_al, t.right_al, S1, Q1, sp[l]);
			dp_forward<BMAX, AFF>(o, t.lenS, t.lenQ, t.b, t.left_al, t.right_al, S2, Q2, ref[l]);
		}
	}
	int bad = 0;
	for (int l = 0; l < LANES; l++) {
		std::vector<int> x, y;
		auto emx = [&](int st) { x.push_back(st); };
		auto emy = [&](int st) { y.push_back(st); };
		auto trace = [&](const auto &ps, auto &em) {
			const Pk16Task &t = T[l];
			return form == 1 ? dp_traceback<kPtrTaggedAffine>(t.lenS, t.lenQ, t.b, ps, em)
			     : form == 2 ? dp_traceback<kPtrTaggedGlobal>(t.lenS, t.lenQ, t.b, ps, em)
			                 : dp_traceback(t.lenS, t.lenQ, t.b, ps, em);
		};
		const int nx = trace(sp[l], emx), ny = trace(ref[l], emy);
		if (nx != ny || x != y || nx < 0) bad++;
	}
	return bad;
}

extern "C" int hostsim_split_ptrs_check(const hs_cfg *cfg, int n, uint64_t seed, int bmax, int max_len, int near_words) {
	const Opts o = to_opts(cfg);
	uint64_t rng = seed | 1;
	int bad = 0;
	for (int t = 0; t < n; t++) {
#define SP_CASE(BM) \
	if (bmax == BM) bad += o.aligner == 0 ? split_check_one<BM, true>(o, rng, max_len, near_words) : split_check_one<BM, false>(o, rng, max_len, near_words);
		SP_CASE(4) SP_CASE(6) SP_CASE(8) SP_CASE(10) SP_CASE(12) SP_CASE(16) SP_CASE(24) SP_CASE(32)
#undef SP_CASE
	}
	return bad;
}

// The split pointer store as storage: LANES lanes put rows 0..len of random words (a last word of
// one cell holds a 2-bit code, as the forward passes leave it), all into one near and far area,
// then every (lane, row, word) is read back. Returns the number of words read back wrong.
template <int BMAX>
static int split_roundtrip_one(uint64_t &rng, int max_len, int near_words) {
	constexpr int LANES = 7;
	constexpr int NW = (2 * BMAX + 1 + 15) / 16;
	constexpr bool SIDE = NW > 1 && (2 * BMAX + 1) % 16 == 1;
	using SP = SplitRowPtrs<NW, SIDE, LANES>;
	auto rnd = [&](uint64_t n) {
		rng ^= rng << 13, rng ^= rng >> 7, rng ^= rng << 17;
		return (uint32_t) (rng % n);
	};
	const int near_rows = SP::rows_for(near_words);
	// the near area with a canary after it: no put may reach past near_words words per lane
	const size_t near_n = (size_t) (near_words > 0 ? near_words : 0) * LANES, canary = 4 * LANES;
	std::vector<uint32_t> near(near_n + canary, 0xA5A5A5A5u);
	std::vector<uint32_t> far((size_t) (max_len + 1) * NW * LANES, 0x5A5A5A5Au);
	std::vector<uint32_t> put[LANES];
	int len[LANES];
	for (int l = 0; l < LANES; l++) {
		len[l] = (int) rnd((uint64_t) max_len + 1);
		SP sp{near.data(), far.data(), l, near_rows, 0u};
		for (int i = 0; i <= len[l]; i++) {
			uint32_t pw[NW];
			for (int w = 0; w < NW; w++) pw[w] = (SIDE && w == NW - 1) ? rnd(4) : (uint32_t) (rng = rng * 6364136223846793005ull + 1442695040888963407ull) ^ (uint32_t) (rng >> 32);
			sp.put_row(i, pw);
			put[l].insert(put[l].end(), pw, pw + NW);
		}
	}
	int bad = 0;
	for (size_t k = near_n; k < near.size(); k++) bad += near[k] != 0xA5A5A5A5u;
	for (int l = 0; l < LANES; l++) {
		const SP sp{near.data(), far.data(), l, near_rows, 0u};
		for (int i = 0; i <= len[l]; i++)
			for (int w = 0; w < NW; w++) {
				const uint32_t want = put[l][(size_t) i * NW + w], got = sp.get(i, w);
				if (SIDE && w == NW - 1 ? (got & 3u) != want : got != want) bad++;
			}
	}
	return bad;
}

// rows_for(words) is the most rows whose near words fit: r rows fit, r + 1 do not
template <int BMAX>
static int split_rows_for_check(int max_words) {
	constexpr int NW = (2 * BMAX + 1 + 15) / 16;
	constexpr bool SIDE = NW > 1 && (2 * BMAX + 1) % 16 == 1;
	using SP = SplitRowPtrs<NW, SIDE, 1>;
	auto need = [](int r) { return r * SP::kNear + (SIDE ? (r + 15) / 16 : 0); };
	int bad = 0;
	for (int words = 0; words <= max_words; words++) {
		const int r = SP::rows_for(words);
		bad += r < 0 || need(r) > words || need(r + 1) <= words;
	}
	return bad;
}

extern "C" int hostsim_split_ptrs_roundtrip(int n, uint64_t seed, int bmax, int max_len, int near_words) {
	uint64_t rng = seed | 1;
	int bad = 0;
#define SF_CASE(BM) \
	if (bmax == BM) bad += split_rows_for_check<BM>(4096);
	SF_CASE(4) SF_CASE(6) SF_CASE(8) SF_CASE(10) SF_CASE(12) SF_CASE(16) SF_CASE(24) SF_CASE(32)
#undef SF_CASE
	for (int t = 0; t < n; t++) {
#define SR_CASE(BM) \
	if (bmax == BM) bad += split_roundtrip_one<BM>(rng, max_len, near_words);
		SR_CASE(4) SR_CASE(6) SR_CASE(8) SR_CASE(10) SR_CASE(12) SR_CASE(16) SR_CASE(24) SR_CASE(32)
#undef SR_CASE
	}
	return bad;
}

extern "C" int hostsim_pk16_check(const hs_cfg *cfg, int n, uint64_t seed, int bmax, int max_len) {
	const Opts o = to_opts(cfg);
	uint64_t rng = seed | 1;
	int bad = 0;
	for (int t = 0; t < n; t++) {
#define PK_CASE(BM)                                                                                          \
	if (bmax == BM) bad += o.aligner == 0 ? pk16_check_one<BM, true>(o, rng, max_len) : pk16_check_one<BM, false>(o, rng, max_len);
		PK_CASE(4) PK_CASE(8) PK_CASE(12)
#undef PK_CASE
	}
	return bad;
}
