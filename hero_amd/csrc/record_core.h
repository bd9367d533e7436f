// Chaining and alignment-record assembly (src/SubjectReadAligner.cpp:197-408,711-852;
// src/AlignmentRecord.cpp:35-100,302-316), written once for the kernels and the CPU harness.
//
// A record's geometry (segment coordinates, band widening, fixed CIGAR pieces) depends only on
// its chain of seeds, never on DP results. walk_record() therefore drives a Visitor through the
// reference's exact sequence of actions; the same walk is run
//   - in plan mode (pass A): emit every DP segment the record could need (speculatively),
//   - in score mode (pass B): consume DP results, apply the early exits in order, accept/reject,
//   - in vote mode (pass B): replay the accepted winner and cast its votes.
#pragma once
#include "hd.h"

namespace hero {

// CIGAR op types inside records: M, I, D, or a mismatch typed by the query base (4 + code).
enum : int { kOpM = 0, kOpI = 1, kOpD = 2, kOpX = 4 };

// ---- chaining ----------------------------------------------------------------------------
// isFromTheSameAlignment (src/SubjectReadAligner.cpp:391-408): the key must lie strictly after
// the chain's last key on both reads, and the drift from the chain's first key must satisfy
// |dq - ds| / (max(dq, ds) + k) <= indelRate, evaluated in IEEE double like the reference.
HD bool same_alignment(const Opts &o, int add_pq, int add_ps, int beg_pq, int beg_ps, int end_pq, int end_ps) {
	if (add_ps <= end_ps) return false;
	if (add_pq <= end_pq) return false;
	double lq = (double) (add_pq - beg_pq);
	double ls = (double) (add_ps - beg_ps);
	lq = lq < 0 ? -lq : lq;
	ls = ls < 0 ? -ls : ls;
	double total = lq >= ls ? lq : ls;
	total += (double) o.k;
	double diff = lq - ls;
	diff = diff < 0 ? -diff : diff;
	if ((diff / total) > o.indel_rate) return false;
	return true;
}

// Greedy first-fit chaining of one query group (processOneSubjectRead2, :817-837) and the
// covered-length score of each chain (calKeyLength, :711-727). `seed(t)` returns the packed
// seed t of the group (sorted). Chain state lives in caller storage of group size:
//   cnext[t] the next seed of t's chain (-1: t is the chain's last)
//   cfirst[c], clast[c] first / last seed of chain c, clen[c] covered length.
// Returns the chain count; *best receives the maximum covered length (markTheBestPossible-
// Alignment, :729-751, keeps exactly the chains with that length, ties included).
template <class FSeed, class FI, class FL>
HD int chain_group(const Opts &o, int gn, const FSeed &seed, FI &cnext, FI &cfirst, FI &clast, FL &clen, int *best) {
	int nch = 0;
	for (int t = 0; t < gn; t++) {
		const uint64_t s = seed(t);
		const int pq = seed_pq(s), ps = seed_ps(s);
		int placed = -1;
		for (int c = 0; c < nch; c++) {
			const uint64_t b = seed(cfirst[c]), e = seed(clast[c]);
			if (same_alignment(o, pq, ps, seed_pq(b), seed_ps(b), seed_pq(e), seed_ps(e))) {
				placed = c;
				break;
			}
		}
		if (placed >= 0) {
			const int eps = seed_ps(seed(clast[placed]));
			clen[placed] += (eps + o.k <= ps) ? o.k : (ps - eps);
			cnext[clast[placed]] = t;
			cnext[t] = -1;
			clast[placed] = t;
		} else {
			cfirst[nch] = t;
			clast[nch] = t;
			clen[nch] = 0;
			cnext[t] = -1;
			nch++;
		}
	}
	int m = 0;
	for (int c = 0; c < nch; c++) {
		const int len = clen[c] + o.k;
		clen[c] = len;
		m = len > m ? len : m;
	}
	*best = m;
	return nch;
}

// Sequential access to the keys of one chain, following the chain links from its first seed.
template <class FSeed, class FI>
struct ChainKeys {
	const FSeed &seed;
	const FI &cnext;
	int t;
	int last_t;
	HD ChainKeys(const FSeed &s, const FI &nx, int first, int last) : seed(s), cnext(nx), t(first), last_t(last) {}
	HD bool next(int &pq, int &ps) {
		if (t < 0) return false;
		const uint64_t s = seed(t);
		t = cnext[t];
		pq = seed_pq(s), ps = seed_ps(s);
		return true;
	}
	HD void last(int &pq, int &ps) const {
		const uint64_t s = seed(last_t);
		pq = seed_pq(s), ps = seed_ps(s);
	}
};

// One DP call as getAlignment2 issues it: setSequences picks the longer side as the aligner's
// subject (the query side keeps the role on ties, src/BandAligner.cpp:559-573) and the call
// widens m_band to lenS - lenQ (:59-62,144-148).
struct DpCall {
	int s0, s1, q0, q1; // record frames
	int flag;           // 0 none, 1 setAlignLeft, 2 setAlignRight
	bool swapped;
	int band;           // m_band used by this call
	int lenS, lenQ;     // aligner frame
};

// ---- getAlignment2 ------------------------------------------------------------------------
// Returns true when the record is accepted (final isGoodAlignment()). The visitor sees, in the
// reference's order: start(pos...), add(count, op) for fixed ops, dp(call) for each DP segment,
// check(maxMis, maxIndel) after the left end and after every key (false = reject now), and
// finish() for the final rate test.
template <class Keys, class V>
HD bool walk_record(const Opts &o, int Ls, int Lq, Keys &keys, V &v) {
	const int k = o.k;
	int pq, ps;
	if (!keys.next(pq, ps)) return false;
	int64_t ov = pq < ps ? pq : ps;
	ov += (Lq - pq < Ls - ps) ? Lq - pq : Ls - ps;
	if (ov < o.min_overlap) return false; // :221-224
	const int64_t extra = (int64_t) round((double) ov * o.indel_rate);  // :225
	const int64_t max_mis = (int64_t) round((double) ov * o.mismatch_rate); // :226
	const int64_t max_indel = extra;
	int band = o.band_override > 0 ? o.band_override : (int) (extra * 2); // setBand (:241)
	auto call = [&](int s0, int s1, int q0, int q1, int flag) {
		DpCall c;
		c.s0 = s0, c.s1 = s1, c.q0 = q0, c.q1 = q1, c.flag = flag;
		c.swapped = !(q1 - q0 <= s1 - s0);
		c.lenS = c.swapped ? q1 - q0 : s1 - s0;
		c.lenQ = c.swapped ? s1 - s0 : q1 - q0;
		if (c.lenS - c.lenQ > band) band = c.lenS - c.lenQ;
		c.band = band;
		v.dp(c);
	};
	// left end (:244-281)
	if (pq != 0 && ps != 0) {
		if (pq < ps) {
			const int s0 = ps - pq - extra < 0 ? 0 : (int) (ps - pq - extra);
			v.start(s0);
			call(s0, ps, 0, pq, 2);
		} else if (pq == ps) {
			v.start(0);
			call(0, ps, 0, pq, 2);
		} else {
			const int q0 = pq - ps - extra < 0 ? 0 : (int) (pq - ps - extra);
			v.start(0);
			if (pq - ps - extra > 0) v.add((int) (pq - ps - extra), kOpI);
			call(0, ps, q0, pq, 2);
		}
		v.add(k, kOpM);
	} else {
		if (ps == 0 && pq == 0) v.start(0);
		else if (ps == 0) v.start(0), v.add(pq, kOpI);
		else v.start(ps);
		v.add(k, kOpM);
	}
	if (!v.check(max_mis, max_indel)) return false;
	// middle (:288-345)
	int cq, cs;
	while (keys.next(cq, cs)) {
		if (cq - pq == cs - ps && ps + k >= cs) {
			if (ps + k > cs) v.add(cs - ps, kOpM);
			else v.add(k, kOpM);
		} else if (ps + k == cs) {
			if (cq - pq - k > 0) {
				v.add(cq - pq - k, kOpI);
				v.add(k, kOpM);
			} else {
				v.add(-cq + pq + k, kOpD);
				v.add(cq - pq, kOpM);
			}
		} else if (ps + k > cs) {
			if (cq - pq < cs - ps) {
				v.add(cs - ps - cq + pq, kOpD);
				v.add(cq - pq, kOpM);
			} else {
				v.add(cq - pq + ps - cs, kOpI);
				v.add(cs - ps, kOpM);
			}
		} else {
			if (pq + k == cq) {
				v.add(cs - ps - k, kOpD);
				v.add(k, kOpM);
			} else if (pq + k > cq) {
				v.add(cs - ps + pq - cq, kOpD);
				v.add(k + pq - cq, kOpM);
			} else {
				call(ps + k, cs, pq + k, cq, 0);
				v.add(k, kOpM);
			}
		}
		pq = cq, ps = cs;
		if (!v.check(max_mis, max_indel)) return false;
	}
	// right end (:348-381)
	keys.last(pq, ps);
	pq += k, ps += k;
	if (pq < Lq) {
		if (ps == Ls) {
			v.add(Lq - pq, kOpI);
		} else if (Ls - ps < Lq - pq) {
			const int64_t qe = Ls - ps + pq + extra;
			call(ps, Ls, pq, qe < Lq ? (int) qe : Lq, 1);
			if (qe < Lq) v.add((int) (Lq - qe), kOpI);
		} else if (Ls - ps == Lq - pq) {
			call(ps, Ls, pq, Lq, 1);
		} else {
			const int64_t se = (int64_t) ps + (Lq - pq) + extra;
			call(ps, se < Ls ? (int) se : Ls, pq, Lq, 1);
		}
	}
	return v.finish();
}

// ---- AlignmentRecord scoring --------------------------------------------------------------
// Incremental form of addCigar merging + calculateScore (src/AlignmentRecord.cpp:35-100):
// mismatch ops count everywhere; I/D and length exclude a leading / trailing I/D op.
struct Score {
	int n;                 // merged ops so far
	int first_t, first_c;  // first merged op
	int last_t, last_c;    // last merged op
	int64_t total, sum_id, sum_mis;
	HD void reset() { n = 0, first_t = last_t = -1, first_c = last_c = 0, total = sum_id = sum_mis = 0; }
	HD void add(int count, int t) {
		if (n > 0 && last_t == t) {
			last_c += count;
			if (n == 1) first_c += count;
		} else {
			n++;
			last_t = t, last_c = count;
			if (n == 1) first_t = t, first_c = count;
		}
		// selects on values, not on which field: a branch per field lets the compiler merge the
		// two updates into one store at a computed offset, which moves the struct to scratch
		total += count;
		sum_id += (t == kOpI || t == kOpD) ? count : 0;
		sum_mis += (t >= kOpX) ? count : 0;
	}
	// One DP call's ops at once, from its run summary (steps: 0 match, 1 mismatch, 2 left, 3 up;
	// left / up are I / D, or D / I when the call's aligner subject is the record's query).
	// Equivalent to add(1, t) per step: the first run merges with the current last op, the
	// inner runs only add to the totals, the last run becomes the last op. Mismatches count as
	// one op type here: merging equal-letter mismatch ops never changes a score (they are never
	// I/D, and every mismatch op counts whole wherever it sits).
	HD static int piece_type(int code, bool swapped) {
		if (code < 2) return code == 0 ? kOpM : kOpX;
		return ((code == 2) != swapped) ? kOpI : kOpD;
	}
	HD void add_piece(const StepSummary &s, bool swapped) {
		const int steps = s.n[0] + s.n[1] + s.n[2] + s.n[3];
		if (steps == 0) return;
		const int ft = piece_type(s.first_code, swapped), lt = piece_type(s.last_code, swapped);
		if (s.first_len == steps) {
			add(steps, ft);
			return;
		}
		add(s.first_len, ft);
		int64_t mid_tot = steps - s.first_len - s.last_len, mid_id = 0, mid_mis = 0;
		for (int c = 0; c < 4; c++) {
			const int cnt = s.n[c] - (c == s.first_code ? s.first_len : 0) - (c == s.last_code ? s.last_len : 0);
			if (c >= 2) mid_id += cnt;
			else if (c == 1) mid_mis += cnt;
		}
		if (mid_tot > 0) {
			total += mid_tot, sum_id += mid_id, sum_mis += mid_mis;
			n++;
			last_t = -2; // inner runs: a type no run can have, so the last run starts a new op
			last_c = 0;
		}
		add(s.last_len, lt);
	}
	HD static bool is_id(int t) { return t == kOpI || t == kOpD; }
	// calculateScore: (nMismatch, nIndel, alignment length)
	HD void counts(double &nmis, double &nindel, double &len) const {
		int64_t ex = 0, ex_id = 0;
		if (n > 0 && is_id(first_t)) ex += first_c, ex_id += first_c;
		if (n > 1 && is_id(last_t)) ex += last_c, ex_id += last_c;
		nmis = (double) sum_mis;
		nindel = (double) (sum_id - ex_id);
		len = (double) (total - ex);
	}
	// isGoodAlignment(_numMisMatch, _numIndel) (:310-316)
	HD bool good_caps(int64_t cap_mis, int64_t cap_indel) const {
		double a, b, c;
		counts(a, b, c);
		return !(a > (double) cap_mis || b > (double) cap_indel);
	}
	// isGoodAlignment() (:302-308) -> rates for dedup
	HD bool good_rates(double mismatch_rate, double indel_rate, double &rate_sum, double &len) const {
		double nmis, nindel;
		counts(nmis, nindel, len);
		double irate = 100, mrate = 100;
		if (len != 0) irate = nindel / len, mrate = nmis / len;
		rate_sum = irate + mrate;
		return mrate <= mismatch_rate && irate <= indel_rate;
	}
};

// SubjectRead::addAlignment decision (src/SubjectRead.cpp:36-50): a later record replaces the
// kept one only when its rate sum is lower, or equal with a strictly longer alignment.
HD bool record_better(double new_sum, double new_len, double old_sum, double old_len) {
	if (old_sum < new_sum) return false;
	if (old_sum == new_sum) return !(old_len >= new_len);
	return true;
}

} // namespace hero
