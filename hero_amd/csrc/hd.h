// Host/device building blocks shared by the HIP kernels and the CPU harness (hostsim.cpp).
// Nothing here uses wave intrinsics; every function is plain per-lane code.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define HD __host__ __device__ __forceinline__
#else
#include <cmath>
#define HD inline
#endif

namespace hero {

// Run options as the kernels see them (mirrors Config fields of the reference,
// src/Config.cpp:17-56, and BandAligner::setCost, src/BandAligner.cpp:528-542).
struct Opts {
	int32_t k;
	int32_t aligner;        // 0 affine, 1 global
	int32_t call_consensus; // 1: vote and correct (-c); 0: alignment records only
	int32_t band_override;  // >0 replaces setBand(2*extra)
	int32_t max_candidates; // >0 keeps the first N query groups per subject strand
	int32_t mapping;        // --mapping: per subject only the best record covering it votes
	int64_t min_overlap;
	double mismatch_rate, indel_rate, consensus_freq;
	int32_t dx, dy, ex, sx, miss, hit, unobserved, _pad2;
	int32_t long_format; // -v (Config::bSpeicialVersion): no self-hit skip (src/SubjectReadAligner.cpp:464-465)
	int32_t merge;       // --merge: subject bases past a record's end vote into the query's extension tally
};

// 2-bit codes of the reference: (c >> 1) & 3 → A0 C1 T2 G3 (src/QueryRead.cpp:124-125).
// Complement A<->T, C<->G is code ^ 2.
HD int ascii_code(char c) { return ((unsigned char) c >> 1) & 3; }
HD char code_ascii(int c) { return "ACTG"[c & 3]; }

// Packed read set: 2-bit codes, MSB-first inside u64 words; read r starts at word woff[r].
// The word array carries two zero words of tail padding so k-mer extraction may read past a
// read's last word.
struct Packed {
	const uint64_t *words;
	const uint64_t *woff;
	const int32_t *len;
	const int64_t *uid;
	const uint8_t *good; // subject/query predicate of the dataset that owns the set
	int64_t n;
};

HD int base_at(const uint64_t *w, int64_t p) { return (int) ((w[p >> 5] >> (62 - 2 * (p & 31))) & 3ull); }

// Base p of read r in strand frame: forward, or of the reverse complement when rc.
HD int strand_base(const Packed &P, int64_t r, int p, bool rc) {
	const uint64_t *w = P.words + P.woff[r];
	if (!rc) return base_at(w, p);
	return base_at(w, P.len[r] - 1 - p) ^ 2;
}

// The k-mer at p (k <= 64) of a packed read, MSB-first: hi = first min(k,32) bases,
// lo = remaining k-32 bases (0 when k <= 32).
HD void kmer_fwd(const uint64_t *w, int p, int k, uint64_t &hi, uint64_t &lo) {
	int wi = p >> 5, sh = 2 * (p & 31);
	uint64_t a = w[wi], b = w[wi + 1], c = w[wi + 2];
	uint64_t x = sh ? ((a << sh) | (b >> (64 - sh))) : a;
	uint64_t y = sh ? ((b << sh) | (c >> (64 - sh))) : b;
	if (k <= 32) {
		hi = k == 32 ? x : (x >> (64 - 2 * k));
		lo = 0;
	} else {
		hi = x;
		lo = (k == 64) ? y : (y >> (64 - 2 * (k - 32)));
	}
}

// Reverse complement of a k-mer given as (hi, lo) — returns the same (hi, lo) convention.
HD uint64_t rev2(uint64_t x) { // reverse the 32 2-bit groups of a word
	x = ((x >> 2) & 0x3333333333333333ull) | ((x & 0x3333333333333333ull) << 2);
	x = ((x >> 4) & 0x0F0F0F0F0F0F0F0Full) | ((x & 0x0F0F0F0F0F0F0F0Full) << 4);
	x = ((x >> 8) & 0x00FF00FF00FF00FFull) | ((x & 0x00FF00FF00FF00FFull) << 8);
	x = ((x >> 16) & 0x0000FFFF0000FFFFull) | ((x & 0x0000FFFF0000FFFFull) << 16);
	x = (x >> 32) | (x << 32);
	return x;
}
HD void kmer_revcomp(int k, uint64_t hi, uint64_t lo, uint64_t &rhi, uint64_t &rlo) {
	// Build the full 2k-bit sequence left-aligned in a 128-bit pair, reverse, complement.
	uint64_t L0, L1; // left-aligned: L0 holds bases 0..31, L1 bases 32..63
	if (k <= 32) {
		L0 = k == 32 ? hi : (hi << (64 - 2 * k));
		L1 = 0;
	} else {
		L0 = hi;
		L1 = k == 64 ? lo : (lo << (64 - 2 * (k - 32)));
	}
	// reversed sequence right-aligned over 128 bits: R = rev(L1):rev(L0)
	uint64_t R0 = rev2(L1), R1 = rev2(L0); // R0 = high word, R1 = low word of the reversed 128 bits
	// The k bases of the RC sit in the low 2k bits of (R0:R1).
	const uint64_t A = 0xAAAAAAAAAAAAAAAAull; // complement = xor 2 per base
	if (k <= 32) {
		rhi = (k == 32 ? R1 : (R1 & ((1ull << (2 * k)) - 1))) ^ (k == 32 ? A : (A & ((1ull << (2 * k)) - 1)));
		rlo = 0;
	} else {
		int r = k - 32; // bases in lo
		// first 32 bases of the RC = bits [2k-1 .. 2r] of R, last r bases = bits [2r-1 .. 0]
		uint64_t first32 = r == 32 ? R0 : ((R0 << (64 - 2 * r)) | (R1 >> (2 * r)));
		uint64_t last = r == 32 ? R1 : (R1 & ((1ull << (2 * r)) - 1));
		rhi = first32 ^ A;
		rlo = last ^ (r == 32 ? A : (A & ((1ull << (2 * r)) - 1)));
	}
}

// HashTableLongKey::hashFunction (src/HashTableLongKey.cpp:188-206) on a packed k-mer:
// sum1 = leading 1 followed by the first 32 bases (the 1 falls off once 32 bases are in),
// sum2 = leading 1 followed by the remaining bases; ((sum1 % P) * (sum2 % P)) % P in u64.
HD uint64_t hash_kmer(int k, uint64_t hi, uint64_t lo, uint64_t P) {
	uint64_t s1 = k >= 32 ? hi : ((1ull << (2 * k)) | hi);
	uint64_t s2 = 1;
	if (k > 32) s2 = (k - 32 >= 32) ? lo : ((1ull << (2 * (k - 32))) | lo);
	return ((s1 % P) * (s2 % P)) % P;
}

HD uint64_t umulhi64(uint64_t a, uint64_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
	return __umul64hi(a, b);
#else
	return (uint64_t) (((unsigned __int128) a * b) >> 64);
#endif
}
// x mod P without a 64-bit division: with m = floor((2^64-1)/P) the quotient estimate
// floor(x*m / 2^64) is at most 2 below floor(x/P) for every 64-bit x, so the remainder
// needs at most two corrections.
struct FastMod {
	uint64_t P, m;
	HD uint64_t mod(uint64_t x) const {
		uint64_t r = x - umulhi64(x, m) * P;
		if (r >= P) r -= P;
		if (r >= P) r -= P;
		return r;
	}
};
HD FastMod make_fastmod(uint64_t P) { return FastMod{P, ~0ull / P}; }
// hash_kmer with FastMod (identical values).
HD uint64_t hash_kmer(int k, uint64_t hi, uint64_t lo, const FastMod &fm) {
	uint64_t s1 = k >= 32 ? hi : ((1ull << (2 * k)) | hi);
	uint64_t s2 = 1;
	if (k > 32) s2 = (k - 32 >= 32) ? lo : ((1ull << (2 * (k - 32))) | lo);
	return fm.mod(fm.mod(s1) * fm.mod(s2));
}

// Key positions of a query read (src/HashTableLongKey.cpp:279-289): 0, k, 2k, ... <= L-k and,
// when that stride leaves the tail uncovered, L-k.
HD int num_keys(int L, int k) {
	if (L < k) return 0;
	int n = (L - k) / k + 1;
	if ((L - k) % k != 0) n++;
	return n;
}
HD int key_pos(int L, int k, int t) {
	int n = (L - k) / k + 1;
	return t < n ? t * k : L - k;
}

// Seed encoding: (innerId << 28) | (posQ << 14) | posS; ascending order is the reference's
// (innerId, posQ, posS) order (src/SubjectRead.cpp:157-173). Reads are limited to 16383 bases.
constexpr int kSeedPosBits = 14;
constexpr int kMaxReadLen = (1 << kSeedPosBits) - 1;
HD uint64_t seed_pack(uint64_t qid, int pq, int ps) {
	return (qid << (2 * kSeedPosBits)) | ((uint64_t) pq << kSeedPosBits) | (uint64_t) ps;
}
HD int64_t seed_qid(uint64_t s) { return (int64_t) (s >> (2 * kSeedPosBits)); }
HD int seed_pq(uint64_t s) { return (int) ((s >> kSeedPosBits) & kMaxReadLen); }
HD int seed_ps(uint64_t s) { return (int) (s & kMaxReadLen); }

// DP task: one BandAligner call after setSequences (src/BandAligner.cpp:559-573) has chosen the
// aligner's subject (the longer side; the query side on ties) and setBand/widening has fixed the
// band. Sides are described in their own frames; `rc` sides read the reverse complement.
enum : uint8_t {
	kTaskAlignLeft = 1,  // setAlignLeft
	kTaskAlignRight = 2, // setAlignRight
	kTaskSwapped = 4,    // aligner subject is the record's query side
	kTaskSubjRC = 8,     // record subject side is a reverse complement (reverse strand under -c)
	kTaskQryRC = 16,     // record query side is a reverse complement (reverse strand without -c)
};
// Each side's segment is located by its first code (in the side's own frame) as the packed word
// that holds it, counted from the start of its read set's words, and its place in that word (code
// 0 at the word's top bits). An `rc` side reads on towards lower forward positions.
struct DpTask {
	uint32_t s_word, q_word; // word of the record subject / query segment's first code
	uint16_t s_bit, s_len;   // its place in the word (0..31); the segment's length
	uint16_t q_bit, q_len;
	uint16_t band;           // effective m_band (already widened)
	uint8_t flags;
	uint8_t _r;
	uint32_t out_off;        // edit-step words at out_off; capacity ceil((s_len+q_len)/16) words
	uint32_t _r2;
	uint32_t _r3;
};
static_assert(sizeof(DpTask) == 32, "DpTask layout");

// The first code of the segment [p0, p0 + len) of a read whose words start at word woff (L bases;
// rc: the segment of its reverse complement, whose first code is forward position L - 1 - p0):
// its word (from the start of the set's words) and its place in that word.
HD void segment_start(uint64_t woff, int L, int p0, int len, bool rc, uint32_t &word, uint16_t &bit) {
	const int f = len <= 0 ? 0 : rc ? L - 1 - p0 : p0;
	word = (uint32_t) (woff + (uint64_t) (f >> 5));
	bit = (uint16_t) (f & 31);
}
// Code p of a task side (0 <= p < len) from the set's words, in the side's frame.
HD int segment_code(const uint64_t *words, uint32_t word, int bit, int p, bool rc) {
	const int64_t f0 = (int64_t) word * 32 + bit;
	return rc ? (base_at(words, f0 - p) ^ 2) : base_at(words, f0 + p);
}

HD int edit_words(int lenS, int lenQ) { return (lenS + lenQ + 15) / 16; }

// Edit step codes written by the DP traceback (aligner orientation).
enum : uint32_t { kStepMatch = 0, kStepMismatch = 1, kStepLeft = 2, kStepUp = 3 };

// Run summary of one DP call's steps in forward order: the count of each step code and the
// code / length of the first and last maximal runs. Record scoring needs nothing else (a record's
// score depends on op totals and on its first / last op when those are I or D).
struct StepSummary {
	uint16_t n[4];
	uint16_t first_len, last_len;
	uint8_t first_code, last_code;
	uint16_t _r;
};
static_assert(sizeof(StepSummary) == 16, "StepSummary layout");

// Builds a StepSummary from steps arriving last-to-first (the traceback's order).
struct SummaryBuilder {
	uint16_t n[4] = {0, 0, 0, 0};
	int cur = -1, cur_len = 0;
	int last_code = -1, last_len = 0; // the run emitted first = last in forward order
	HD void add(uint32_t code) {
		n[code]++;
		if ((int) code == cur) {
			cur_len++;
			return;
		}
		if (cur >= 0 && last_code < 0) last_code = cur, last_len = cur_len;
		cur = (int) code, cur_len = 1;
	}
	HD StepSummary finish() const {
		StepSummary s;
		for (int i = 0; i < 4; i++) s.n[i] = n[i];
		s.first_code = (uint8_t) (cur < 0 ? 0 : cur);
		s.first_len = (uint16_t) cur_len;
		s.last_code = (uint8_t) (last_code < 0 ? s.first_code : last_code);
		s.last_len = (uint16_t) (last_code < 0 ? cur_len : last_len);
		s._r = 0;
		return s;
	}
};

// Writes traceback steps last-to-first into a task's word range: step m goes to slot cap-1-m.
struct StepWriter {
	uint32_t *w;
	int e;
	int cur;
	uint32_t acc;
	SummaryBuilder sb;
	HD void operator()(uint32_t st) {
		sb.add(st);
		e--;
		const int wi = e >> 4;
		if (wi != cur) {
			if (cur >= 0) w[cur] = acc;
			acc = 0;
			cur = wi;
		}
		acc |= st << (2 * (e & 15));
	}
	HD void flush() {
		if (cur >= 0) w[cur] = acc;
	}
};

// In-band cells of one DP call: sum over rows of the row width, minus the origin.
// Cells of one banded DP call: sum over rows i = 0..lenS of (min(lenQ, i+b) - max(0, i-b) + 1),
// minus the origin cell. Closed form of that sum (each part summed separately, so the identity holds
// for any arguments).
HD uint64_t band_cells(int lenS, int lenQ, int b) {
	const int64_t S = lenS, Lq = lenQ, B = b;
	// sum of min(lenQ, i + b): i + b for i <= lenQ - b, lenQ after
	const int64_t A = (Lq - B < S) ? Lq - B : S;
	int64_t hi_sum;
	if (A < 0) hi_sum = (S + 1) * Lq;
	else hi_sum = (A + 1) * B + A * (A + 1) / 2 + (S - A) * Lq;
	// sum of max(0, i - b): i - b from i = max(0, b + 1)
	const int64_t j0 = B + 1 > 0 ? B + 1 : 0;
	const int64_t lo_sum = j0 > S ? 0 : (S - j0 + 1) * ((j0 - B) + (S - B)) / 2;
	return (uint64_t) (hi_sum - lo_sum + (S + 1) - 1);
}

} // namespace hero
