// Host-side read preparation shared by the C-ABI library and the CPU harness: the reference's
// dataset predicates and 2-bit packing into the device layout.
#pragma once
#include <stdint.h>

#include <cstring>
#include <vector>

#include "hd.h"

namespace hero {

// QueryDataset::qualityFilter (src/QueryDataset.cpp:60-75): only A/C/G/T and no base at >= 80%.
// Branch-free per block of 255 bytes (8-bit counters the compiler vectorizes): a block whose four
// letter counts do not add up to its length holds another byte, which fails the read.
inline bool quality_filter(const char *s, int64_t n) {
	const unsigned char *u = (const unsigned char *) s;
	uint64_t na = 0, nc = 0, ng = 0, nt = 0;
	for (int64_t i0 = 0; i0 < n; i0 += 255) {
		const int m = (int) (n - i0 < 255 ? n - i0 : 255);
		uint8_t a = 0, c = 0, g = 0, t = 0;
		for (int j = 0; j < m; j++) {
			const unsigned char b = u[i0 + j];
			a += b == 'A', c += b == 'C', g += b == 'G', t += b == 'T';
		}
		if (a + c + g + t != m) return false;
		na += a, nc += c, ng += g, nt += t;
	}
	const uint64_t thr = (uint64_t) ((double) n * .8);
	return !(na >= thr || nc >= thr || ng >= thr || nt >= thr);
}
// Query reads are indexed and corrected iff len > minOverlap && qualityFilter (src/QueryDataset.cpp:226).
inline bool query_good(const char *s, int64_t n, int64_t min_overlap) {
	return (uint64_t) n > (uint16_t) min_overlap && quality_filter(s, n);
}
// Subject reads are streamed iff len < minOverlap || qualityFilter (src/SubjectDataset.cpp:157-162);
// those shorter than minOverlap can never pass getAlignment2's overlap test (src/SubjectReadAligner
// .cpp:216-224), so only the others need any device work.
inline bool subject_active(const char *s, int64_t n, int64_t min_overlap, int k) {
	const bool good = (uint64_t) n < (uint16_t) min_overlap || quality_filter(s, n);
	return good && n >= min_overlap && n >= k;
}
// Both predicates of one read (a query set is its own subject set) from one scan: each needs the
// quality filter only when len >= (u16) minOverlap.
inline void read_flags(const char *s, int64_t n, int64_t min_overlap, int k, uint8_t &good, uint8_t &active) {
	const uint64_t mo = (uint16_t) min_overlap;
	const bool qf = (uint64_t) n >= mo && quality_filter(s, n);
	good = (uint64_t) n > mo && qf;
	active = ((uint64_t) n < mo || qf) && n >= min_overlap && n >= k;
}

static_assert(__BYTE_ORDER__ == __ORDER_LITTLE_ENDIAN__, "pack8 loads bases little-endian");
// 8 ASCII bases -> 16 bits of 2-bit codes, the first base in the top two bits: after the byte
// swap the first base sits in the top byte, and three shift-or steps close the gaps pairwise.
inline uint64_t pack8(const char *p) {
	uint64_t x;
	std::memcpy(&x, p, 8);
	x = (__builtin_bswap64(x) >> 1) & 0x0303030303030303ull; // ascii_code of every byte
	x = (x | (x >> 6)) & 0x000F000F000F000Full;
	x = (x | (x >> 12)) & 0x000000FF000000FFull;
	return (x | (x >> 24)) & 0xFFFFull;
}
// One read into 2-bit MSB-first words: whole 8-base groups by pack8, the last < 8 bases one by one
// (no load crosses the read's end).
inline void pack_read(const char *s, int L, uint64_t *w) {
	int i = 0;
	for (; i + 32 <= L; i += 32)
		w[i >> 5] = pack8(s + i) << 48 | pack8(s + i + 8) << 32 | pack8(s + i + 16) << 16 | pack8(s + i + 24);
	if (i < L) {
		uint64_t x = 0;
		int t = 0;
		for (; i + t + 8 <= L; t += 8) x |= pack8(s + i + t) << (48 - 2 * t);
		for (; i + t < L; t++) x |= (uint64_t) ascii_code(s[i + t]) << (62 - 2 * t);
		w[i >> 5] = x;
	}
}

// Packs reads into 2-bit MSB-first u64 words (one word boundary per read) plus 2 zero words of
// tail padding. Returns per-read word offsets.
inline void pack_reads(const char *seq, const int64_t *off, const int32_t *len, int64_t n, std::vector<uint64_t> &words,
		std::vector<uint64_t> &woff) {
	woff.resize((size_t) n);
	uint64_t acc = 0;
	for (int64_t r = 0; r < n; r++) woff[(size_t) r] = acc, acc += (uint64_t) ((len[r] + 31) / 32);
	words.assign((size_t) acc + 2, 0ull);
#pragma omp parallel for schedule(static, 4096)
	for (int64_t r = 0; r < n; r++) pack_read(seq + off[r], len[r], &words[(size_t) woff[(size_t) r]]);
}

// pack_reads_into and read_flags in one pass over each read (its bytes are still in the core's
// cache for the second use): good[r] and active[r] as read_flags gives them.
inline void pack_reads_flags_into(const char *seq, const int64_t *off, const int32_t *len, int64_t n,
		const uint64_t *woff, uint64_t *out, size_t words, int64_t min_overlap, int k, uint8_t *good,
		uint8_t *active) {
	const size_t used = n ? (size_t) (woff[n - 1] + (uint64_t) ((len[n - 1] + 31) / 32)) : 0;
	for (size_t i = used; i < words; i++) out[i] = 0;
#pragma omp parallel for schedule(static, 4096)
	for (int64_t r = 0; r < n; r++) {
		read_flags(seq + off[r], len[r], min_overlap, k, good[r], active[r]);
		pack_read(seq + off[r], len[r], out + woff[r]);
	}
}

// pack_reads into a caller buffer of `words` u64 (woff from the caller; tail words cleared).
inline void pack_reads_into(const char *seq, const int64_t *off, const int32_t *len, int64_t n, const uint64_t *woff,
		uint64_t *out, size_t words) {
	const size_t used = n ? (size_t) (woff[n - 1] + (uint64_t) ((len[n - 1] + 31) / 32)) : 0;
	for (size_t i = used; i < words; i++) out[i] = 0;
#pragma omp parallel for schedule(static, 4096)
	for (int64_t r = 0; r < n; r++) pack_read(seq + off[r], len[r], out + woff[r]);
}

} // namespace hero
