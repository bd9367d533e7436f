// Host-side read preparation shared by the C-ABI library and the CPU harness: the reference's
// dataset predicates and 2-bit packing into the device layout.
#pragma once
#include <stdint.h>

#include <vector>

#include "hd.h"

namespace hero {

// QueryDataset::qualityFilter (src/QueryDataset.cpp:60-75): only A/C/G/T and no base at >= 80%.
inline bool quality_filter(const char *s, int64_t n) {
	uint64_t cnt[4] = {0, 0, 0, 0};
	for (int64_t i = 0; i < n; i++) {
		const char c = s[i];
		if (c != 'A' && c != 'C' && c != 'G' && c != 'T') return false;
		cnt[ascii_code(c)]++;
	}
	const uint64_t thr = (uint64_t) ((double) n * .8);
	return !(cnt[0] >= thr || cnt[1] >= thr || cnt[2] >= thr || cnt[3] >= thr);
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

// Packs reads into 2-bit MSB-first u64 words (one word boundary per read) plus 2 zero words of
// tail padding. Returns per-read word offsets.
inline void pack_reads(const char *seq, const int64_t *off, const int32_t *len, int64_t n, std::vector<uint64_t> &words,
		std::vector<uint64_t> &woff) {
	woff.resize((size_t) n);
	uint64_t acc = 0;
	for (int64_t r = 0; r < n; r++) woff[(size_t) r] = acc, acc += (uint64_t) ((len[r] + 31) / 32);
	words.assign((size_t) acc + 2, 0ull);
#pragma omp parallel for schedule(static, 4096)
	for (int64_t r = 0; r < n; r++) {
		const char *s = seq + off[r];
		uint64_t *w = &words[(size_t) woff[(size_t) r]];
		const int L = len[r];
		for (int i = 0; i < L; i += 32) {
			uint64_t x = 0;
			const int m = L - i < 32 ? L - i : 32;
			for (int t = 0; t < m; t++) x |= (uint64_t) ascii_code(s[i + t]) << (62 - 2 * t);
			w[i >> 5] = x;
		}
	}
}

// pack_reads into a caller buffer of `words` u64 (woff from the caller; tail words cleared).
inline void pack_reads_into(const char *seq, const int64_t *off, const int32_t *len, int64_t n, const uint64_t *woff,
		uint64_t *out, size_t words) {
	const size_t used = n ? (size_t) (woff[n - 1] + (uint64_t) ((len[n - 1] + 31) / 32)) : 0;
	for (size_t i = used; i < words; i++) out[i] = 0;
#pragma omp parallel for schedule(static, 4096)
	for (int64_t r = 0; r < n; r++) {
		const char *s = seq + off[r];
		uint64_t *w = out + woff[r];
		const int L = len[r];
		for (int i = 0; i < L; i += 32) {
			uint64_t x = 0;
			const int m = L - i < 32 ? L - i : 32;
			for (int t = 0; t < m; t++) x |= (uint64_t) ascii_code(s[i + t]) << (62 - 2 * t);
			w[i >> 5] = x;
		}
	}
}

} // namespace hero
