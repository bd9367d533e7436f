// Deterministic synthetic sequencing reads for the correction workloads (host C++, OpenMP).
//
// Genome: i.i.d. uniform A/C/G/T from one splitmix64 stream. Optional repeat families: copies of
// `repeat_seg_len`-base source segments pasted over `repeat_frac` of the genome, each copy with
// `repeat_div` per-base substitutions. Reads: one splitmix64 stream per read index (so any read
// can be regenerated alone and the loop is parallel): length uniform in [len_min, len_max], start
// uniform over the genome, strand Bernoulli(1/2) (reverse complement), then substitutions with a
// rate rising linearly from sub_5p at the read's first base to sub_3p at its last, and 1-bp
// indels (50/50 insertion of a random base / deletion) with per-base rate indel_rate; the read
// keeps its length by consuming further template bases. Read names are their indices 0..N-1.
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <vector>

#include "hero_synth.h"

namespace {

inline uint64_t splitmix64(uint64_t &x) {
	uint64_t z = (x += 0x9E3779B97F4A7C15ull);
	z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
	z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
	return z ^ (z >> 31);
}
inline double uniform01(uint64_t &x) { return (double) (splitmix64(x) >> 11) * (1.0 / 9007199254740992.0); }
inline uint64_t below(uint64_t &x, uint64_t n) { return n ? splitmix64(x) % n : 0; }

const char kBase[4] = {'A', 'C', 'G', 'T'};
inline char comp(char c) {
	switch (c) {
	case 'A': return 'T';
	case 'C': return 'G';
	case 'G': return 'C';
	default: return 'A';
	}
}

struct Synth {
	hero_synth_params p;
	std::vector<char> genome;
	std::vector<char> seq;
	std::vector<int64_t> offset;
	std::vector<int32_t> length;
};

void make_genome(Synth &s) {
	const hero_synth_params &p = s.p;
	s.genome.resize(p.genome_len);
	uint64_t st = p.seed ^ 0x5EEDull;
	for (uint64_t i = 0; i < p.genome_len; i++) s.genome[i] = kBase[splitmix64(st) & 3];
	if (p.repeat_frac > 0 && p.repeat_seg_len > 0 && p.genome_len > 2 * p.repeat_seg_len) {
		uint64_t seg = p.repeat_seg_len;
		uint64_t copies = (uint64_t) std::llround(p.repeat_frac * (double) p.genome_len / (double) seg);
		uint64_t per_family = p.repeat_copies_per_family ? p.repeat_copies_per_family : 20;
		uint64_t families = copies / per_family + (copies % per_family ? 1 : 0);
		std::vector<std::vector<char>> src(families, std::vector<char>(seg));
		for (uint64_t f = 0; f < families; f++)
			for (uint64_t i = 0; i < seg; i++) src[f][i] = kBase[splitmix64(st) & 3];
		for (uint64_t c = 0; c < copies; c++) {
			uint64_t f = c % families;
			uint64_t at = below(st, p.genome_len - seg);
			for (uint64_t i = 0; i < seg; i++) {
				char b = src[f][i];
				if (uniform01(st) < p.repeat_div) b = kBase[(((b == 'A') ? 0 : (b == 'C') ? 1 : (b == 'G') ? 2 : 3) + 1 + below(st, 3)) & 3];
				s.genome[at + i] = b;
			}
		}
	}
}

int make_reads(Synth &s) {
	const hero_synth_params &p = s.p;
	uint64_t lmin = p.len_min, lmax = p.len_max < p.len_min ? p.len_min : p.len_max;
	uint64_t n = p.n_reads;
	if (n == 0) n = (uint64_t) std::llround(p.coverage * (double) p.genome_len / (0.5 * (double) (lmin + lmax)));
	uint64_t margin = lmax + lmax / 4 + 64; // template slack for deletions
	if (p.genome_len <= margin + 1) return -1;
	s.length.resize(n);
	s.offset.resize(n);
	for (uint64_t r = 0; r < n; r++) {
		uint64_t st = (p.seed * 0x100000001B3ull) ^ (r * 0x9E3779B97F4A7C15ull) ^ 0xA11CEull;
		s.length[r] = (int32_t) (lmin + below(st, lmax - lmin + 1));
	}
	int64_t acc = 0;
	for (uint64_t r = 0; r < n; r++) s.offset[r] = acc, acc += s.length[r];
	s.seq.resize((size_t) acc);
#pragma omp parallel for schedule(static, 1024)
	for (int64_t r = 0; r < (int64_t) n; r++) {
		uint64_t st = (p.seed * 0x100000001B3ull) ^ ((uint64_t) r * 0x9E3779B97F4A7C15ull) ^ 0xA11CEull;
		uint64_t L = (uint64_t) s.length[r];
		below(st, lmax - lmin + 1); // length draw, consumed above
		uint64_t tl = L + L / 4 + 64;
		uint64_t start = below(st, p.genome_len - tl);
		int reverse = (int) (splitmix64(st) & 1);
		std::vector<char> tmpl(tl);
		for (uint64_t i = 0; i < tl; i++)
			tmpl[i] = reverse ? comp(s.genome[start + tl - 1 - i]) : s.genome[start + i];
		char *out = &s.seq[(size_t) s.offset[r]];
		uint64_t t = 0, o = 0;
		while (o < L) {
			double frac = L > 1 ? (double) o / (double) (L - 1) : 0.0;
			double sub = p.sub_5p + (p.sub_3p - p.sub_5p) * frac;
			if (p.indel_rate > 0 && uniform01(st) < p.indel_rate) {
				if (splitmix64(st) & 1) { // insertion
					out[o++] = kBase[splitmix64(st) & 3];
					continue;
				}
				t++; // deletion
				if (t >= tl) t = tl - 1;
			}
			char b = tmpl[t < tl ? t : tl - 1];
			t++;
			if (uniform01(st) < sub) {
				int c = (b == 'A') ? 0 : (b == 'C') ? 1 : (b == 'G') ? 2 : 3;
				b = kBase[(c + 1 + (int) below(st, 3)) & 3];
			}
			out[o++] = b;
		}
	}
	return 0;
}

} // namespace

extern "C" {

void *hero_synth_create(const hero_synth_params *params) {
	if (!params || params->len_min == 0 || params->genome_len == 0) return nullptr;
	Synth *s = new Synth();
	s->p = *params;
	make_genome(*s);
	if (make_reads(*s) != 0) {
		delete s;
		return nullptr;
	}
	return s;
}

int64_t hero_synth_num_reads(const void *h) { return h ? (int64_t) ((const Synth *) h)->length.size() : -1; }
int64_t hero_synth_num_bases(const void *h) { return h ? (int64_t) ((const Synth *) h)->seq.size() : -1; }

int hero_synth_copy(const void *h, char *seq, int64_t *offset, int32_t *length) {
	if (!h) return -1;
	const Synth *s = (const Synth *) h;
	if (seq) std::memcpy(seq, s->seq.data(), s->seq.size());
	if (offset) std::memcpy(offset, s->offset.data(), s->offset.size() * sizeof(int64_t));
	if (length) std::memcpy(length, s->length.data(), s->length.size() * sizeof(int32_t));
	return 0;
}

int hero_synth_genome(const void *h, char *out, uint64_t cap) {
	if (!h) return -1;
	const Synth *s = (const Synth *) h;
	if (cap < s->genome.size()) return -2;
	std::memcpy(out, s->genome.data(), s->genome.size());
	return 0;
}

void hero_synth_destroy(void *h) { delete (Synth *) h; }

// Two-line FASTA of reads [r0, r1) named by their unique ids (the layout the reference parses,
// src/QueryDataset.cpp:103-244). Returns 0 or -1 when the file cannot be written.
int hero_write_fasta(const char *path, const char *seq, const int64_t *offset, const int32_t *length,
		const int64_t *unique_id, int64_t r0, int64_t r1) {
	FILE *f = std::fopen(path, "wb");
	if (!f) return -1;
	std::vector<char> buf;
	buf.reserve(1 << 22);
	char name[32];
	for (int64_t r = r0; r < r1; r++) {
		const int m = std::snprintf(name, sizeof name, ">%lld\n", (long long) unique_id[r]);
		buf.insert(buf.end(), name, name + m);
		buf.insert(buf.end(), seq + offset[r], seq + offset[r] + length[r]);
		buf.push_back('\n');
		if (buf.size() > (1u << 22)) {
			std::fwrite(buf.data(), 1, buf.size(), f);
			buf.clear();
		}
	}
	std::fwrite(buf.data(), 1, buf.size(), f);
	return std::fclose(f) == 0 ? 0 : -1;
}

} // extern "C"
