// Synthetic read generator (test and benchmark inputs; not part of the drop-in boundary).
#ifndef HERO_SYNTH_H
#define HERO_SYNTH_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
	uint64_t seed;
	uint64_t genome_len;
	uint32_t len_min, len_max;   /* read lengths uniform in [len_min, len_max] */
	double coverage;             /* used when n_reads == 0: n = coverage * G / mean length */
	uint64_t n_reads;
	double sub_5p, sub_3p;       /* substitution rate at the read's first and last base */
	double indel_rate;           /* 1-bp indels per base */
	double repeat_frac;          /* share of the genome covered by repeat copies */
	uint32_t repeat_seg_len;     /* length of a repeat segment */
	uint32_t repeat_copies_per_family;
	double repeat_div;           /* per-base divergence of each copy */
} hero_synth_params;

void *hero_synth_create(const hero_synth_params *params);
int64_t hero_synth_num_reads(const void *h);
int64_t hero_synth_num_bases(const void *h);
int hero_synth_copy(const void *h, char *seq, int64_t *offset, int32_t *length);
int hero_synth_genome(const void *h, char *out, uint64_t cap);
void hero_synth_destroy(void *h);
int hero_write_fasta(const char *path, const char *seq, const int64_t *offset, const int32_t *length,
		const int64_t *unique_id, int64_t r0, int64_t r1);

#ifdef __cplusplus
}
#endif
#endif
