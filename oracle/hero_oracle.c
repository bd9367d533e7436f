/*
 * hero_oracle.c — CPU restatement of HERO's read-correction hot path (C11 + OpenMP).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path in hero_amd/ and the "port" CPU
 * baseline of bench.py. Nothing in the product links or calls this file.
 *
 * Each function restates the reference algorithm (guo-xuan/Hero, src/) and cites the lines it
 * follows. Reference scores are doubles that only ever hold small integers
 * (src/BandAligner.cpp:528-542), so the DP runs in int; the few genuinely floating-point decisions
 * (chain ratio test, rounding of rates, acceptance, dedup, consensus) are kept in IEEE double,
 * in the same operation order, compiled with -ffp-contract=off.
 */
#include "hero_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define MAXDUL 100000 /* src/BandAligner.h:18 */
#define BIT_OF_READ_ID 40 /* src/HashTableLongKey.cpp:15-20 (no -v / --mapping) */

static void *xmalloc(size_t n) {
	void *p = malloc(n ? n : 1);
	if (!p) {
		fprintf(stderr, "hero_oracle: out of memory (%zu bytes)\n", n);
		abort();
	}
	return p;
}
static void *xrealloc(void *p, size_t n) {
	void *q = realloc(p, n ? n : 1);
	if (!q) {
		fprintf(stderr, "hero_oracle: out of memory (%zu bytes)\n", n);
		abort();
	}
	return q;
}
void oracle_free(void *p) { free(p); }

void oracle_default_cfg(oracle_cfg *c) {
	memset(c, 0, sizeof *c);
	c->k = 39;                  /* Config.cpp:17 */
	c->aligner = 1;             /* Config.cpp:35 IGLOBAL */
	c->call_consensus = 0;      /* Config.cpp:45 */
	c->hash_factor = 32;        /* Config.cpp:37 */
	c->min_overlap = 40;        /* Config.cpp:32 */
	c->mismatch_rate = 1.0;     /* Config.cpp:25 */
	c->indel_rate = 1.0;        /* Config.cpp:29 */
	c->consensus_freq = 0.6;    /* Config.cpp:47 */
	c->dx = -4, c->dy = -4, c->miss = -4, c->hit = 1, c->unobserved = -1, c->sx = 0, c->ex = -1; /* BandAligner.cpp:528-542 */
}

/* ------------------------------------------------------------------------------------------ */
/* 2-bit code and complement (src/QueryRead.cpp:83-95,124-125; SubjectRead.cpp:131-141)       */
/* ------------------------------------------------------------------------------------------ */
static inline int code_of(char c) { return ((unsigned char) c >> 1) & 3; }

static void reverse_complement(const char *s, int64_t n, char *out) {
	for (int64_t i = 0; i < n; i++) {
		char c = s[i];
		out[n - 1 - i] = (c & 0x02) ? (char) (c ^ 0x04) : (char) (c ^ 0x15);
	}
}

/* QueryDataset::qualityFilter (src/QueryDataset.cpp:60-75). */
static int quality_filter(const char *s, int64_t n) {
	uint64_t cnt[4] = {0, 0, 0, 0};
	for (int64_t i = 0; i < n; i++) {
		char c = s[i];
		if (c != 'A' && c != 'C' && c != 'G' && c != 'T') return 0;
		cnt[code_of(c)]++;
	}
	uint64_t thr = (uint64_t) ((double) n * .8);
	if (cnt[0] >= thr || cnt[1] >= thr || cnt[2] >= thr || cnt[3] >= thr) return 0;
	return 1;
}

/* Query predicate (src/QueryDataset.cpp:226): good iff len > minOverlap && qualityFilter. */
static int query_is_good(const oracle_cfg *c, const char *s, int64_t n) {
	uint16_t mo = (uint16_t) c->min_overlap; /* UINT16 dataset_minimumoverlaplength */
	return (uint64_t) n > mo && quality_filter(s, n);
}
/* Subject predicate (src/SubjectDataset.cpp:157-162): good iff len < minOverlap || qualityFilter. */
static int subject_is_good(const oracle_cfg *c, const char *s, int64_t n) {
	uint16_t mo = (uint16_t) c->min_overlap; /* Config::getminimumOverlapLength() returns UINT16 */
	return (uint64_t) n < mo || quality_filter(s, n);
}

/* Both dataset predicates of every read of a set. */
void oracle_read_flags(const oracle_cfg *cfg, const oracle_reads *r, uint8_t *query_good, uint8_t *subject_good) {
	for (int64_t i = 0; i < r->n; i++) {
		query_good[i] = (uint8_t) query_is_good(cfg, r->seq + r->offset[i], r->length[i]);
		subject_good[i] = (uint8_t) subject_is_good(cfg, r->seq + r->offset[i], r->length[i]);
	}
}

/* ------------------------------------------------------------------------------------------ */
/* Hash function and table size                                                              */
/* ------------------------------------------------------------------------------------------ */
/* HashTableLongKey::hashFunction (src/HashTableLongKey.cpp:188-206): sum1 takes the first 32
 * bases MSB-first after a leading 1 (shifted out once 32 bases are in), sum2 the rest; the
 * product of the two residues wraps in u64 before the final modulus. */
uint64_t oracle_hash_kmer(const char *s, int64_t k, uint64_t P) {
	uint64_t s1 = 1, s2 = 1;
	for (int64_t i = 0; i < k; i++) {
		uint64_t c = ((uint64_t) (unsigned char) s[i] >> 1) & 3ull;
		if (i < 32) s1 = (s1 << 2) | c;
		else s2 = (s2 << 2) | c;
	}
	return ((s1 % P) * (s2 % P)) % P;
}

static uint64_t mulmod64(uint64_t a, uint64_t b, uint64_t m) { return (uint64_t) (((__uint128_t) a * b) % m); }
static uint64_t powmod64(uint64_t a, uint64_t e, uint64_t m) {
	uint64_t r = 1;
	a %= m;
	while (e) {
		if (e & 1) r = mulmod64(r, a, m);
		a = mulmod64(a, a, m);
		e >>= 1;
	}
	return r;
}
/* Deterministic Miller-Rabin for 64-bit n (bases 2..37 cover all n < 3.3e24). */
static int is_prime64(uint64_t n) {
	static const uint64_t bases[12] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
	if (n < 2) return 0;
	for (int i = 0; i < 12; i++) {
		if (n == bases[i]) return 1;
		if (n % bases[i] == 0) return 0;
	}
	uint64_t d = n - 1;
	int r = 0;
	while ((d & 1) == 0) d >>= 1, r++;
	for (int i = 0; i < 12; i++) {
		uint64_t x = powmod64(bases[i], d, n);
		if (x == 1 || x == n - 1) continue;
		int comp = 1;
		for (int j = 1; j < r; j++) {
			x = mulmod64(x, x, n);
			if (x == n - 1) {
				comp = 0;
				break;
			}
		}
		if (comp) return 0;
	}
	return 1;
}
/* Role of HashTableLongKey::createHashTables + getPrimeLargerThanNumber
 * (src/HashTableLongKey.cpp:41-68,101-166): a prime strictly larger than factor*N. The
 * reference takes it from a fixed list; slot placement never reaches the output (lookups return
 * exactly the bucket of the probed k-mer), so the smallest prime is used here. */
uint64_t oracle_table_size(uint64_t n_reads, uint64_t factor) {
	uint64_t n = n_reads * factor;
	uint64_t p = n + 1;
	while (!is_prime64(p)) p++;
	return p;
}

/* ------------------------------------------------------------------------------------------ */
/* Query index (HashTableLongKey)                                                             */
/* ------------------------------------------------------------------------------------------ */
typedef struct {
	uint64_t P;
	int64_t *slot;       /* bucket id per slot, -1 = empty (NULL / empty vector) */
	int64_t *bstart;     /* CSR start per bucket */
	int64_t *bcount;
	int64_t *entries;    /* (pos << 40) | innerId, insertion order (src/HashTableLongKey.cpp:334-337) */
	int64_t nbuckets;
	uint64_t max_probe;  /* maxSingleHashCollision */
	uint64_t total_keys;
	const oracle_reads *q;
} index_t;

static inline const char *qseq(const oracle_reads *r, int64_t id) { return r->seq + r->offset[id]; }

/* Key positions of one query read (src/HashTableLongKey.cpp:279-289): 0, k, 2k, ... <= L-k, then
 * L-k when the stride loop left the read's tail uncovered. */
static int key_positions(int64_t L, int64_t k, int64_t *out) {
	int n = 0;
	int64_t i;
	if (L < k) return 0;
	for (i = 0; i <= L - k; i += k) out[n++] = i;
	if (i < L) out[n++] = L - k;
	return n;
}

/* HashTableLongKey::insertQueryDataset / insertQueryRead (src/HashTableLongKey.cpp:250-342):
 * good reads in innerId order; probe linearly from the hash until an empty slot or a slot whose
 * first entry's k-mer (read back from its query) equals the key; append. Two passes (count, fill)
 * give the same bucket contents and order as the reference's push_back. */
static int index_build(const oracle_cfg *c, const oracle_reads *q, index_t *ix) {
	int64_t k = c->k;
	memset(ix, 0, sizeof *ix);
	ix->q = q;
	/* --mapping sizes the table by the k-mers of the bases, not by the reads
	 * (src/HashTableLongKey.cpp:52-60) */
	uint64_t units = (uint64_t) q->n;
	if (c->mapping) {
		uint64_t total = 0;
		for (int64_t i = 0; i < q->n; i++) total += (uint64_t) q->length[i];
		units = total / (uint64_t) k;
	}
	ix->P = oracle_table_size(units, (uint64_t) c->hash_factor);
	ix->slot = (int64_t *) xmalloc(sizeof(int64_t) * ix->P);
	for (uint64_t s = 0; s < ix->P; s++) ix->slot[s] = -1;
	int64_t cap_b = 1024;
	int64_t *first = (int64_t *) xmalloc(sizeof(int64_t) * cap_b); /* first entry of each bucket */
	ix->bcount = (int64_t *) xmalloc(sizeof(int64_t) * cap_b);
	int64_t *key_slot = NULL, nkeys = 0, cap_k = 0;
	int64_t pos[512];
	for (int pass = 0; pass < 2; pass++) {
		int64_t kk = 0;
		for (int64_t id = 0; id < q->n; id++) {
			const char *s = qseq(q, id);
			int64_t L = q->length[id];
			if (!query_is_good(c, s, L)) continue; /* src/HashTableLongKey.cpp:272-275 */
			if (L - k >= 512 * k) return -3;
			int np = key_positions(L, k, pos);
			for (int t = 0; t < np; t++) {
				int64_t p = pos[t];
				int64_t entry = (p << BIT_OF_READ_ID) | id;
				if (pass == 0) {
					uint64_t h = oracle_hash_kmer(s + p, k, ix->P);
					uint64_t coll = 0;
					while (ix->slot[h] >= 0) {
						int64_t f = first[ix->slot[h]];
						int64_t fid = f & ((1ll << BIT_OF_READ_ID) - 1), fpos = f >> BIT_OF_READ_ID;
						if (memcmp(qseq(q, fid) + fpos, s + p, (size_t) k) == 0) break;
						coll++;
						h = (h == ix->P - 1) ? 0 : h + 1;
						if (coll >= ix->P) return -2; /* "The size of hash table is too small." */
					}
					if (ix->slot[h] < 0) {
						if (ix->nbuckets == cap_b) {
							cap_b *= 2;
							first = (int64_t *) xrealloc(first, sizeof(int64_t) * cap_b);
							ix->bcount = (int64_t *) xrealloc(ix->bcount, sizeof(int64_t) * cap_b);
						}
						first[ix->nbuckets] = entry;
						ix->bcount[ix->nbuckets] = 0;
						ix->slot[h] = ix->nbuckets++;
					}
					ix->bcount[ix->slot[h]]++;
					if (coll > ix->max_probe) ix->max_probe = coll;
					if (nkeys == cap_k) {
						cap_k = cap_k ? 2 * cap_k : 4096;
						key_slot = (int64_t *) xrealloc(key_slot, sizeof(int64_t) * cap_k);
					}
					key_slot[nkeys++] = ix->slot[h];
				} else {
					int64_t b = key_slot[kk++];
					ix->entries[ix->bstart[b] + ix->bcount[b]++] = entry;
				}
			}
		}
		if (pass == 0) {
			ix->bstart = (int64_t *) xmalloc(sizeof(int64_t) * (ix->nbuckets + 1));
			int64_t acc = 0;
			for (int64_t b = 0; b < ix->nbuckets; b++) {
				ix->bstart[b] = acc;
				acc += ix->bcount[b];
				ix->bcount[b] = 0;
			}
			ix->bstart[ix->nbuckets] = acc;
			ix->entries = (int64_t *) xmalloc(sizeof(int64_t) * (acc ? acc : 1));
			ix->total_keys = (uint64_t) acc;
		}
	}
	free(first);
	free(key_slot);
	return 0;
}

static void index_free(index_t *ix) {
	free(ix->slot);
	free(ix->bstart);
	free(ix->bcount);
	free(ix->entries);
	memset(ix, 0, sizeof *ix);
}

/* HashTableLongKey::getListOfReads + compareSequence (src/HashTableLongKey.cpp:70-99): probe from
 * the hash; stop at an empty slot (no bucket), at a slot whose first k-mer equals the window
 * (that bucket), or after more than maxSingleHashCollision non-matching slots (no bucket). */
static int64_t index_lookup(const index_t *ix, int64_t k, const char *w) {
	uint64_t h = oracle_hash_kmer(w, k, ix->P);
	uint64_t coll = 0;
	while (ix->slot[h] >= 0) {
		int64_t b = ix->slot[h];
		int64_t f = ix->entries[ix->bstart[b]];
		int64_t fid = f & ((1ll << BIT_OF_READ_ID) - 1), fpos = f >> BIT_OF_READ_ID;
		if (memcmp(qseq(ix->q, fid) + fpos, w, (size_t) k) == 0) return b;
		coll++;
		if (coll > ix->max_probe) return -1;
		h = (h == ix->P - 1) ? 0 : h + 1;
	}
	return -1;
}

/* ------------------------------------------------------------------------------------------ */
/* Seeds                                                                                      */
/* ------------------------------------------------------------------------------------------ */
typedef struct {
	int64_t qid, ps, pq; /* PairKeyPosition: iQueryReadInnerId, iPositionOnSubject, iPositionOnQuery */
} seed_t;

/* comparePairKeyPosition (src/SubjectRead.cpp:157-173): (innerId, posQ, posS) ascending. */
static int seed_cmp(const void *a, const void *b) {
	const seed_t *x = (const seed_t *) a, *y = (const seed_t *) b;
	if (x->qid != y->qid) return x->qid < y->qid ? -1 : 1;
	if (x->pq != y->pq) return x->pq < y->pq ? -1 : 1;
	if (x->ps != y->ps) return x->ps < y->ps ? -1 : 1;
	return 0;
}

typedef struct {
	seed_t *v;
	int64_t n, cap;
} seedvec;

static void seedvec_push(seedvec *sv, seed_t s) {
	if (sv->n == sv->cap) {
		sv->cap = sv->cap ? 2 * sv->cap : 256;
		sv->v = (seed_t *) xrealloc(sv->v, sizeof(seed_t) * sv->cap);
	}
	sv->v[sv->n++] = s;
}

/* SubjectReadAligner::searchHashTable2 (src/SubjectReadAligner.cpp:444-491): every window
 * 0..L-k (inclusive) of the strand sequence; skip hits on the query with the subject's uniqueId;
 * without -c, reverse-strand coordinates are mapped back to forward frames. Then sort. */
static void search_seeds(const oracle_cfg *c, const index_t *ix, const char *strand_seq, int64_t L,
		int64_t subject_uid, int reverse, seedvec *out, uint64_t *lookups) {
	int64_t k = c->k;
	out->n = 0;
	for (int64_t i = 0; i <= L - k; i++) {
		(*lookups)++;
		int64_t b = index_lookup(ix, k, strand_seq + i);
		if (b < 0) continue;
		for (int64_t e = ix->bstart[b]; e < ix->bstart[b] + ix->bcount[b]; e++) {
			int64_t data = ix->entries[e];
			int64_t qid = data & ((1ll << BIT_OF_READ_ID) - 1);
			if (ix->q->unique_id[qid] == subject_uid) continue;
			seed_t s;
			s.qid = qid;
			s.ps = i;
			s.pq = data >> BIT_OF_READ_ID;
			if (!c->call_consensus && reverse) {
				s.ps = L - (s.ps + k);
				s.pq = (int64_t) ix->q->length[qid] - (s.pq + k);
			}
			seedvec_push(out, s);
		}
	}
	if (out->n > 1) qsort(out->v, (size_t) out->n, sizeof(seed_t), seed_cmp);
}

/* ------------------------------------------------------------------------------------------ */
/* BandAligner                                                                                */
/* ------------------------------------------------------------------------------------------ */
typedef struct {
	int *FM, *IX, *IY;
	unsigned char *BM;
	size_t cap;
	int *edit;
	size_t edit_cap;
	int32_t *cnt;
	char *typ;
	int nops, ops_cap;
} dp_ws;

static void dp_ws_reserve(dp_ws *w, size_t cells, size_t edits) {
	if (cells > w->cap) {
		w->cap = cells + cells / 2;
		w->FM = (int *) xrealloc(w->FM, sizeof(int) * w->cap);
		w->IX = (int *) xrealloc(w->IX, sizeof(int) * w->cap);
		w->IY = (int *) xrealloc(w->IY, sizeof(int) * w->cap);
		w->BM = (unsigned char *) xrealloc(w->BM, w->cap);
	}
	if (edits > w->edit_cap) {
		w->edit_cap = edits + edits / 2;
		w->edit = (int *) xrealloc(w->edit, sizeof(int) * w->edit_cap);
	}
	if ((int) edits + 1 > w->ops_cap) {
		w->ops_cap = (int) edits + 16;
		w->cnt = (int32_t *) xrealloc(w->cnt, sizeof(int32_t) * w->ops_cap);
		w->typ = (char *) xrealloc(w->typ, (size_t) w->ops_cap);
	}
}
static void dp_ws_free(dp_ws *w) {
	free(w->FM), free(w->IX), free(w->IY), free(w->BM), free(w->edit), free(w->cnt), free(w->typ);
	memset(w, 0, sizeof *w);
}

/* Band storage: cell (i,j) of the (lenS+1)x(lenQ+1) matrix lives at row i, column j-i+b. */
#define AT(M, i, j) (M)[(size_t) (i) * W + (size_t) ((j) - (i) + b)]

static inline int64_t imax64(int64_t a, int64_t b) { return a > b ? a : b; }
static inline int64_t imin64(int64_t a, int64_t b) { return a < b ? a : b; }
static inline int imax(int a, int b) { return a > b ? a : b; }

/* In-band cells of one call: sum over rows of the row width, minus the origin. */
static uint64_t band_cells(int64_t lenS, int64_t lenQ, int64_t b) {
	uint64_t n = 0;
	for (int64_t i = 0; i <= lenS; i++) n += (uint64_t) (imin64(lenQ, i + b) - imax64(0, i - b) + 1);
	return n - 1;
}

/* BandAligner::bandedAlign (src/BandAligner.cpp:56-112): linear gaps; the upward move is free in
 * column 0 under setAlignRight and in column lenQ under setAlignLeft; N against a base scores
 * unobserved_match; ties prefer diagonal, then left, then up. */
static void dp_global(const oracle_cfg *c, const char *S, int64_t lenS, const char *Q, int64_t lenQ, int64_t b,
		int left_al, int right_al, dp_ws *w) {
	int64_t W = 2 * b + 1;
	int *FM = w->FM;
	unsigned char *BM = w->BM;
	AT(FM, 0, 0) = 0;
	for (int64_t i = 0; i <= lenS; i++) {
		int64_t lo = imax64(0, i - b), hi = imin64(lenQ, i + b);
		for (int64_t j = lo; j <= hi; j++) {
			if (i == 0 && j == 0) continue;
			int up = -MAXDUL, lf = -MAXDUL, dg = -MAXDUL;
			if (i > 0 && j < i + b) {
				int g;
				if (j > 0 && j < lenQ) g = c->dx;
				else if (j == 0 && right_al) g = c->sx;
				else if (j >= lenQ && left_al) g = c->sx;
				else g = c->dx;
				up = AT(FM, i - 1, j) + g;
			}
			if (j > lo) lf = c->dy + AT(FM, i, j - 1);
			if (i > 0 && j > 0) {
				char a = S[i - 1], q = Q[j - 1];
				int s = (a != q) ? ((a == 'N' || q == 'N') ? c->unobserved : c->miss) : c->hit;
				dg = AT(FM, i - 1, j - 1) + s;
			}
			if (dg >= lf && dg >= up) {
				AT(FM, i, j) = dg;
				AT(BM, i, j) = 2;
			} else if (lf >= up) {
				AT(FM, i, j) = lf;
				AT(BM, i, j) = 1;
			} else {
				AT(FM, i, j) = up;
				AT(BM, i, j) = 0;
			}
		}
	}
}

/* BandAligner::bandedAlignAffineGap (src/BandAligner.cpp:139-225): three matrices, boundary
 * initialisation of column 0 (free under setAlignRight) and row 0 within the band, IY cleared at
 * the band's left edge, IX at its right edge, free trailing subject gap under setAlignLeft, and a
 * single pointer per cell: 2 if FM > IX and FM >= IY, else 1 if IY > IX, else 0. */
static void dp_affine(const oracle_cfg *c, const char *S, int64_t lenS, const char *Q, int64_t lenQ, int64_t b,
		int left_al, int right_al, dp_ws *w) {
	int64_t W = 2 * b + 1;
	int *FM = w->FM, *IX = w->IX, *IY = w->IY;
	unsigned char *BM = w->BM;
	AT(FM, 0, 0) = 0;
	AT(IX, 0, 0) = c->dx - c->ex;
	AT(IY, 0, 0) = c->dx - c->ex;
	int64_t hi0 = imin64(lenS, b);
	for (int64_t i = 1; i <= hi0; i++) {
		if (right_al) {
			AT(IX, i, 0) = AT(FM, i - 1, 0) + c->sx;
			AT(IY, i, 0) = c->dy;
			AT(FM, i, 0) = AT(IX, i, 0);
		} else {
			AT(IX, i, 0) = AT(IX, i - 1, 0) + c->ex;
			AT(IY, i, 0) = AT(IX, i, 0);
			AT(FM, i, 0) = AT(IX, i, 0);
		}
		AT(BM, i, 0) = 0;
	}
	int64_t hj0 = imin64(lenQ, b);
	for (int64_t j = 1; j <= hj0; j++) {
		AT(IY, 0, j) = AT(IY, 0, j - 1) + c->ex;
		AT(IX, 0, j) = AT(IY, 0, j);
		AT(FM, 0, j) = AT(IY, 0, j);
		AT(BM, 0, j) = 1;
	}
	for (int64_t i = 1; i <= lenS; i++) {
		int64_t lo = imax64(0, i - b), hi = imin64(lenQ, i + b);
		for (int64_t j = lo; j <= hi; j++) {
			if (j == 0) continue;
			int iy, ix, fm;
			if (j == lo) iy = -MAXDUL;
			else iy = imax(AT(FM, i, j - 1) + c->dy, AT(IY, i, j - 1) + c->ex);
			if (j == hi && i + b <= hi) ix = -MAXDUL;
			else if (left_al && j >= lenQ) ix = imax(AT(FM, i - 1, j) + c->sx, AT(IX, i - 1, j) + c->sx);
			else ix = imax(AT(FM, i - 1, j) + c->dx, AT(IX, i - 1, j) + c->ex);
			int s = (S[i - 1] != Q[j - 1]) ? c->miss : c->hit;
			fm = imax(imax(AT(FM, i - 1, j - 1) + s, AT(IX, i - 1, j - 1) + s), AT(IY, i - 1, j - 1) + s);
			AT(IY, i, j) = iy;
			AT(IX, i, j) = ix;
			AT(FM, i, j) = fm;
			AT(BM, i, j) = (fm > ix && fm >= iy) ? 2 : (iy > ix ? 1 : 0);
		}
	}
}

/* Traceback (src/BandAligner.cpp:114-125 / 227-238) with si={-1,0,-1}, sj={0,-1,-1}
 * (src/BandAligner.cpp:536-541): edit = sj-si, i.e. 1 (up), -1 (left), 0 (diagonal), written
 * backwards; returns ei, the index of the first edit. */
static int64_t traceback(const unsigned char *BM, int64_t lenS, int64_t lenQ, int64_t b, int *edit) {
	int64_t W = 2 * b + 1;
	static const int si[3] = {-1, 0, -1}, sj[3] = {0, -1, -1};
	int64_t ei = lenQ + lenS, i = lenS, j = lenQ;
	int id = AT(BM, i, j);
	for (;;) {
		edit[--ei] = sj[id] - si[id];
		i += si[id];
		j += sj[id];
		if (i == 0 && j == 0) break;
		id = AT(BM, i, j);
	}
	return ei;
}

/* BandAligner::calculateCigarAndPosition[Reversed] (src/BandAligner.cpp:258-419): run-length over
 * the edits; a mismatch is a count-1 op typed by the original query's base (the aligner's query,
 * or its subject when the pair was swapped); I/D swap when reversed. Runs of M, I, D merge,
 * mismatches never do. */
static void cigar_from_edits(const char *S, const char *Q, const int *edit, int64_t ei, int64_t n, int reversed,
		dp_ws *w) {
	int64_t si_ = 0, qi = 0;
	w->nops = 0;
	char cur = 0;
	int32_t cnt = 0;
	for (int64_t e = ei; e < n; e++) {
		char t;
		int mismatch = 0;
		switch (edit[e]) {
		case -1:
			t = reversed ? 'D' : 'I';
			qi++;
			break;
		case 0:
			if (Q[qi] == S[si_]) t = 'M';
			else {
				t = reversed ? S[si_] : Q[qi];
				mismatch = 1;
			}
			qi++;
			si_++;
			break;
		default:
			t = reversed ? 'I' : 'D';
			si_++;
			break;
		}
		if (e == ei) {
			cur = t;
			cnt = 1;
		} else if (!mismatch && t == cur) {
			cnt++;
		} else {
			w->cnt[w->nops] = cnt;
			w->typ[w->nops] = cur;
			w->nops++;
			cur = t;
			cnt = 1;
		}
	}
	w->cnt[w->nops] = cnt;
	w->typ[w->nops] = cur;
	w->nops++;
}

/* One BandAligner call from setSequences to the CIGAR (src/BandAligner.cpp:241-256,559-573):
 * the longer side becomes the aligner's subject (query side wins ties); the left/right flag then
 * applies to the aligner's own subject/query; m_band widens to lenS-lenQ. */
static void band_align_call(const oracle_cfg *c, const char *subj, int64_t s0, int64_t s1, const char *qry, int64_t q0,
		int64_t q1, int64_t *band, int flag, dp_ws *w, uint64_t *cells) {
	const char *S, *Q;
	int64_t lenS, lenQ;
	int reversed;
	if (q1 - q0 <= s1 - s0) {
		Q = qry + q0, lenQ = q1 - q0, S = subj + s0, lenS = s1 - s0, reversed = 0;
	} else {
		Q = subj + s0, lenQ = s1 - s0, S = qry + q0, lenS = q1 - q0, reversed = 1;
	}
	if (lenS - lenQ > *band) *band = lenS - lenQ;
	int64_t b = *band;
	dp_ws_reserve(w, (size_t) (lenS + 1) * (size_t) (2 * b + 1), (size_t) (lenS + lenQ + 1));
	int left_al = flag == 1, right_al = flag == 2;
	if (c->aligner == 0) dp_affine(c, S, lenS, Q, lenQ, b, left_al, right_al, w);
	else dp_global(c, S, lenS, Q, lenQ, b, left_al, right_al, w);
	int64_t ei = traceback(w->BM, lenS, lenQ, b, w->edit);
	cigar_from_edits(S, Q, w->edit, ei, lenS + lenQ, reversed, w);
	if (cells) *cells += band_cells(lenS, lenQ, b);
}

int oracle_band_align(const oracle_cfg *cfg, const char *subject, int64_t s0, int64_t s1, const char *query,
		int64_t q0, int64_t q1, int64_t *band, int flag, int32_t *counts, char *types, int cap, uint64_t *cells) {
	if (s1 - s0 <= 0 && q1 - q0 <= 0) return -2; /* the reference would index edit[-1] */
	if (s0 < 0 || q0 < 0) return -3;            /* src/BandAligner.cpp:546-549,577-580 exit(1) */
	dp_ws w;
	memset(&w, 0, sizeof w);
	band_align_call(cfg, subject, s0, s1, query, q0, q1, band, flag, &w, cells);
	int n = w.nops;
	if (n > cap) {
		dp_ws_free(&w);
		return -1;
	}
	memcpy(counts, w.cnt, sizeof(int32_t) * (size_t) n);
	memcpy(types, w.typ, (size_t) n);
	dp_ws_free(&w);
	return n;
}

/* ------------------------------------------------------------------------------------------ */
/* AlignmentRecord                                                                            */
/* ------------------------------------------------------------------------------------------ */
typedef struct {
	int32_t *cnt;
	char *typ;
	int n, cap;
	int64_t pos;       /* iPositionOfQueryOnSubject */
	int orient;        /* iOrientation */
	int64_t qid;       /* iQueryReadUniqueId (innerId under -c, src/SubjectReadAligner.cpp:230-235) */
	const char *qseq;  /* querySequence used in the alignment */
	double nmis, nindel, len, irate, mrate;
} rec_t;

static void rec_init(rec_t *r) { memset(r, 0, sizeof *r), r->mrate = 100, r->irate = 100; }
static void rec_free(rec_t *r) {
	if (r->orient & 2) free((void *) r->qseq); /* own copy of a query RC (see process_strand) */
	free(r->cnt), free(r->typ), memset(r, 0, sizeof *r);
}

/* AlignmentRecord::addCigar (src/AlignmentRecord.cpp:35-48): merge into the last op if equal type. */
static void rec_add(rec_t *r, int64_t count, char type) {
	if (r->n > 0 && r->typ[r->n - 1] == type) {
		r->cnt[r->n - 1] += (int32_t) count;
		return;
	}
	if (r->n == r->cap) {
		r->cap = r->cap ? 2 * r->cap : 16;
		r->cnt = (int32_t *) xrealloc(r->cnt, sizeof(int32_t) * (size_t) r->cap);
		r->typ = (char *) xrealloc(r->typ, (size_t) r->cap);
	}
	r->cnt[r->n] = (int32_t) count;
	r->typ[r->n] = type;
	r->n++;
}

/* AlignmentRecord::calculateScore (src/AlignmentRecord.cpp:56-100): mismatch ops count
 * everywhere; I/D only in interior ops; leading/trailing I/D do not add to the length. */
static void rec_score(rec_t *r) {
	r->nmis = 0, r->nindel = 0, r->len = 0;
	int n = r->n;
	if (n > 0) {
		char t = r->typ[0];
		if (t != 'I' && t != 'D') {
			if (t == 'M') r->len += r->cnt[0];
			else r->nmis += r->cnt[0], r->len += r->cnt[0];
		}
		for (int i = 1; i < n - 1; i++) {
			t = r->typ[i];
			if (t == 'M') r->len += r->cnt[i];
			else if (t == 'D' || t == 'I') r->nindel += r->cnt[i], r->len += r->cnt[i];
			else r->nmis += r->cnt[i], r->len += r->cnt[i];
		}
	}
	if (n > 1) {
		char t = r->typ[n - 1];
		if (t != 'I' && t != 'D') {
			if (t == 'M') r->len += r->cnt[n - 1];
			else r->nmis += r->cnt[n - 1], r->len += r->cnt[n - 1];
		}
	}
	if (r->len == 0) r->irate = 100, r->mrate = 100;
	else r->irate = r->nindel / r->len, r->mrate = r->nmis / r->len;
}

/* isGoodAlignment(_numMisMatch, _numIndel) and isGoodAlignment() (src/AlignmentRecord.cpp:302-316). */
static int rec_good_caps(rec_t *r, int64_t cap_mis, int64_t cap_indel) {
	rec_score(r);
	return !(r->nmis > (double) cap_mis || r->nindel > (double) cap_indel);
}
static int rec_good_rates(const oracle_cfg *c, rec_t *r) {
	rec_score(r);
	return r->mrate <= c->mismatch_rate && r->irate <= c->indel_rate;
}

int oracle_record_score(const oracle_cfg *cfg, int32_t *counts, char *types, int n, int64_t cap_mis, int64_t cap_indel,
		double *out) {
	rec_t r;
	rec_init(&r);
	for (int i = 0; i < n; i++) rec_add(&r, counts[i], types[i]);
	out[3] = rec_good_caps(&r, cap_mis, cap_indel);
	out[4] = rec_good_rates(cfg, &r);
	out[0] = r.nmis, out[1] = r.nindel, out[2] = r.len, out[5] = r.irate + r.mrate;
	for (int i = 0; i < r.n; i++) counts[i] = r.cnt[i], types[i] = r.typ[i];
	int m = r.n;
	rec_free(&r);
	return m;
}

/* ------------------------------------------------------------------------------------------ */
/* getAlignment2 (src/SubjectReadAligner.cpp:197-389)                                          */
/* ------------------------------------------------------------------------------------------ */
typedef struct {
	uint64_t dp_calls, dp_cells, records_built;
} run_counters;

static void rec_add_dp(rec_t *r, const dp_ws *w) { /* AlignmentRecord::addAlignment (:50-54) */
	for (int i = 0; i < w->nops; i++) rec_add(r, w->cnt[i], w->typ[i]);
}

/* Builds one record for a chain of seeds; returns 1 and fills *out when the record is accepted
 * (caller owns it), 0 otherwise. Segment geometry, band persistence across the record's DP calls
 * (setBand once, widened in place by each call) and early exits follow the reference. */
static int get_alignment(const oracle_cfg *c, const char *subj, int64_t Ls, const char *qs, int64_t Lq, int64_t qid,
		int orient, const seed_t *keys, int nkeys, dp_ws *w, run_counters *rc, rec_t *out) {
	int64_t k = c->k;
	int64_t pq = keys[0].pq, ps = keys[0].ps;
	int64_t ov = (pq < ps ? pq : ps);
	ov += (Lq - pq < Ls - ps) ? Lq - pq : Ls - ps;
	if (ov < c->min_overlap) return 0; /* :221-224 */
	int64_t extra = (int64_t) round((double) ov * c->indel_rate);    /* :225 */
	int64_t max_mis = (int64_t) round((double) ov * c->mismatch_rate); /* :226 */
	int64_t max_indel = extra;
	rec_t *r = out;
	rec_init(r);
	r->orient = orient;
	r->qid = qid;
	r->qseq = qs;
	int64_t band = c->band_override > 0 ? c->band_override : extra * 2; /* setBand (:241) */
	/* left end (:244-281) */
	if (pq != 0 && ps != 0) {
		if (pq < ps) {
			int64_t s0 = ps - pq - extra < 0 ? 0 : ps - pq - extra;
			r->pos = s0;
			band_align_call(c, subj, s0, ps, qs, 0, pq, &band, 2, w, &rc->dp_cells);
		} else if (pq == ps) {
			r->pos = 0;
			band_align_call(c, subj, 0, ps, qs, 0, pq, &band, 2, w, &rc->dp_cells);
		} else {
			int64_t q0 = pq - ps - extra < 0 ? 0 : pq - ps - extra;
			r->pos = 0;
			if (pq - ps - extra > 0) rec_add(r, pq - ps - extra, 'I');
			band_align_call(c, subj, 0, ps, qs, q0, pq, &band, 2, w, &rc->dp_cells);
		}
		rc->dp_calls++;
		rec_add_dp(r, w);
		rec_add(r, k, 'M');
	} else {
		if (ps == 0 && pq == 0) r->pos = 0;
		else if (ps == 0) r->pos = 0, rec_add(r, pq, 'I');
		else r->pos = ps;
		rec_add(r, k, 'M');
	}
	if (!rec_good_caps(r, max_mis, max_indel)) goto reject;
	/* middle (:288-345) */
	for (int i = 1; i < nkeys; i++) {
		int64_t cq = keys[i].pq, cs = keys[i].ps;
		if (cq - pq == cs - ps && ps + k >= cs) {
			if (ps + k > cs) rec_add(r, cs - ps, 'M');
			else rec_add(r, k, 'M');
		} else if (ps + k == cs) {
			if (cq - pq - k > 0) {
				rec_add(r, cq - pq - k, 'I');
				rec_add(r, k, 'M');
			} else {
				rec_add(r, -cq + pq + k, 'D');
				rec_add(r, cq - pq, 'M');
			}
		} else if (ps + k > cs) {
			if (cq - pq < cs - ps) {
				rec_add(r, cs - ps - cq + pq, 'D');
				rec_add(r, cq - pq, 'M');
			} else {
				rec_add(r, cq - pq + ps - cs, 'I');
				rec_add(r, cs - ps, 'M');
			}
		} else {
			if (pq + k == cq) {
				rec_add(r, cs - ps - k, 'D');
				rec_add(r, k, 'M');
			} else if (pq + k > cq) {
				rec_add(r, cs - ps + pq - cq, 'D');
				rec_add(r, k + pq - cq, 'M');
			} else {
				band_align_call(c, subj, ps + k, cs, qs, pq + k, cq, &band, 0, w, &rc->dp_cells);
				rc->dp_calls++;
				rec_add_dp(r, w);
				rec_add(r, k, 'M');
			}
		}
		pq = cq, ps = cs;
		if (!rec_good_caps(r, max_mis, max_indel)) goto reject;
	}
	/* right end (:348-381) */
	pq = keys[nkeys - 1].pq + k;
	ps = keys[nkeys - 1].ps + k;
	if (pq < Lq) {
		if (ps == Ls) {
			rec_add(r, Lq - pq, 'I');
		} else if (Ls - ps < Lq - pq) {
			int64_t qe = Ls - ps + pq + extra < Lq ? Ls - ps + pq + extra : Lq;
			band_align_call(c, subj, ps, Ls, qs, pq, qe, &band, 1, w, &rc->dp_cells);
			rc->dp_calls++;
			rec_add_dp(r, w);
			if (Ls - ps + pq + extra < Lq) rec_add(r, Lq - (Ls - ps + pq + extra), 'I');
		} else if (Ls - ps == Lq - pq) {
			band_align_call(c, subj, ps, Ls, qs, pq, Lq, &band, 1, w, &rc->dp_cells);
			rc->dp_calls++;
			rec_add_dp(r, w);
		} else {
			int64_t se = ps + (Lq - pq) + extra < Ls ? ps + (Lq - pq) + extra : Ls;
			band_align_call(c, subj, ps, se, qs, pq, Lq, &band, 1, w, &rc->dp_cells);
			rc->dp_calls++;
			rec_add_dp(r, w);
		}
	}
	if (!rec_good_rates(c, r)) goto reject; /* :383-387 */
	rc->records_built++;
	return 1;
reject:
	rec_free(r);
	return 0;
}

/* ------------------------------------------------------------------------------------------ */
/* SubjectRead record list                                                                    */
/* ------------------------------------------------------------------------------------------ */
typedef struct {
	rec_t *v;
	int n, cap;
	uint64_t largest; /* iLargestQueryId */
} reclist;

static void reclist_push(reclist *l, rec_t *r) {
	if (l->n == l->cap) {
		l->cap = l->cap ? 2 * l->cap : 32;
		l->v = (rec_t *) xrealloc(l->v, sizeof(rec_t) * (size_t) l->cap);
	}
	l->v[l->n++] = *r;
}
static void reclist_clear(reclist *l) {
	for (int i = 0; i < l->n; i++) rec_free(&l->v[i]);
	l->n = 0;
	l->largest = 0;
}

/* One "keep the better record" decision of SubjectRead::addAlignment: the existing record stays
 * when its rate sum is lower, or equal with alignment length >= the new one's. */
static void keep_better(rec_t *slot, rec_t *r) {
	double a = slot->irate + slot->mrate, b = r->irate + r->mrate;
	if (a < b) rec_free(r);
	else if (a == b) {
		if (slot->len >= r->len) rec_free(r);
		else rec_free(slot), *slot = *r;
	} else {
		rec_free(slot), *slot = *r;
	}
}

/* SubjectRead::addAlignment (src/SubjectRead.cpp:25-79), without -v. */
static void reclist_add(reclist *l, rec_t *r) {
	uint64_t id = (uint64_t) r->qid;
	if (l->n == 0) {
		l->largest = id;
		reclist_push(l, r);
		return;
	}
	if ((uint64_t) l->v[l->n - 1].qid == id) {
		keep_better(&l->v[l->n - 1], r);
	} else if (l->largest >= id) {
		for (int i = 0; i < l->n; i++) {
			if ((uint64_t) l->v[i].qid > id) break;
			if ((uint64_t) l->v[i].qid == id) {
				keep_better(&l->v[i], r);
				return;
			}
		}
		reclist_push(l, r);
	} else {
		l->largest = id;
		reclist_push(l, r);
	}
}

/* AlignmentRecord::isSubjectCoveredByQuery (src/AlignmentRecord.cpp:343-370): the query starts at
 * the subject's first base, neither end op is a deletion, and the non-insertion ops span at least
 * the subject. */
static int rec_covers_subject(const rec_t *r, int64_t Ls) {
	if (r->pos != 0 || r->n == 0) return 0;
	int64_t len = 0;
	if (r->typ[0] == 'D') return 0;
	if (r->typ[0] != 'I') len += r->cnt[0];
	if (r->n > 1) {
		for (int i = 1; i < r->n - 1; i++)
			if (r->typ[i] != 'I') len += r->cnt[i];
		if (r->typ[r->n - 1] == 'D') return 0;
		if (r->typ[r->n - 1] != 'I') len += r->cnt[r->n - 1];
	}
	return len >= Ls;
}
/* AlignmentRecord::getAccuracy (:372-381): substituted bases (ops other than I, D, M). */
static double rec_substitutions(const rec_t *r) {
	double s = 0;
	for (int i = 0; i < r->n; i++)
		if (r->typ[i] != 'I' && r->typ[i] != 'D' && r->typ[i] != 'M') s += r->cnt[i];
	return s;
}
/* SubjectRead::chooseQueryWithBestAlignment (src/SubjectRead.cpp:463-499), --mapping: no record
 * covers the subject -> none votes; several -> only the first with the fewest substitutions
 * votes; exactly one -> every record stays (the reference keeps the list as it is). */
static void choose_best(reclist *l, int64_t Ls) {
	if (l->n <= 0) return;
	int nq = 0, best = -1;
	double least = (double) Ls;
	for (int i = 0; i < l->n; i++) {
		if (!rec_covers_subject(&l->v[i], Ls)) continue;
		nq++;
		const double s = rec_substitutions(&l->v[i]);
		if (s < least) least = s, best = i;
	}
	if (nq == 0 || (nq > 1 && best < 0)) {
		reclist_clear(l);
		return;
	}
	if (nq == 1) return;
	for (int i = 0; i < l->n; i++)
		if (i != best) rec_free(&l->v[i]);
	l->v[0] = l->v[best];
	l->n = 1;
}

/* ------------------------------------------------------------------------------------------ */
/* Chaining (src/SubjectReadAligner.cpp:391-408,711-751,805-852)                               */
/* ------------------------------------------------------------------------------------------ */
/* isFromTheSameAlignment: strictly after the chain's last key on both reads, and the diagonal
 * drift from the chain's first key within the indel rate (IEEE double, as the reference). */
static int same_alignment(const oracle_cfg *c, const seed_t *add, const seed_t *begin, const seed_t *end) {
	if (add->ps <= end->ps) return 0;
	if (add->pq <= end->pq) return 0;
	double lq = fabs((double) (add->pq - begin->pq));
	double ls = fabs((double) (add->ps - begin->ps));
	double total = lq >= ls ? lq : ls;
	total += (double) c->k;
	double diff = fabs(lq - ls);
	if ((diff / total) > c->indel_rate) return 0;
	return 1;
}

/* calKeyLength: covered subject length of a chain. */
static int64_t chain_key_length(const oracle_cfg *c, const seed_t *keys, const int *idx, int n) {
	int64_t k = c->k;
	if (n == 1) return k;
	int64_t len = 0;
	for (int i = 0; i + 1 < n; i++) {
		int64_t a = keys[idx[i]].ps, b = keys[idx[i + 1]].ps;
		len += (a + k <= b) ? k : (b - a);
	}
	return len + k;
}

typedef struct {
	int *idx; /* seed indices of every chain, chain-major with stride = group size */
	int *len;
	int cap;
	seed_t *tmp;
	int tmp_cap;
} chain_ws;

static void chain_ws_free(chain_ws *cw) { free(cw->idx), free(cw->len), free(cw->tmp), memset(cw, 0, sizeof *cw); }

/* processOneSubjectRead2 for one strand: greedy first-fit chaining per query group, then every
 * chain whose covered length equals the group maximum (ties kept, markTheBestPossibleAlignment)
 * goes through getAlignment2 in chain order; accepted records enter the subject's list. */
static void process_strand(const oracle_cfg *c, const oracle_reads *q, const char *subj, int64_t Ls, int orient,
		const seedvec *sv, dp_ws *w, chain_ws *cw, run_counters *rc, reclist *recs, char *qrc_buf) {
	int64_t n = sv->n;
	int64_t groups = 0;
	for (int64_t g0 = 0; g0 < n;) {
		int64_t g1 = g0 + 1;
		while (g1 < n && sv->v[g1].qid == sv->v[g0].qid) g1++;
		groups++;
		if (c->max_candidates > 0 && groups > c->max_candidates) break;
		int gn = (int) (g1 - g0);
		if (gn * gn > cw->cap) {
			cw->cap = gn * gn;
			cw->idx = (int *) xrealloc(cw->idx, sizeof(int) * (size_t) cw->cap);
		}
		if (gn > cw->tmp_cap) {
			cw->tmp_cap = gn;
			cw->len = (int *) xrealloc(cw->len, sizeof(int) * (size_t) gn);
			cw->tmp = (seed_t *) xrealloc(cw->tmp, sizeof(seed_t) * (size_t) gn);
		}
		const seed_t *s = sv->v + g0;
		int nch = 0;
		for (int i = 0; i < gn; i++) {
			int placed = 0;
			for (int j = 0; j < nch; j++) {
				int *ch = cw->idx + (size_t) j * gn;
				if (same_alignment(c, &s[i], &s[ch[0]], &s[ch[cw->len[j] - 1]])) {
					ch[cw->len[j]++] = i;
					placed = 1;
					break;
				}
			}
			if (!placed) {
				cw->idx[(size_t) nch * gn] = i;
				cw->len[nch++] = 1;
			}
		}
		int64_t best = 0;
		int64_t *kl = (int64_t *) xmalloc(sizeof(int64_t) * (size_t) nch);
		for (int j = 0; j < nch; j++) {
			kl[j] = chain_key_length(c, s, cw->idx + (size_t) j * gn, cw->len[j]);
			if (kl[j] > best) best = kl[j];
		}
		int64_t qid = s[0].qid;
		const char *qs = qseq(q, qid);
		int64_t Lq = q->length[qid];
		if (!c->call_consensus && orient == 1) { /* query RC for reverse strand without -c (:207-208) */
			reverse_complement(qs, Lq, qrc_buf);
			qs = qrc_buf;
		}
		int64_t rec_id = c->call_consensus ? qid : q->unique_id[qid];
		for (int j = 0; j < nch; j++) {
			if (kl[j] != best) continue;
			for (int t = 0; t < cw->len[j]; t++) cw->tmp[t] = s[cw->idx[(size_t) j * gn + t]];
			rec_t r;
			if (get_alignment(c, subj, Ls, qs, Lq, rec_id, orient, cw->tmp, cw->len[j], w, rc, &r)) {
				if (qs == qrc_buf) {
					/* the reference keeps a pointer to the query's stored RC; keep our own copy */
					char *cp = (char *) xmalloc((size_t) Lq);
					memcpy(cp, qs, (size_t) Lq);
					r.qseq = cp;
					r.orient |= 2; /* owns qseq */
				}
				reclist_add(recs, &r);
			}
		}
		free(kl);
		g0 = g1;
	}
}

/* ------------------------------------------------------------------------------------------ */
/* Votes and consensus                                                                        */
/* ------------------------------------------------------------------------------------------ */
/* SubjectRead::updateQueryCoverage (src/SubjectRead.cpp:208-330) with
 * QueryRead::updateCoveragematrix (src/QueryRead.cpp:124-146): walk the CIGAR from
 * (pos on subject, 0 on query); M adds one vote for the query's own base, a mismatch op adds one
 * vote per position for the subject's base (subject RC for reverse records), I/D skip. */
static void vote_record(const oracle_reads *q, uint32_t *tally, const rec_t *r, const char *subj_strand) {
	int64_t qid = r->qid;
	uint32_t *t = tally + 4 * q->offset[qid];
	const char *qs = qseq(q, qid);
	int64_t ps = r->pos, pq = 0;
	for (int i = 0; i < r->n; i++) {
		char ty = r->typ[i];
		int32_t n = r->cnt[i];
		if (ty == 'I') pq += n;
		else if (ty == 'D') ps += n;
		else if (ty == 'M') {
			for (int32_t x = 0; x < n; x++) {
				uint32_t *p = &t[4 * (pq + x) + code_of(qs[pq + x])];
#pragma omp atomic
				(*p)++;
			}
			pq += n, ps += n;
		} else {
			for (int32_t x = 0; x < n; x++) {
				uint32_t *p = &t[4 * pq + code_of(subj_strand[ps])];
#pragma omp atomic
				(*p)++;
				pq++, ps++;
			}
		}
	}
}

/* QueryRead::callConsensus (src/QueryRead.cpp:148-180) with aDNA = {A,C,T,G} (:11). */
static const char ADNA[4] = {'A', 'C', 'T', 'G'};

int oracle_consensus_read(const oracle_cfg *c, char *seq, int32_t len, uint32_t *t) {
	int changed = 0;
	for (int32_t i = 0; i < len; i++) t[4 * i + code_of(seq[i])]++; /* self vote (:149) */
	for (int32_t i = 0; i < len; i++) {
		int id = -1;
		double largest = 0, larger = 0, sum = 0, f;
		for (int j = 0; j < 4; j++) sum += t[4 * i + j];
		for (int j = 0; j < 4; j++) {
			f = ((double) t[4 * i + j]) / sum;
			if (largest < f) {
				larger = largest;
				largest = f;
				id = j;
			}
		}
		if (largest - larger >= c->consensus_freq) {
			if (seq[i] != ADNA[id]) changed++;
			seq[i] = ADNA[id];
		}
	}
	return changed;
}

/* ------------------------------------------------------------------------------------------ */
/* .aln writer (src/AlignmentRecord.cpp:102-149, SubjectRead.cpp:179-193, SubjectReadAligner.cpp:854-910) */
/* ------------------------------------------------------------------------------------------ */
typedef struct {
	char *s;
	size_t n, cap;
} strbuf;
static void sb_put(strbuf *b, const char *s, size_t n) {
	if (b->n + n + 1 > b->cap) {
		b->cap = (b->n + n + 1) * 2;
		b->s = (char *) xrealloc(b->s, b->cap);
	}
	memcpy(b->s + b->n, s, n);
	b->n += n;
	b->s[b->n] = 0;
}
static void sb_printf(strbuf *b, const char *fmt, long long v) {
	char tmp[64];
	int m = snprintf(tmp, sizeof tmp, fmt, v);
	sb_put(b, tmp, (size_t) m);
}

static void aln_record(strbuf *b, const rec_t *r) {
	sb_printf(b, "%lld,", (long long) r->qid);
	sb_printf(b, "%lld,", (long long) (r->orient & 1));
	sb_printf(b, "%lld,", (long long) r->pos);
	for (int i = 0; i < r->n; i++) {
		sb_printf(b, "%lld", (long long) r->cnt[i]);
		sb_put(b, &r->typ[i], 1);
	}
	sb_put(b, ",", 1);
	size_t index = 0;
	for (int i = 0; i < r->n; i++) {
		if (r->typ[i] == 'I') {
			for (int j = 0; j < r->cnt[i]; j++) {
				if (i != 0 && i != r->n - 1) sb_put(b, &r->qseq[index], 1);
				index++;
			}
			continue;
		}
		if (r->typ[i] == 'D') continue;
		index += (size_t) r->cnt[i];
	}
	sb_put(b, ",", 1);
}

/* ------------------------------------------------------------------------------------------ */
/* Driver (src/main.cpp:40-106, src/SubjectReadAligner.cpp:493-645)                            */
/* ------------------------------------------------------------------------------------------ */
int64_t oracle_seeds(const oracle_cfg *cfg, const oracle_reads *queries, const char *subject, int32_t len,
		int64_t subject_unique_id, int reverse, int64_t *out, int64_t cap) {
	index_t ix;
	if (index_build(cfg, queries, &ix) != 0) return -2;
	char *rc = (char *) xmalloc((size_t) len + 1);
	const char *seq = subject;
	if (reverse) reverse_complement(subject, len, rc), seq = rc;
	seedvec sv = {0};
	uint64_t lookups = 0;
	search_seeds(cfg, &ix, seq, len, subject_unique_id, reverse, &sv, &lookups);
	int64_t n = sv.n;
	if (n <= cap)
		for (int64_t i = 0; i < n; i++) out[3 * i] = sv.v[i].qid, out[3 * i + 1] = sv.v[i].ps, out[3 * i + 2] = sv.v[i].pq;
	free(sv.v);
	free(rc);
	index_free(&ix);
	return n <= cap ? n : -1;
}

int oracle_run(const oracle_cfg *cfg, const oracle_reads *q, const oracle_reads *s, char *corrected, uint32_t *tally,
		char **aln, char **records, int threads, oracle_stats *st) {
	oracle_stats local;
	if (!st) st = &local;
	memset(st, 0, sizeof *st);
	if (threads > 0) omp_set_num_threads(threads);
	if (cfg->k <= 0 || cfg->k > 1000) return -1;
	int64_t qbases = 0;
	for (int64_t i = 0; i < q->n; i++) qbases = imax64(qbases, q->offset[i] + q->length[i]);
	int own_tally = 0;
	if (cfg->call_consensus && !tally) {
		tally = (uint32_t *) xmalloc(sizeof(uint32_t) * 4 * (size_t) qbases);
		own_tally = 1;
	}
	if (tally) memset(tally, 0, sizeof(uint32_t) * 4 * (size_t) qbases);

	double t0 = omp_get_wtime();
	index_t ix;
	int rcode = index_build(cfg, q, &ix);
	if (rcode != 0) {
		if (own_tally) free(tally);
		return rcode;
	}
	st->table_size = ix.P;
	st->distinct_keys = (uint64_t) ix.nbuckets;
	st->total_keys = ix.total_keys;
	st->max_probe = ix.max_probe;
	double t1 = omp_get_wtime();
	st->t_index = t1 - t0;

	int64_t ns = s->n;
	char **aln_parts = (!cfg->call_consensus && aln) ? (char **) calloc((size_t) ns, sizeof(char *)) : NULL;
	char **rec_parts = records ? (char **) calloc((size_t) ns, sizeof(char *)) : NULL;
	uint64_t lookups = 0, nseeds = 0, dpc = 0, dpcells = 0, built = 0, kept = 0;
	int64_t maxlen = 0;
	for (int64_t i = 0; i < q->n; i++) maxlen = imax64(maxlen, q->length[i]);
	for (int64_t i = 0; i < ns; i++) maxlen = imax64(maxlen, s->length[i]);

#pragma omp parallel reduction(+ : lookups, nseeds, dpc, dpcells, built, kept)
	{
		dp_ws w;
		memset(&w, 0, sizeof w);
		chain_ws cw;
		memset(&cw, 0, sizeof cw);
		seedvec sv = {0};
		reclist recs = {0};
		char *rcbuf = (char *) xmalloc((size_t) maxlen + 1);
		char *qrcbuf = (char *) xmalloc((size_t) maxlen + 1);
		run_counters rc = {0, 0, 0};
#pragma omp for schedule(dynamic, 16)
		for (int64_t si = 0; si < ns; si++) {
			const char *fw = s->seq + s->offset[si];
			int64_t Ls = s->length[si];
			strbuf sb = {0};
			strbuf rb = {0};
			if (!subject_is_good(cfg, fw, Ls)) { /* :526-528 */
				if (aln_parts) sb_printf(&sb, ">%lld\n\n", (long long) s->unique_id[si]), aln_parts[si] = sb.s;
				continue;
			}
			reverse_complement(fw, Ls, rcbuf);
			recs.n = 0;
			recs.largest = 0;
			for (int orient = 0; orient < 2; orient++) {
				const char *strand = orient ? rcbuf : fw;
				search_seeds(cfg, &ix, strand, Ls, s->unique_id[si], orient, &sv, &lookups);
				nseeds += (uint64_t) sv.n;
				if (sv.n == 0) continue;
				/* -c aligns the reverse strand against the subject's RC; otherwise against the forward
				 * subject with the query's RC (:807-812) */
				const char *subj = (cfg->call_consensus && orient) ? rcbuf : fw;
				process_strand(cfg, q, subj, Ls, orient, &sv, &w, &cw, &rc, &recs, qrcbuf);
			}
			kept += (uint64_t) recs.n;
			if (cfg->call_consensus && cfg->mapping) choose_best(&recs, Ls); /* :557-560 */
			if (cfg->call_consensus) {
				for (int r = 0; r < recs.n; r++) vote_record(q, tally, &recs.v[r], (recs.v[r].orient & 1) ? rcbuf : fw);
			} else if (aln_parts) {
				sb_printf(&sb, ">%lld\n", (long long) s->unique_id[si]);
				for (int r = 0; r < recs.n; r++) aln_record(&sb, &recs.v[r]);
				sb_put(&sb, "\n", 1);
				aln_parts[si] = sb.s;
			}
			if (rec_parts) {
				for (int r = 0; r < recs.n; r++) {
					const rec_t *x = &recs.v[r];
					sb_printf(&rb, "%lld ", (long long) si);
					sb_printf(&rb, "%lld ", (long long) (x->orient & 1));
					sb_printf(&rb, "%lld ", (long long) x->qid);
					sb_printf(&rb, "%lld ", (long long) x->pos);
					for (int t = 0; t < x->n; t++) {
						sb_printf(&rb, "%lld", (long long) x->cnt[t]);
						sb_put(&rb, &x->typ[t], 1);
					}
					sb_put(&rb, "\n", 1);
				}
				rec_parts[si] = rb.s;
			}
			reclist_clear(&recs);
		}
		dpc += rc.dp_calls;
		dpcells += rc.dp_cells;
		built += rc.records_built;
		dp_ws_free(&w);
		chain_ws_free(&cw);
		free(sv.v);
		free(recs.v);
		free(rcbuf);
		free(qrcbuf);
	}
	st->lookups = lookups, st->seeds = nseeds, st->dp_calls = dpc, st->dp_cells = dpcells;
	st->records_built = built, st->records_kept = kept;
	double t2 = omp_get_wtime();
	st->t_align = t2 - t1;
	index_free(&ix);

	if (aln_parts) {
		strbuf all = {0};
		for (int64_t i = 0; i < ns; i++)
			if (aln_parts[i]) sb_put(&all, aln_parts[i], strlen(aln_parts[i])), free(aln_parts[i]);
		free(aln_parts);
		*aln = all.s ? all.s : (char *) calloc(1, 1);
	}
	if (rec_parts) {
		strbuf all = {0};
		for (int64_t i = 0; i < ns; i++)
			if (rec_parts[i]) sb_put(&all, rec_parts[i], strlen(rec_parts[i])), free(rec_parts[i]);
		free(rec_parts);
		*records = all.s ? all.s : (char *) calloc(1, 1);
	}

	/* QueryDataset::callConsensus (src/QueryDataset.cpp:393-412): good reads only; the others are
	 * written unchanged (isGoodRead() returns true, src/QueryRead.cpp:214-215). */
	uint64_t changed = 0;
	if (cfg->call_consensus && !cfg->mapping && corrected) { /* --mapping stops before consensus (main.cpp:82-87) */
		memcpy(corrected, q->seq, (size_t) qbases);
		uint32_t *tw = tally;
		uint32_t *scratch = NULL;
		if (!own_tally) { /* keep the caller's tally as the pre-consensus vote state */
			scratch = (uint32_t *) xmalloc(sizeof(uint32_t) * 4 * (size_t) qbases);
			memcpy(scratch, tally, sizeof(uint32_t) * 4 * (size_t) qbases);
			tw = scratch;
		}
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : changed)
		for (int64_t i = 0; i < q->n; i++) {
			const char *src = q->seq + q->offset[i];
			if (!query_is_good(cfg, src, q->length[i])) continue;
			char *dst = corrected + q->offset[i];
			if (oracle_consensus_read(cfg, dst, q->length[i], tw + 4 * q->offset[i]) > 0) changed++;
		}
		free(scratch);
	}
	st->reads_changed = changed;
	st->t_consensus = omp_get_wtime() - t2;
	if (own_tally) free(tally);
	return 0;
}

/* QueryDataset::getCoverageAndWriteResults (src/QueryDataset.cpp:560-600), --mapping: a base's
 * coverage is the sum of its four counters (QueryRead::getCoverageAt, src/QueryRead.cpp:435-442),
 * over every base of every query read; the text is "COV\tFRE" and one line per coverage value that
 * some base has, ascending. Returns malloc'd text (free with oracle_free). */
char *oracle_coverage_text(const oracle_reads *q, const uint32_t *tally) {
	uint32_t maxc = 0;
	for (int64_t i = 0; i < q->n; i++)
		for (int32_t j = 0; j < q->length[i]; j++) {
			const uint32_t *t = tally + 4 * (q->offset[i] + j);
			const uint32_t c = t[0] + t[1] + t[2] + t[3];
			if (c > maxc) maxc = c;
		}
	maxc++;
	uint64_t *h = (uint64_t *) calloc((size_t) maxc, sizeof(uint64_t));
	for (int64_t i = 0; i < q->n; i++)
		for (int32_t j = 0; j < q->length[i]; j++) {
			const uint32_t *t = tally + 4 * (q->offset[i] + j);
			h[t[0] + t[1] + t[2] + t[3]]++;
		}
	strbuf sb = {0};
	sb_put(&sb, "COV\tFRE\n", 8);
	for (uint32_t c = 0; c < maxc; c++) {
		if (!h[c]) continue;
		sb_printf(&sb, "%lld\t", (long long) c);
		sb_printf(&sb, "%lld\n", (long long) (uint32_t) h[c]);
	}
	free(h);
	return sb.s;
}
