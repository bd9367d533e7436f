// HIP kernels of the correction path for gfx950 (CDNA4): seed index build, seed search and
// sort, chaining and planning (pass A), banded DP (batched lane-per-pair and plain
// wave-per-pair forms), record scoring / dedup / votes (pass B) and consensus.
// Every kernel that works per subject runs one 64-lane wavefront per subject read.
#include <hip/hip_runtime.h>

#include <cstring>
#include <map>
#include <mutex>

#include "dp_pk16.h"
#include "kernels.h"
#include "pipeline_core.h"

namespace hero {


__device__ __forceinline__ int lane_id() { return (int) threadIdx.x & 63; }

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
	for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
	return v;
}
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
#pragma unroll
	for (int off = 32; off > 0; off >>= 1) {
		const uint32_t o = __shfl_xor(v, off, 64);
		v = o > v ? o : v;
	}
	return v;
}
__device__ __forceinline__ uint32_t wave_excl_scan(uint32_t v) {
	const int lane = lane_id();
	uint32_t x = v;
#pragma unroll
	for (int off = 1; off < 64; off <<= 1) {
		const uint32_t t = __shfl_up(x, off, 64);
		if (lane >= off) x += t;
	}
	return x - v;
}
__device__ __forceinline__ uint64_t wave_sum64(uint64_t v) {
	uint32_t lo = (uint32_t) v, hi = (uint32_t) (v >> 32);
#pragma unroll
	for (int off = 32; off > 0; off >>= 1) {
		const uint32_t l2 = __shfl_xor(lo, off, 64), h2 = __shfl_xor(hi, off, 64);
		const uint32_t s = lo + l2;
		hi = hi + h2 + (s < lo ? 1u : 0u);
		lo = s;
	}
	return ((uint64_t) hi << 32) | lo;
}
// Adds v to counter `slot` in this workgroup's stripe.
__device__ __forceinline__ void stat_add(unsigned long long *stats, int slot, unsigned long long v) {
	atomicAdd(&stats[(blockIdx.x % kStatStripes) * kStatStride + slot], v);
}

__device__ __forceinline__ uint64_t lanes_below_mask() {
	const int lane = lane_id();
	return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// ---------------------------------------------------------------------------------------------
// Seed index (roles of HashTableLongKey::createHashTables / insertQueryDataset / getListOfReads,
// src/HashTableLongKey.cpp:41-99,250-342). Slots are keyed by canonical k-mers (Slot); the home
// slot is the reference hash of the canonical form; linear probing. Each half of a bucket holds
// exactly the occurrences of one k-mer, so lookups return what getListOfReads returns, whatever
// the slot placement.
// ---------------------------------------------------------------------------------------------
// Build-time occurrence rank: bit 31 marks an occurrence of the reverse complement of its slot's
// canonical k-mer (set by k_index_gen), the low bits are its rank in its half (k_index_count).
constexpr uint32_t kOccRc = 0x80000000u;
__global__ void k_index_gen(Packed Q, const uint64_t *koff, int k, int all_positions, uint64_t *occ_hi, uint64_t *occ_lo,
		uint64_t *occ_entry, uint32_t *occ_rank) {
	const int64_t q = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
	if (q >= Q.n || !Q.good[q]) return;
	const int L = Q.len[q];
	const uint64_t *w = Q.words + Q.woff[q];
	uint64_t o = koff[q];
	const int nk = all_positions ? (L >= k ? L - k + 1 : 0) : num_keys(L, k);
	for (int t = 0; t < nk; t++) {
		const int p = all_positions ? t : key_pos(L, k, t);
		uint64_t hi, lo, chi, clo;
		kmer_fwd(w, p, k, hi, lo);
		const bool fwd = kmer_canonical(k, hi, lo, chi, clo);
		occ_hi[o] = chi;
		occ_lo[o] = clo;
		occ_entry[o] = ((uint64_t) uid_tag(Q.uid[q]) << 54) | ((uint64_t) p << 40) | (uint64_t) q;
		occ_rank[o] = fwd ? 0u : kOccRc;
		o++;
	}
}
hipError_t launch_index_gen(const Packed &Q, const uint64_t *koff, int k, bool all_positions, uint64_t *occ_hi,
		uint64_t *occ_lo, uint64_t *occ_entry, uint32_t *occ_rank, hipStream_t st) {
	if (Q.n == 0) return hipSuccess;
	k_index_gen<<<(unsigned) ((Q.n + 255) / 256), 256, 0, st>>>(Q, koff, k, all_positions ? 1 : 0, occ_hi, occ_lo,
			occ_entry, occ_rank);
	return hipGetLastError();
}

// One occurrence per thread: claim the first free slot from the home slot with a CAS of the
// occurrence id, or stop at the slot already claimed by an occurrence of the same k-mer (whose
// key is readable in occ_* since they were written by the previous kernel).
__global__ void k_index_claim(const uint64_t *occ_hi, const uint64_t *occ_lo, uint64_t nocc, uint64_t P, int k,
		uint32_t *owner, uint32_t *occ_slot, uint32_t *max_probe, int *err) {
	const uint64_t o = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
	uint32_t longest = 0; // probes past the home slot (most claims take none)
	if (o < nocc) {
		const uint64_t hi = occ_hi[o], lo = occ_lo[o];
		uint64_t h = hash_kmer(k, hi, lo, make_fastmod(P));
		bool placed = false;
		for (uint64_t probe = 0; probe < P; probe++) {
			const uint32_t cur = atomicCAS(&owner[h], 0u, (uint32_t) (o + 1));
			bool done = cur == 0;
			if (!done) {
				const uint64_t other = cur - 1;
				done = occ_hi[other] == hi && (k <= 32 || occ_lo[other] == lo);
			}
			if (done) {
				occ_slot[o] = (uint32_t) h;
				longest = (uint32_t) probe;
				placed = true;
				break;
			}
			h = (h + 1 == P) ? 0 : h + 1;
		}
		if (!placed) *err = 1; // table full ("The size of hash table is too small.")
	}
	// one atomic per wavefront, and none when no lane probed past its home slot: every thread's
	// atomic on the one address serialized the kernel
	longest = wave_max_u32(longest);
	if (lane_id() == 0 && longest) atomicMax(max_probe, longest);
}
hipError_t launch_index_claim(const uint64_t *occ_hi, const uint64_t *occ_lo, uint64_t nocc, uint64_t P, int k,
		uint32_t *owner, uint32_t *occ_slot, uint32_t *max_probe, int *err, hipStream_t st) {
	if (nocc == 0) return hipSuccess;
	k_index_claim<<<(unsigned) ((nocc + 255) / 256), 256, 0, st>>>(occ_hi, occ_lo, nocc, P, k, owner, occ_slot, max_probe, err);
	return hipGetLastError();
}

__global__ void k_index_count(const uint32_t *occ_slot, uint64_t nocc, Slot *slots, uint32_t *occ_rank) {
	const uint64_t o = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
	if (o >= nocc) return;
	const uint32_t rc = occ_rank[o] & kOccRc;
	Slot &s = slots[occ_slot[o]];
	occ_rank[o] = rc | atomicAdd(rc ? &s.nrc : &s.nfwd, 1u);
}
hipError_t launch_index_count(const uint32_t *occ_slot, uint64_t nocc, Slot *slots, uint32_t *occ_rank, hipStream_t st) {
	if (nocc == 0) return hipSuccess;
	k_index_count<<<(unsigned) ((nocc + 255) / 256), 256, 0, st>>>(occ_slot, nocc, slots, occ_rank);
	return hipGetLastError();
}

// Bucket ranges: one cursor atomic per block of kAllocPerThread * 256 slots (the slots of a block
// get consecutive ranges; where buckets lie in `entries` is invisible to lookups).
constexpr int kAllocPerThread = 16;
// cursor[0]: the next free entry; cursor[1] += the reference's distinct keys (each half of a
// canonical slot that holds occurrences is one of the reference's keys).
__global__ __launch_bounds__(256) void k_index_alloc(const uint32_t *owner, uint64_t P, const uint64_t *occ_hi,
		const uint64_t *occ_lo, int k, Slot *slots, unsigned long long *cursor) {
	__shared__ uint32_t wave_tot[4];
	__shared__ uint32_t block_base;
	const uint64_t base = (uint64_t) blockIdx.x * 256 * kAllocPerThread + threadIdx.x;
	uint32_t cnt = 0, keys = 0;
	for (int i = 0; i < kAllocPerThread; i++) {
		const uint64_t s = base + (uint64_t) i * 256;
		if (s < P && owner[s]) {
			const uint32_t nf = slots[s].nfwd, nr = slots[s].nrc;
			cnt += nf + nr;
			keys += (nf ? 1u : 0u) + (nr ? 1u : 0u);
		}
	}
	keys = wave_sum(keys);
	if (lane_id() == 0 && keys) atomicAdd(&cursor[1], (unsigned long long) keys);
	const uint32_t pre = wave_excl_scan(cnt);
	const int wv = threadIdx.x >> 6;
	if (lane_id() == 63) wave_tot[wv] = pre + cnt;
	__syncthreads();
	if (threadIdx.x == 0) {
		uint32_t t = 0;
		for (int i = 0; i < 4; i++) {
			const uint32_t v = wave_tot[i];
			wave_tot[i] = t;
			t += v;
		}
		block_base = t ? (uint32_t) atomicAdd(cursor, (unsigned long long) t) : 0u;
	}
	__syncthreads();
	uint32_t b = block_base + wave_tot[wv] + pre;
	for (int i = 0; i < kAllocPerThread; i++) {
		const uint64_t s = base + (uint64_t) i * 256;
		if (s >= P) break;
		const uint32_t own = owner[s];
		if (!own) continue;
		const uint64_t o = own - 1;
		Slot sl = slots[s];
		sl.key = occ_hi[o];
		sl.key_lo = k > 32 ? occ_lo[o] : 0;
		sl.start = b;
		sl.count = sl.nfwd + sl.nrc;
		b += sl.count;
		slots[s] = sl;
	}
}
hipError_t launch_index_alloc(const uint32_t *owner, uint64_t P, const uint64_t *occ_hi, const uint64_t *occ_lo, int k,
		Slot *slots, unsigned long long *cursor, hipStream_t st) {
	const uint64_t per_block = 256ull * kAllocPerThread;
	k_index_alloc<<<(unsigned) ((P + per_block - 1) / per_block), 256, 0, st>>>(owner, P, occ_hi, occ_lo, k, slots,
			cursor);
	return hipGetLastError();
}

__global__ void k_index_fill(const uint32_t *occ_slot, const uint32_t *occ_rank, const uint64_t *occ_entry, uint64_t nocc,
		const Slot *slots, uint64_t *entries) {
	const uint64_t o = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
	if (o >= nocc) return;
	const uint32_t r = occ_rank[o];
	const Slot &s = slots[occ_slot[o]];
	entries[(uint64_t) s.start + ((r & kOccRc) ? s.nfwd : 0u) + (r & ~kOccRc)] = occ_entry[o];
}
hipError_t launch_index_fill(const uint32_t *occ_slot, const uint32_t *occ_rank, const uint64_t *occ_entry, uint64_t nocc,
		const Slot *slots, uint64_t *entries, hipStream_t st) {
	if (nocc == 0) return hipSuccess;
	k_index_fill<<<(unsigned) ((nocc + 255) / 256), 256, 0, st>>>(occ_slot, occ_rank, occ_entry, nocc, slots, entries);
	return hipGetLastError();
}

// Puts each half bucket in the reference's insertion order (innerId, then position). Seeds are
// re-sorted by (innerId, posQ, posS) downstream, so this only makes the index deterministic.
__global__ void k_index_sort_buckets(const Slot *slots, uint64_t P, uint64_t *entries) {
	const uint64_t s = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
	if (s >= P) return;
	const Slot sl = slots[s];
	auto key = [](uint64_t v) { return ((v & kIdMask) << 14) | (uint64_t) entry_pos(v); };
	for (int half = 0; half < 2; half++) {
		const uint32_t n = half ? sl.count - sl.nfwd : sl.nfwd;
		if (n < 2 || n > 512) continue;
		uint64_t *e = entries + sl.start + (half ? sl.nfwd : 0u);
		for (uint32_t i = 1; i < n; i++) {
			const uint64_t v = e[i], kv = key(v);
			int j = (int) i - 1;
			while (j >= 0 && key(e[j]) > kv) {
				e[j + 1] = e[j];
				j--;
			}
			e[j + 1] = v;
		}
	}
}
hipError_t launch_index_sort_buckets(const Slot *slots, uint64_t P, uint64_t *entries, hipStream_t st) {
	k_index_sort_buckets<<<(unsigned) ((P + 255) / 256), 256, 0, st>>>(slots, P, entries);
	return hipGetLastError();
}

// The k-mer of subject window w in strand frame (reverse strand: window of the RC).
__device__ __forceinline__ void window_key(const Packed &S, int64_t s, int Ls, int w, int strand, int k, uint64_t &hi,
		uint64_t &lo) {
	const uint64_t *ws = S.words + S.woff[s];
	if (strand == 0) {
		kmer_fwd(ws, w, k, hi, lo);
	} else {
		uint64_t fh, fl;
		kmer_fwd(ws, Ls - w - k, k, fh, fl);
		kmer_revcomp(k, fh, fl, hi, lo);
	}
}

// ---------------------------------------------------------------------------------------------
// Seed search (SubjectReadAligner::searchHashTable2, src/SubjectReadAligner.cpp:444-491)
// ---------------------------------------------------------------------------------------------
// Seeds of one window's bucket as they are emitted: skip hits on the query with the subject's
// uniqueId; without -c, reverse-strand coordinates map back to forward frames (:474-481).
template <class F>
__device__ __forceinline__ uint32_t bucket_seeds(const Packed &S, const Packed &Q, const DevIndex &ix, const Opts &o,
		int Ls, int64_t uid, int w, int strand, uint32_t start, uint32_t count, F &&emit) {
	(void) S;
	uint32_t n = 0;
	const uint32_t tag = uid_tag(uid);
	for (uint32_t e = start; e < start + count; e++) {
		const uint64_t data = ix.entries[e];
		const uint64_t qid = data & kIdMask;
		if (!o.long_format && entry_tag(data) == tag && Q.uid[qid] == uid) continue;
		int pq = entry_pos(data), ps = w;
		if (!o.call_consensus && strand == 1) ps = Ls - (ps + o.k), pq = Q.len[qid] - (pq + o.k);
		emit(seed_pack(qid, pq, ps));
		n++;
	}
	return n;
}

template <class F>
__device__ __forceinline__ uint32_t window_seeds(const Packed &S, const Packed &Q, const DevIndex &ix, const Opts &o,
		int64_t s, int Ls, int64_t uid, int w, int strand, F &&emit) {
	uint64_t hi, lo;
	window_key(S, s, Ls, w, strand, o.k, hi, lo);
	uint32_t start = 0, count = 0;
	if (!ix_lookup(ix, hi, lo, start, count)) return 0;
	return bucket_seeds(S, Q, ix, o, Ls, uid, w, strand, start, count, emit);
}

// Seed count per subject strand. With c.win_info set it also keeps every window's bucket and each
// lane's seed count, so pass A1 emits the seeds without hashing or probing again.
__global__ __launch_bounds__(64) void k_seed_count(Packed S, Packed Q, DevIndex ix, Opts o, ChunkDev c,
		unsigned long long *stats) {
	const int sl = (int) xcd_item(blockIdx.x, gridDim.x);
	const int64_t s = c.s_begin + sl;
	const int lane = lane_id();
	if (!S.good[s]) {
		if (lane == 0) c.seg_n[2 * sl] = 0, c.seg_n[2 * sl + 1] = 0;
		return;
	}
	const int Ls = S.len[s];
	const int64_t uid = S.uid[s];
	const int nwin = Ls - o.k + 1;
	const uint64_t *ws = S.words + S.woff[s];
	// Window w of the forward strand and window nwin - 1 - w of the reverse strand are the two
	// orientations of one k-mer: one probe of its canonical form finds both windows' buckets (the
	// reverse-strand window is the lane's too; pass A1 follows the same order).
	// A lane takes windows lane, lane + 64, lane + 128, ... two at a time: both keys are hashed and
	// both home slots loaded before either is used, so their loads overlap (the kernel waits on
	// these chains of dependent loads far more than on bandwidth).
	uint32_t cnt[2] = {0, 0};
	for (int w0 = lane; w0 < nwin; w0 += 128) {
		const int nw = w0 + 64 < nwin ? 2 : 1;
		uint64_t ch[2], cl[2], h[2];
		bool fwd[2], fwd_r[2];
#pragma unroll
		for (int u = 0; u < 2; u++) {
			const int w = w0 + 64 * (u < nw ? u : 0);
			uint64_t hi, lo, rhi, rlo;
			kmer_fwd(ws, w, o.k, hi, lo);
			kmer_revcomp(o.k, hi, lo, rhi, rlo); // the reverse strand's window nwin - 1 - w
			fwd[u] = hi < rhi || (hi == rhi && lo <= rlo);
			fwd_r[u] = !fwd[u] || (hi == rhi && lo == rlo); // the reverse window's test
			ch[u] = fwd[u] ? hi : rhi, cl[u] = fwd[u] ? lo : rlo;
			h[u] = hash_kmer(ix.k, ch[u], cl[u], ix.fm);
		}
		const Slot home[2] = {ix.slots[h[0]], ix.slots[h[1]]};
#pragma unroll
		for (int u = 0; u < 2; u++) {
			if (u >= nw) break;
			const int w = w0 + 64 * u;
			Slot sl_ix;
			const bool found = ix_probe_from(ix, ch[u], cl[u], h[u], home[u], sl_ix);
#pragma unroll
			for (int strand = 0; strand < 2; strand++) {
				const int wi = strand ? nwin - 1 - w : w;
				uint32_t start = 0, count = 0;
				if (found) ix_half(sl_ix, strand ? fwd_r[u] : fwd[u], start, count);
				if (c.win_info)
					c.win_info[(int64_t) (2 * sl + strand) * c.win_stride + wi] = count ? ((uint64_t) start << 32) | count : 0;
				if (count) cnt[strand] += bucket_seeds(S, Q, ix, o, Ls, uid, wi, strand, start, count, [](uint64_t) {});
			}
		}
	}
#pragma unroll
	for (int strand = 0; strand < 2; strand++) {
		const int slot = 2 * sl + strand;
		if (c.win_info) c.lane_n[(int64_t) slot * 64 + lane] = cnt[strand];
		const uint32_t t = wave_sum(cnt[strand]);
		if (lane == 0) c.seg_n[slot] = t;
	}
	if (lane == 0) stat_add(stats, kStLookups, (unsigned long long) (2 * nwin));
}
hipError_t launch_seed_count(const Packed &S, const Packed &Q, const DevIndex &ix, const Opts &o, const ChunkDev &c,
		unsigned long long *stats, hipStream_t st) {
	if (c.nsub == 0) return hipSuccess;
	k_seed_count<<<c.nsub, 64, 0, st>>>(S, Q, ix, o, c, stats);
	return hipGetLastError();
}

// Ascending bitonic network over n keys padded (virtually) with +inf to a power of two: every
// compare-exchange touching a pad element is a no-op, so skipping it sorts any n.
template <class A>
__device__ void bitonic_sort(A &a, int n) {
	int N = 1;
	while (N < n) N <<= 1;
	const int lane = lane_id();
	for (int kk = 2; kk <= N; kk <<= 1) {
		for (int i = lane; i < N / 2; i += 64) { // flip stage
			const int blk = i / (kk / 2), off = i % (kk / 2);
			const int x = blk * kk + off, y = blk * kk + kk - 1 - off;
			if (y < n) {
				const uint64_t u = a[x], v = a[y];
				if (u > v) a[x] = v, a[y] = u;
			}
		}
		__syncthreads();
		for (int j = kk / 4; j > 0; j >>= 1) {
			for (int i = lane; i < N / 2; i += 64) {
				const int blk = i / j, off = i % j;
				const int x = blk * 2 * j + off, y = x + j;
				if (y < n) {
					const uint64_t u = a[x], v = a[y];
					if (u > v) a[x] = v, a[y] = u;
				}
			}
			__syncthreads();
		}
	}
}
struct LdsArr {
	uint64_t *p;
	__device__ uint64_t &operator[](int i) const { return p[i]; }
};
struct GlobArr {
	uint64_t *p;
	__device__ uint64_t &operator[](int i) const { return p[i]; }
};

// Seeds of a subject strand sorted in LDS (more take the global-memory path). 4 KB per wavefront,
// and passes A1 / A2 / B are register-allocated for 8 / 7 / 8 waves per SIMD: these kernels wait on
// chains of dependent loads, so more waves in flight is what speeds them up (C2: a strand has 26
// seeds on average).
constexpr int kLdsSeeds = 512;

__device__ __forceinline__ PairReads make_pair(const Packed &S, const Packed &Q, const Opts &o, int64_t s, int Ls,
		int64_t q, int strand) {
	return make_pair_reads(S, Q, o.call_consensus != 0, s, Ls, q, strand);
}

// A group's seeds and chain arrays start at its first seed. Pass A1 chains it once and keeps the
// chain count and best covered length there; later passes reload them.
__device__ __forceinline__ Group make_group(const ChunkDev &c, uint32_t base, int gn) {
	Group g;
	g.seeds = c.seeds + base;
	g.cnext = c.cnext + base, g.cfirst = c.cfirst + base, g.clast = c.clast + base, g.clen = c.clen + base;
	g.gn = gn, g.nch = 0, g.best = 0;
	return g;
}
__device__ __forceinline__ Group chained_group(const ChunkDev &c, uint32_t base, int gn) {
	Group g = make_group(c, base, gn);
	g.nch = c.g_nch[base];
	g.best = c.g_best[base];
	return g;
}

// Pass A1: per subject strand, emit the seeds, sort them, chain every kept query group and count
// the DP tasks / edit words its surviving chains need.
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_pass_a1(Packed S, Packed Q, DevIndex ix, Opts o, ChunkDev c,
		unsigned long long *stats) {
	__shared__ uint64_t lds[kLdsSeeds];
	const int sl = (int) xcd_item(blockIdx.x, gridDim.x);
	const int64_t s = c.s_begin + sl;
	const int lane = lane_id();
	if (!S.good[s]) {
		if (c.sub_pairs && lane == 0) c.sub_pairs[sl] = 0;
		return;
	}
	const int Ls = S.len[s];
	const int64_t uid = S.uid[s];
	const int nwin = Ls - o.k + 1;
	uint32_t kept_groups = 0; // both strands: bounds the subject's (subject, query) pairs
	for (int strand = 0; strand < 2; strand++) {
		const uint32_t n = c.seg_n[2 * sl + strand], off = c.seg_off[2 * sl + strand];
		if (n == 0) continue;
		const bool in_lds = n <= (uint32_t) kLdsSeeds;
		uint64_t *dst = in_lds ? lds : c.seeds + off;
		if (c.win_info) {
			const int slot = 2 * sl + strand;
			uint32_t pos = wave_excl_scan(c.lane_n[(int64_t) slot * 64 + lane]);
			const uint64_t *wi = c.win_info + (int64_t) slot * c.win_stride;
			// the lane's windows as k_seed_count counted them: reverse-strand windows from the end
			for (int wl = lane; wl < nwin; wl += 64) {
				const int w = strand ? nwin - 1 - wl : wl;
				const uint64_t b = wi[w];
				if (b) bucket_seeds(S, Q, ix, o, Ls, uid, w, strand, (uint32_t) (b >> 32), (uint32_t) b,
						[&](uint64_t v) { dst[pos++] = v; });
			}
		} else {
			uint32_t cnt = 0;
			for (int w = lane; w < nwin; w += 64) cnt += window_seeds(S, Q, ix, o, s, Ls, uid, w, strand, [](uint64_t) {});
			uint32_t pos = wave_excl_scan(cnt);
			for (int w = lane; w < nwin; w += 64)
				window_seeds(S, Q, ix, o, s, Ls, uid, w, strand, [&](uint64_t v) { dst[pos++] = v; });
		}
		__syncthreads();
		if (in_lds) {
			LdsArr a{lds};
			bitonic_sort(a, (int) n);
			for (uint32_t t = lane; t < n; t += 64) c.seeds[off + t] = lds[t];
		} else {
			GlobArr a{c.seeds + off};
			bitonic_sort(a, (int) n);
		}
		__syncthreads();
		// the sorted seeds as this pass reads them: from LDS when they were sorted there (the global
		// copy is for the later passes)
		const uint64_t *sd = in_lds ? (const uint64_t *) lds : c.seeds + off;
		// groups: a seed starts a group when its query differs from the previous seed's
		uint32_t groups_before = 0;
		for (uint32_t t0 = 0; t0 < n; t0 += 64) {
			const uint32_t t = t0 + lane;
			bool start = false;
			if (t < n) start = t == 0 || seed_qid(sd[t]) != seed_qid(sd[t - 1]);
			const uint64_t m = __ballot(start);
			const uint32_t rank = groups_before + (uint32_t) __popcll(m & lanes_below_mask());
			groups_before += (uint32_t) __popcll(m);
			uint32_t ntask = 0, nword = 0, nrec = 0, nprog = 0;
			uint8_t flag = 0;
			if (start) {
				const bool kept = o.max_candidates <= 0 || rank < (uint32_t) o.max_candidates;
				flag = kept ? 3 : 1;
				if (kept) {
					const int64_t q = seed_qid(sd[t]);
					uint32_t e = t + 1;
					while (e < n && seed_qid(sd[e]) == q) e++;
					Group g = make_group(c, off + t, (int) (e - t));
					g.seeds = sd + t;
					chain(o, g);
					c.g_nch[off + t] = g.nch;
					c.g_best[off + t] = g.best;
					const PairReads pr = make_pair(S, Q, o, s, Ls, q, strand);
					int nt, nw, np, nsc;
					plan_count(o, g, pr, nt, nw, np, nsc);
					ntask = (uint32_t) nt, nword = (uint32_t) nw, nprog = (uint32_t) np, nrec = (uint32_t) nsc;
				}
			}
			if (t < n) {
				c.gflag[off + t] = flag;
				c.g_tasks[off + t] = ntask;
				c.g_words[off + t] = nword;
				c.g_prog[off + t] = nprog;
				if (c.g_recs) c.g_recs[off + t] = nrec;
			}
		}
		kept_groups += o.max_candidates > 0 ? min(groups_before, (uint32_t) o.max_candidates) : groups_before;
		__syncthreads(); // the next strand's seeds go to the same LDS
	}
	if (c.sub_pairs && lane == 0) c.sub_pairs[sl] = kept_groups;
	if (lane == 0) stat_add(stats, kStSeeds, (unsigned long long) (c.seg_n[2 * sl] + c.seg_n[2 * sl + 1]));
}
hipError_t launch_pass_a1(const Packed &S, const Packed &Q, const DevIndex &ix, const Opts &o, const ChunkDev &c,
		unsigned long long *stats, hipStream_t st) {
	if (c.nsub == 0) return hipSuccess;
	k_pass_a1<<<c.nsub, 64, 0, st>>>(S, Q, ix, o, c, stats);
	return hipGetLastError();
}

// Pass A2: write the DP tasks of every kept group at its scanned task / word offsets.
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(6, 8))) void k_pass_a2(Packed S, Packed Q, Opts o, ChunkDev c) {
	const int sl = (int) xcd_item(blockIdx.x, gridDim.x);
	const int64_t s = c.s_begin + sl;
	const int lane = lane_id();
	if (!S.good[s]) return;
	const int Ls = S.len[s];
	const uint32_t s_woff = (uint32_t) S.woff[s];
	for (int strand = 0; strand < 2; strand++) {
		const uint32_t n = c.seg_n[2 * sl + strand], off = c.seg_off[2 * sl + strand];
		for (uint32_t t = lane; t < n; t += 64) {
			if (c.gflag[off + t] != 3) continue;
			const int64_t q = seed_qid(c.seeds[off + t]);
			uint32_t e = t + 1;
			while (e < n && seed_qid(c.seeds[off + e]) == q) e++;
			const Group g = chained_group(c, off + t, (int) (e - t));
			const PairReads pr = make_pair(S, Q, o, s, Ls, q, strand);
			plan_emit(o, g, pr, c.tasks, c.g_task_off[off + t], c.g_word_off[off + t], c.prog + c.g_prog_off[off + t], s_woff,
					(uint32_t) Q.woff[q]);
		}
	}
}
hipError_t launch_pass_a2(const Packed &S, const Packed &Q, const Opts &o, const ChunkDev &c, hipStream_t st) {
	if (c.nsub == 0) return hipSuccess;
	k_pass_a2<<<c.nsub, 64, 0, st>>>(S, Q, o, c);
	return hipGetLastError();
}

// First seed of query q's group in a sorted segment, or -1.
__device__ __forceinline__ int find_group(const uint64_t *seeds, uint32_t n, int64_t q) {
	uint32_t lo = 0, hi = n;
	while (lo < hi) {
		const uint32_t mid = (lo + hi) >> 1;
		if (seed_qid(seeds[mid]) < q) lo = mid + 1;
		else hi = mid;
	}
	return (lo < n && seed_qid(seeds[lo]) == q) ? (int) lo : -1;
}
__device__ __forceinline__ uint32_t group_end(const uint64_t *seeds, uint32_t n, uint32_t t) {
	const int64_t q = seed_qid(seeds[t]);
	uint32_t e = t + 1;
	while (e < n && seed_qid(seeds[e]) == q) e++;
	return e;
}

struct PassBCounters {
	uint64_t cells = 0, calls = 0, built = 0, kept = 0;
};

// One query of a subject: its forward group (seed index gf of segment 0, or -1) and reverse group
// (gr of segment 1, or -1). Records are built from the DP results in the reference's order
// (forward chains, then reverse chains), the first best is kept (SubjectRead::addAlignment), and
// that record casts its votes.
template <bool RECORDS>
__device__ __forceinline__ void pass_b_pair(const Packed &S, const Packed &Q, const Opts &o, const ChunkDev &c,
		const DpResults &R, int64_t s, int Ls, int64_t q, uint32_t base0, int gn0, uint32_t base1, int gn1,
		uint32_t vbase, uint32_t *vcount, PassBCounters &k) {
	Best best;
	best.have = false, best.rate_sum = 0, best.len = 0, best.strand = 0, best.script = 0, best.task0 = 0;
	bool fwd_seen = false; // a forward record was accepted: the query's entry sits among the forward ones
#pragma unroll 1
	for (int strand = 0; strand < 2; strand++) {
		const int gn = strand == 0 ? gn0 : gn1;
		if (strand == 1) fwd_seen = best.have;
		if (gn <= 0) continue;
		const uint32_t base = strand == 0 ? base0 : base1;
		const uint32_t gp = c.g_prog_off[base], task0 = c.g_task_off[base];
		if constexpr (RECORDS) {
			RecHead *slots = c.recs + c.g_rec_off[base];
			const int sub = (int) (s - c.s_begin); // record mode runs on contiguous chunks only
			score_group(o, c.prog, gp, R, task0, strand, best, k.cells, k.calls, k.built,
					[&](int ks, uint32_t script, uint32_t task, double rate_sum, double len) {
						RecHead h;
						h.qid = (uint32_t) q, h.script = script, h.task0 = task, h.sub = sub;
						h.strand = strand, h._pad = 0, h.rate_sum = rate_sum, h.len = len;
						slots[ks] = h;
					});
		} else {
			score_group(o, c.prog, gp, R, task0, strand, best, k.cells, k.calls, k.built);
		}
	}
	if (!best.have) return;
	k.kept++;
	if (RECORDS || !o.call_consensus) return;
	// the winner's votes are cast by k_votes
	VoteItem v;
	v.q = (uint32_t) q, v.script = best.script, v.task0 = best.task0;
	v.sub = (int32_t) (s - c.s_begin), v.strand = best.strand, v.first_fwd = fwd_seen ? 1 : 0;
	c.votes[vbase + atomicAdd(vcount, 1u)] = v;
}

// Votes of the kept records (SubjectRead::updateQueryCoverage, src/SubjectRead.cpp:208-330): one
// lane per record; the record is replayed from its chain and DP results.
__global__ __launch_bounds__(64) void k_votes(Packed S, Packed Q, Opts o, ChunkDev c, uint32_t nvotes,
		const uint64_t *q_boff, uint32_t *tally, uint32_t *diff, uint32_t *ext, unsigned long long *stats) {
	const uint32_t i = xcd_item(blockIdx.x, gridDim.x) * 64u + threadIdx.x;
	if (i >= nvotes) return;
	const VoteItem v = c.votes[i];
	if (v.q == kNoRecord) return;
	const int64_t s = c.s_begin + v.sub;
	// the pair's reads as make_pair gives them, except that the query's bases are located only for
	// --merge (the plain votes read the subject's bases and the query's length)
	PairReads pr;
	pr.s_read = s, pr.q_read = (int64_t) v.q, pr.Ls = S.len[s], pr.Lq = Q.len[v.q];
	pr.subj_rc = o.call_consensus && v.strand == 1;
	pr.qry_rc = !o.call_consensus && v.strand == 1;
	pr.sw = S.words + S.woff[s];
	pr.qw = ext ? Q.words + Q.woff[v.q] : nullptr;
	const DpResults R{c.tasks, c.words, c.nsteps, c.sums};
	const uint64_t b = q_boff[v.q];
	VoteSink sink{tally + 4 * b, diff + b, pr, &stats[kStOob]};
	if (!ext) {
		vote_record(c.prog + v.script, R, v.task0, sink);
		return;
	}
	// --merge: the same votes, then the subject bases after the record's end into the extension tally
	vote_record_merging(c.prog + v.script, R, v.task0, sink, ext + (size_t) kExtPositions * 4 * v.q);
}
hipError_t launch_votes(const Packed &S, const Packed &Q, const Opts &o, const ChunkDev &c, uint32_t nvotes,
		const uint64_t *q_boff, uint32_t *tally, uint32_t *diff, uint32_t *ext, unsigned long long *stats, hipStream_t st) {
	if (nvotes == 0) return hipSuccess;
	k_votes<<<(nvotes + 63) / 64, 64, 0, st>>>(S, Q, o, c, nvotes, q_boff, tally, diff, ext, stats);
	return hipGetLastError();
}

// --mapping (SubjectRead::chooseQueryWithBestAlignment, src/SubjectRead.cpp:463-499), one wavefront
// per subject over its kept records, before the votes. A record covers the subject when it starts
// at the subject's first base, neither end op of its CIGAR is a deletion and its non-insertion ops
// span the subject (AlignmentRecord::isSubjectCoveredByQuery, src/AlignmentRecord.cpp:343-370).
// None covers: no record votes. Several: only the first one, in record-list order, with the fewest
// substituted bases (getAccuracy, :372-381) votes. Exactly one: every record keeps its votes (the
// reference leaves the list as it is then).
// The CIGAR facts come from the op stream the record writer merges (RecordV): first and last op
// type (a zero-length op counts, except a mismatch run of no bases, which adds no op), the
// non-insertion length and the mismatched bases.
struct MapStatsV {
	const DpResults &R;
	int first, last, pos;
	int64_t len, subs;
	__device__ void note(int c, int t) {
		if (t == kOpX && c == 0) return;
		if (first < 0) first = t;
		last = t;
		if (t != kOpI) len += c;
		if (t == kOpX) subs += c;
	}
	__device__ void start(int p) { pos = p; }
	__device__ void add(int c, int t) { note(c, t); }
	__device__ void run(uint32_t task, bool swapped) {
		feed_runs(R, task, swapped, [&](int cnt, int t) { note(cnt, t); });
	}
	__device__ void finish() {}
};
__global__ __launch_bounds__(64) void k_map_select(Packed S, ChunkDev c) {
	const int sl = (int) xcd_item(blockIdx.x, gridDim.x);
	const int lane = lane_id();
	const uint32_t b0 = c.sub_pair_off[sl], n = c.sub_pairs[sl];
	if (n == 0) return;
	const int Ls = S.len[c.s_begin + sl];
	const DpResults R{c.tasks, c.words, c.nsteps, c.sums};
	// order of a covering record: fewest substitutions, then the list order (entries first added by
	// a forward record, by innerId, then those first added by a reverse record, by innerId)
	uint32_t nq = 0;
	uint64_t best = ~0ull;
	for (uint32_t i = lane; i < n; i += 64) {
		const VoteItem v = c.votes[b0 + i];
		if (v.q == kNoRecord) continue;
		MapStatsV m{R, -1, -1, -1, 0, 0};
		replay_script(c.prog + v.script, v.task0, m);
		if (!(m.pos == 0 && m.first >= 0 && m.first != kOpD && m.last != kOpD && m.len >= Ls)) continue;
		nq++;
		const uint64_t key = ((uint64_t) (m.subs < 0x7FFFFFFF ? m.subs : 0x7FFFFFFF) << 33) |
				((uint64_t) (v.first_fwd ? 0u : 1u) << 32) | v.q;
		best = key < best ? key : best;
	}
	nq = wave_sum(nq);
#pragma unroll
	for (int off = 32; off > 0; off >>= 1) {
		const uint64_t o2 = __shfl_xor(best, off, 64);
		best = o2 < best ? o2 : best;
	}
	if (nq == 1) return;
	for (uint32_t i = lane; i < n; i += 64) {
		const uint32_t q = c.votes[b0 + i].q;
		if (q != kNoRecord && (nq == 0 || q != (uint32_t) best)) c.votes[b0 + i].q = kNoRecord;
	}
}
hipError_t launch_map_select(const Packed &S, const ChunkDev &c, hipStream_t st) {
	if (c.nsub == 0) return hipSuccess;
	k_map_select<<<c.nsub, 64, 0, st>>>(S, c);
	return hipGetLastError();
}

// --mapping output (QueryDataset::getCoverageAndWriteResults, src/QueryDataset.cpp:560-600): bases
// by coverage, the sum of a base's four counters (QueryRead::getCoverageAt, src/QueryRead.cpp:435-442).
// With the vote layout of VoteSink that sum is the prefix sum of the difference array (the records
// aligning a subject base to it: its matches plus its mismatches). A wavefront per read; counts below
// kCovLds go to a block histogram in LDS, the rest straight to the device histogram (cap bins, the
// last taking every larger coverage); *max_cov receives the largest coverage.
constexpr int kCovLds = 2048;
__global__ __launch_bounds__(256) void k_coverage_hist(const int32_t *len, const uint64_t *q_boff, const uint32_t *diff,
		int64_t nq, uint32_t cap, unsigned long long *hist, uint32_t *max_cov) {
	__shared__ uint32_t h[kCovLds];
	for (int i = threadIdx.x; i < kCovLds; i += blockDim.x) h[i] = 0;
	__syncthreads();
	const int lane = lane_id();
	uint32_t mx = 0;
	for (int64_t r = (int64_t) blockIdx.x * 4 + (threadIdx.x >> 6); r < nq; r += (int64_t) gridDim.x * 4) {
		const int L = len[r];
		const uint64_t b = q_boff[r];
		uint32_t carry = 0;
		for (int p0 = 0; p0 < L; p0 += 64) {
			const int p = p0 + lane;
			const uint32_t x = p < L ? diff[b + p] : 0u;
			const uint32_t run = carry + wave_excl_scan(x) + x;
			carry = __shfl(run, 63, 64);
			if (p >= L) continue;
			mx = run > mx ? run : mx;
			if (run < (uint32_t) kCovLds) atomicAdd(&h[run], 1u);
			else atomicAdd(&hist[run < cap ? run : cap - 1], 1ull);
		}
	}
	__syncthreads();
	for (int i = threadIdx.x; i < kCovLds; i += blockDim.x)
		if (h[i]) atomicAdd(&hist[(uint32_t) i < cap ? (uint32_t) i : cap - 1], (unsigned long long) h[i]);
	mx = wave_max_u32(mx);
	if (lane == 0) atomicMax(max_cov, mx);
}
hipError_t launch_coverage_hist(const int32_t *len, const uint64_t *q_boff, const uint32_t *diff, int64_t nq, uint32_t cap,
		unsigned long long *hist, uint32_t *max_cov, hipStream_t st) {
	if (nq == 0) return hipSuccess;
	const int64_t blocks = (nq + 3) / 4 < 2048 ? (nq + 3) / 4 : 2048;
	k_coverage_hist<<<(unsigned) blocks, 256, 0, st>>>(len, q_boff, diff, nq, cap, hist, max_cov);
	return hipGetLastError();
}

// Record mode: one lane per record slot; count pass (ops == nullptr) then emit pass.
__global__ __launch_bounds__(64) void k_record_ops(Packed S, Packed Q, Opts o, ChunkDev c, uint32_t nslots,
		uint32_t *op_n, const uint32_t *op_off, int32_t *pos, uint32_t *ops) {
	const uint32_t i = blockIdx.x * 64u + threadIdx.x;
	if (i >= nslots) return;
	const RecHead h = c.recs[i];
	if (h.qid == kNoRecord) {
		if (!ops) op_n[i] = 0;
		return;
	}
	const int64_t s = c.s_begin + h.sub;
	const PairReads pr = make_pair(S, Q, o, s, S.len[s], (int64_t) h.qid, h.strand);
	const DpResults R{c.tasks, c.words, c.nsteps, c.sums};
	int p = 0;
	const int n = record_ops(c.prog + h.script, pr, R, h.task0, ops ? ops + op_off[i] : nullptr, &p);
	if (ops) pos[i] = p;
	else op_n[i] = (uint32_t) n;
}
hipError_t launch_record_ops(const Packed &S, const Packed &Q, const Opts &o, const ChunkDev &c, uint32_t nslots,
		uint32_t *op_n, const uint32_t *op_off, int32_t *pos, uint32_t *ops, hipStream_t st) {
	if (nslots == 0) return hipSuccess;
	k_record_ops<<<(nslots + 63) / 64, 64, 0, st>>>(S, Q, o, c, nslots, op_n, op_off, pos, ops);
	return hipGetLastError();
}

// ---------------------------------------------------------------------------------------------
// Serving orientation: arriving query keys probe the resident subjects' all-position index. A
// query key K at posQ matching subject s forward at p is the reference's strand-0 window p; its
// reverse complement matching s at p is strand-1 window Ls - p - k (searchHashTable2,
// src/SubjectReadAligner.cpp:444-491, with the index and the probe sides exchanged).
// ---------------------------------------------------------------------------------------------
// Batch intake on the device: one wavefront per read packs its bytes into 2-bit MSB-first words
// (lane per 32-base word; every word written in full, so no clearing is needed), evaluates the
// query predicate len > (uint16) minOverlap && qualityFilter (src/QueryDataset.cpp:60-75,226) with
// wave-reduced base counts, and writes the read's key references (pos 0, k, 2k, ... and L - k,
// src/HashTableLongKey.cpp:279-289) at koff. Keys are written for every read longer than the
// minimum overlap; the probe skips those of reads that fail the quality filter.
__global__ __launch_bounds__(256) void k_read_intake(const char *raw, const uint64_t *boff, const int32_t *len,
		const uint64_t *woff, const uint64_t *koff, int64_t n, int k, uint32_t min_ov, uint64_t *words, uint64_t nwords,
		uint8_t *good, uint64_t *kref) {
	const int lane = lane_id();
	if (blockIdx.x == 0 && threadIdx.x < 2) words[nwords + threadIdx.x] = 0; // tail padding
	for (int64_t r = (int64_t) blockIdx.x * 4 + (threadIdx.x >> 6); r < n; r += (int64_t) gridDim.x * 4) {
		const int L = len[r];
		const char *s = raw + boff[r];
		uint64_t *wd = words + woff[r];
		uint32_t cnt[4] = {0, 0, 0, 0}, invalid = 0;
		for (int wi = lane; wi * 32 < L; wi += 64) {
			const int m = L - wi * 32 < 32 ? L - wi * 32 : 32;
			uint64_t x = 0;
			for (int t = 0; t < m; t++) {
				const char ch = s[wi * 32 + t];
				const int code = ascii_code(ch);
				x |= (uint64_t) code << (62 - 2 * t);
				const bool ok = ch == 'A' || ch == 'C' || ch == 'G' || ch == 'T';
				invalid += ok ? 0u : 1u;
				cnt[code] += ok ? 1u : 0u;
			}
			wd[wi] = x;
		}
		invalid = wave_sum(invalid);
#pragma unroll
		for (int c = 0; c < 4; c++) cnt[c] = wave_sum(cnt[c]);
		const bool longer = (uint64_t) L > (uint64_t) min_ov;
		if (lane == 0) {
			const uint64_t thr = (uint64_t) ((double) L * .8);
			const bool skewed = cnt[0] >= thr || cnt[1] >= thr || cnt[2] >= thr || cnt[3] >= thr;
			good[r] = longer && invalid == 0 && !skewed ? 1 : 0;
		}
		if (longer) {
			const int nk = num_keys(L, k);
			for (int t = lane; t < nk; t += 64) kref[koff[r] + t] = ((uint64_t) r << 14) | (uint64_t) key_pos(L, k, t);
		}
	}
}
hipError_t launch_read_intake(const char *raw, const uint64_t *boff, const int32_t *len, const uint64_t *woff,
		const uint64_t *koff, int64_t n, int k, uint32_t min_ov, uint64_t *words, uint64_t nwords, uint8_t *good,
		uint64_t *kref, hipStream_t st) {
	const int64_t blocks = n / 4 + 1 < 16384 ? n / 4 + 1 : 16384;
	k_read_intake<<<(unsigned) blocks, 256, 0, st>>>(raw, boff, len, woff, koff, n, k, min_ov, words, nwords, good, kref);
	return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_query_probe(Packed Q, Packed S, DevIndex six, Opts o, const uint64_t *kref,
		uint32_t nkeys, uint32_t *cnt, const uint32_t *off, uint64_t *keys_out, uint64_t *vals_out) {
	const uint32_t t = blockIdx.x * 256u + threadIdx.x;
	if (t >= nkeys) return;
	const uint64_t r = kref[t];
	const int64_t q = (int64_t) (r >> 14);
	const int pos = (int) (r & 0x3FFF);
	if (!Q.good[q]) { // a read that fails the quality filter has no seeds
		if (!keys_out) cnt[t] = 0;
		return;
	}
	const int k = o.k;
	uint64_t fh, fl;
	kmer_fwd(Q.words + Q.woff[q], pos, k, fh, fl);
	const int64_t quid = Q.uid[q];
	const uint32_t qtag = uid_tag(quid);
	const uint32_t base = keys_out ? off[t] : 0u;
	uint32_t n = 0;
	for (int strand = 0; strand < 2; strand++) {
		uint64_t hi = fh, lo = fl;
		if (strand) kmer_revcomp(k, fh, fl, hi, lo);
		uint32_t start = 0, count = 0;
		if (!ix_lookup(six, hi, lo, start, count)) continue;
		for (uint32_t e = start; e < start + count; e++) {
			const uint64_t d = six.entries[e];
			const uint64_t sr = d & kIdMask;
			if (entry_tag(d) == qtag && S.uid[sr] == quid) continue;
			if (keys_out) {
				const int p = entry_pos(d);
				const int ps = strand ? S.len[sr] - p - k : p;
				keys_out[base + n] = ((uint64_t) strand << 28) | ((uint64_t) pos << 14) | (uint64_t) ps;
				vals_out[base + n] = (sr << 22) | (uint64_t) q;
			}
			n++;
		}
	}
	if (!keys_out) cnt[t] = n;
}
hipError_t launch_query_probe(const Packed &Q, const Packed &S, const DevIndex &six, const Opts &o, const uint64_t *kref,
		uint32_t nkeys, uint32_t *cnt, const uint32_t *off, uint64_t *keys_out, uint64_t *vals_out, hipStream_t st) {
	if (nkeys == 0) return hipSuccess;
	k_query_probe<<<(nkeys + 255) / 256, 256, 0, st>>>(Q, S, six, o, kref, nkeys, cnt, off, keys_out, vals_out);
	return hipGetLastError();
}

// Kept group starts of one segment into an LDS list (wave-wide ballot compaction). Kept groups
// are a prefix of the segment's groups, so list[cnt] receives the end of the last kept group (the
// first dropped group's start, or n).
__device__ __forceinline__ uint32_t collect_groups(const uint8_t *gflag, uint32_t n, uint32_t *list, uint32_t cap) {
	uint32_t cnt = 0, end = n;
	for (uint32_t t0 = 0; t0 < n; t0 += 64) {
		const uint32_t t = t0 + lane_id();
		const uint8_t f = t < n ? gflag[t] : 0;
		const uint64_t m = __ballot(f == 3);
		const uint32_t pos = cnt + (uint32_t) __popcll(m & lanes_below_mask());
		if (f == 3 && pos < cap) list[pos] = t;
		cnt += (uint32_t) __popcll(m);
		const uint64_t dropped = __ballot(f == 1);
		if (dropped) {
			end = t0 + (uint32_t) __ffsll((unsigned long long) dropped) - 1;
			break;
		}
	}
	if (cnt <= cap && lane_id() == 0) list[cnt] = end; // list holds cap + 1 entries
	return cnt;
}

// Pass B: one wavefront per subject. The kept groups of both strands are listed in LDS with their
// query ids and paired by query (binary search over the reverse list); lanes then take the pairs
// densely. More kept groups than the lists hold: the segments are walked directly instead.
constexpr uint32_t kPassBGroups = 128;
template <bool RECORDS>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_pass_b(Packed S, Packed Q, Opts o, ChunkDev c, unsigned long long *stats) {
	__shared__ uint32_t vcount;
	__shared__ uint32_t list0[kPassBGroups + 1], list1[kPassBGroups + 1];
	__shared__ uint32_t q0s[kPassBGroups], q1s[kPassBGroups];
	__shared__ int32_t match0[kPassBGroups];
	__shared__ uint32_t rev_only[kPassBGroups];
	__shared__ uint8_t used1[kPassBGroups];
	const int sl = (int) xcd_item(blockIdx.x, gridDim.x);
	const int64_t s = c.s_begin + sl;
	const int lane = lane_id();
	if (!S.good[s]) return;
	const int Ls = S.len[s];
	const uint32_t n0 = c.seg_n[2 * sl], off0 = c.seg_off[2 * sl];
	const uint32_t n1 = c.seg_n[2 * sl + 1], off1 = c.seg_off[2 * sl + 1];
	DpResults R{c.tasks, c.words, c.nsteps, c.sums};
	PassBCounters k;
	const uint32_t vbase = c.sub_pair_off ? c.sub_pair_off[sl] : 0u;
	if (lane == 0) vcount = 0;
	const uint32_t c0 = collect_groups(c.gflag + off0, n0, list0, kPassBGroups);
	const uint32_t c1 = collect_groups(c.gflag + off1, n1, list1, kPassBGroups);
	const bool in_lds = c0 <= kPassBGroups && c1 <= kPassBGroups;
	uint32_t total = n0 + n1; // fallback: one lane per seed of both segments
	if (in_lds) {
		__syncthreads();
		for (uint32_t i = lane; i < c0; i += 64) q0s[i] = (uint32_t) seed_qid(c.seeds[off0 + list0[i]]);
		for (uint32_t j = lane; j < c1; j += 64) q1s[j] = (uint32_t) seed_qid(c.seeds[off1 + list1[j]]), used1[j] = 0;
		__syncthreads();
		for (uint32_t i = lane; i < c0; i += 64) {
			const uint32_t q = q0s[i];
			uint32_t lo = 0, hi = c1;
			while (lo < hi) {
				const uint32_t mid = (lo + hi) >> 1;
				if (q1s[mid] < q) lo = mid + 1;
				else hi = mid;
			}
			const bool found = lo < c1 && q1s[lo] == q;
			match0[i] = found ? (int32_t) lo : -1;
			if (found) used1[lo] = 1;
		}
		__syncthreads();
		uint32_t nrev = 0;
		for (uint32_t j0 = 0; j0 < c1; j0 += 64) {
			const uint32_t j = j0 + lane;
			const bool r = j < c1 && !used1[j];
			const uint64_t m = __ballot(r);
			if (r) rev_only[nrev + (uint32_t) __popcll(m & lanes_below_mask())] = j;
			nrev += (uint32_t) __popcll(m);
		}
		__syncthreads();
		total = c0 + nrev;
	}
	for (uint32_t p = lane; p < total; p += 64) {
		int64_t q = -1;
		uint32_t b0 = 0, b1 = 0;
		int g0 = 0, g1 = 0;
		if (in_lds) {
			if (p < c0) {
				q = q0s[p];
				b0 = off0 + list0[p], g0 = (int) (list0[p + 1] - list0[p]);
				if (match0[p] >= 0) {
					const uint32_t j = (uint32_t) match0[p];
					b1 = off1 + list1[j], g1 = (int) (list1[j + 1] - list1[j]);
				}
			} else {
				const uint32_t j = rev_only[p - c0];
				q = q1s[j];
				b1 = off1 + list1[j], g1 = (int) (list1[j + 1] - list1[j]);
			}
		} else if (p < n0) {
			if (c.gflag[off0 + p] == 3) {
				q = seed_qid(c.seeds[off0 + p]);
				b0 = off0 + p, g0 = (int) (group_end(c.seeds + off0, n0, p) - p);
				const int r = find_group(c.seeds + off1, n1, q);
				if (r >= 0 && c.gflag[off1 + r] == 3)
					b1 = off1 + (uint32_t) r, g1 = (int) (group_end(c.seeds + off1, n1, (uint32_t) r) - (uint32_t) r);
			}
		} else {
			const uint32_t t = p - n0;
			if (c.gflag[off1 + t] == 3) {
				const int64_t qq = seed_qid(c.seeds[off1 + t]);
				const int f = find_group(c.seeds + off0, n0, qq);
				if (!(f >= 0 && c.gflag[off0 + f] == 3)) { // else handled with the forward group
					q = qq;
					b1 = off1 + t, g1 = (int) (group_end(c.seeds + off1, n1, t) - t);
				}
			}
		}
		if (q >= 0) pass_b_pair<RECORDS>(S, Q, o, c, R, s, Ls, q, b0, g0, b1, g1, vbase, &vcount, k);
	}
	uint64_t cells = k.cells, calls = k.calls, built = k.built, kept = k.kept;
	cells = wave_sum64(cells);
	calls = wave_sum64(calls);
	built = wave_sum64(built);
	kept = wave_sum64(kept);

	if (lane == 0) {
		stat_add(stats, kStCellsExec, (unsigned long long) cells);
		stat_add(stats, kStCallsExec, (unsigned long long) calls);
		stat_add(stats, kStBuilt, (unsigned long long) built);
		stat_add(stats, kStKept, (unsigned long long) kept);

	}
}
hipError_t launch_pass_b(const Packed &S, const Packed &Q, const Opts &o, const ChunkDev &c, unsigned long long *stats,
		hipStream_t st) {
	if (c.nsub == 0) return hipSuccess;
	if (c.recs) k_pass_b<true><<<c.nsub, 64, 0, st>>>(S, Q, o, c, stats);
	else k_pass_b<false><<<c.nsub, 64, 0, st>>>(S, Q, o, c, stats);
	return hipGetLastError();
}

// ---------------------------------------------------------------------------------------------
// Serving passes. A batch's seeds are sorted by (subject, query, strand, posQ, posS) as
// hi = subject << 22 | query and lo = strand << 28 | posQ << 14 | posS, so each (subject, query,
// strand) group and each (subject, query) pair is one contiguous run. Pass A chains every group
// and plans its DP (one lane per group); pass B scores both strands of every pair, keeps the
// first best record and casts its votes (one lane per pair). Same functions as the offline
// passes, with lanes over groups instead of wavefronts over subjects.
// ---------------------------------------------------------------------------------------------
constexpr uint64_t kServeQMask = (1ull << 22) - 1;

__global__ void k_serve_mark(const uint64_t *hi, const uint64_t *lo, uint32_t n, uint32_t *gflag, uint32_t *pflag) {
	const uint32_t i = blockIdx.x * 256u + threadIdx.x;
	if (i >= n) return;
	const bool p = i == 0 || hi[i] != hi[i - 1];
	gflag[i] = (p || (lo[i] >> 28) != (lo[i - 1] >> 28)) ? 1u : 0u;
	pflag[i] = p ? 1u : 0u;
}
__global__ void k_serve_lists(const uint64_t *hi, const uint64_t *lo, uint32_t n, const uint32_t *gflag,
		const uint32_t *gx, const uint32_t *pflag, const uint32_t *px, uint32_t *glist, uint32_t *plist, uint64_t *seeds) {
	const uint32_t i = blockIdx.x * 256u + threadIdx.x;
	if (i >= n) return;
	if (gflag[i]) glist[gx[i]] = i;
	if (pflag[i]) plist[px[i]] = i;
	seeds[i] = ((hi[i] & kServeQMask) << 28) | (lo[i] & ((1ull << 28) - 1));
}
hipError_t launch_serve_mark(const uint64_t *hi, const uint64_t *lo, uint32_t n, uint32_t *gflag, uint32_t *pflag,
		hipStream_t st) {
	if (n == 0) return hipSuccess;
	k_serve_mark<<<(n + 255) / 256, 256, 0, st>>>(hi, lo, n, gflag, pflag);
	return hipGetLastError();
}
hipError_t launch_serve_lists(const uint64_t *hi, const uint64_t *lo, uint32_t n, const uint32_t *gflag,
		const uint32_t *gx, const uint32_t *pflag, const uint32_t *px, uint32_t *glist, uint32_t *plist, uint64_t *seeds,
		hipStream_t st) {
	if (n == 0) return hipSuccess;
	k_serve_lists<<<(n + 255) / 256, 256, 0, st>>>(hi, lo, n, gflag, gx, pflag, px, glist, plist, seeds);
	return hipGetLastError();
}

__device__ __forceinline__ PairReads serve_pair(const Packed &S, const Packed &Q, const Opts &o, const ServeDev &v,
		uint32_t base, int64_t &s, int &strand) {
	s = (int64_t) (v.hi[base] >> 22);
	strand = (int) (v.lo[base] >> 28);
	return make_pair(S, Q, o, s, S.len[s], (int64_t) (v.hi[base] & kServeQMask), strand);
}

__global__ __launch_bounds__(64) void k_serve_a1(Packed S, Packed Q, Opts o, ChunkDev c, ServeDev v,
		unsigned long long *stats) {
	const uint32_t g = blockIdx.x * 64u + threadIdx.x;
	if (g < v.ngroups) {
		const uint32_t base = v.glist[g];
		const uint32_t end = g + 1 < v.ngroups ? v.glist[g + 1] : v.nseeds;
		Group gr = make_group(c, base, (int) (end - base));
		chain(o, gr);
		c.g_nch[base] = gr.nch;
		c.g_best[base] = gr.best;
		int64_t s;
		int strand;
		const PairReads pr = serve_pair(S, Q, o, v, base, s, strand);
		int nt, nw, np, nsc;
		plan_count(o, gr, pr, nt, nw, np, nsc);
		c.g_tasks[g] = (uint32_t) nt;
		c.g_words[g] = (uint32_t) nw;
		c.g_prog[g] = (uint32_t) np;
	}
	if (blockIdx.x == 0 && threadIdx.x == 0) stat_add(stats, kStSeeds, (unsigned long long) v.nseeds);
}
__global__ __launch_bounds__(64) void k_serve_a2(Packed S, Packed Q, Opts o, ChunkDev c, ServeDev v) {
	const uint32_t g = blockIdx.x * 64u + threadIdx.x;
	if (g >= v.ngroups) return;
	const uint32_t base = v.glist[g];
	const uint32_t end = g + 1 < v.ngroups ? v.glist[g + 1] : v.nseeds;
	const Group gr = chained_group(c, base, (int) (end - base));
	int64_t s;
	int strand;
	const PairReads pr = serve_pair(S, Q, o, v, base, s, strand);
	plan_emit(o, gr, pr, c.tasks, c.g_task_off[g], c.g_word_off[g], c.prog + c.g_prog_off[g], (uint32_t) S.woff[s],
			(uint32_t) Q.woff[(int64_t) (v.hi[base] & kServeQMask)]);
}
__global__ __launch_bounds__(64) void k_serve_b(Packed S, Packed Q, Opts o, ChunkDev c, ServeDev v,
		const uint64_t *q_boff, uint32_t *tally, uint32_t *diff, unsigned long long *stats) {
	const uint32_t p = blockIdx.x * 64u + threadIdx.x;
	PassBCounters k;
	if (p < v.npairs) {
		const uint32_t base = v.plist[p];
		const uint32_t end = p + 1 < v.npairs ? v.plist[p + 1] : v.nseeds;
		const DpResults R{c.tasks, c.words, c.nsteps, c.sums};
		Best best;
		best.have = false, best.rate_sum = 0, best.len = 0, best.strand = 0, best.script = 0, best.task0 = 0;
		// the pair's groups: forward (strand 0) first, then reverse
		uint32_t i = base;
		while (i < end) {
			const uint32_t gi = v.gx[i]; // group index of the group starting at seed i
			const uint32_t ge = gi + 1 < v.ngroups ? v.glist[gi + 1] : v.nseeds;
			const int strand = (int) (v.lo[i] >> 28);
			score_group(o, c.prog, c.g_prog_off[gi], R, c.g_task_off[gi], strand, best, k.cells, k.calls, k.built);
			i = ge;
		}
		if (best.have) {
			k.kept++;
			int64_t s;
			int strand;
			const PairReads pr = serve_pair(S, Q, o, v, base, s, strand);
			const int64_t q = (int64_t) (v.hi[base] & kServeQMask);
			const PairReads prb = make_pair(S, Q, o, s, S.len[s], q, best.strand);
			const uint64_t b = q_boff[q];
			VoteSink sink{tally + 4 * b, diff + b, prb, &stats[kStOob]};
			vote_record(c.prog + best.script, R, best.task0, sink);
		}
	}
	uint64_t cells = wave_sum64(k.cells), calls = wave_sum64(k.calls), built = wave_sum64(k.built), kept = wave_sum64(k.kept);
	if (lane_id() == 0) {
		stat_add(stats, kStCellsExec, (unsigned long long) cells);
		stat_add(stats, kStCallsExec, (unsigned long long) calls);
		stat_add(stats, kStBuilt, (unsigned long long) built);
		stat_add(stats, kStKept, (unsigned long long) kept);
	}
}
hipError_t launch_serve_a1(const Packed &S, const Packed &Q, const Opts &o, const ChunkDev &c, const ServeDev &v,
		unsigned long long *stats, hipStream_t st) {
	if (v.ngroups == 0) return hipSuccess;
	k_serve_a1<<<(v.ngroups + 63) / 64, 64, 0, st>>>(S, Q, o, c, v, stats);
	return hipGetLastError();
}
hipError_t launch_serve_a2(const Packed &S, const Packed &Q, const Opts &o, const ChunkDev &c, const ServeDev &v,
		hipStream_t st) {
	if (v.ngroups == 0) return hipSuccess;
	k_serve_a2<<<(v.ngroups + 63) / 64, 64, 0, st>>>(S, Q, o, c, v);
	return hipGetLastError();
}
hipError_t launch_serve_b(const Packed &S, const Packed &Q, const Opts &o, const ChunkDev &c, const ServeDev &v,
		const uint64_t *q_boff, uint32_t *tally, uint32_t *diff, unsigned long long *stats, hipStream_t st) {
	if (v.npairs == 0) return hipSuccess;
	k_serve_b<<<(v.npairs + 63) / 64, 64, 0, st>>>(S, Q, o, c, v, q_boff, tally, diff, stats);
	return hipGetLastError();
}

// ---------------------------------------------------------------------------------------------
// Banded DP (BandAligner::bandedAlign / bandedAlignAffineGap)
// ---------------------------------------------------------------------------------------------
// Sequential reader of one side of a task in its frame (reverse complement when rc).
// A read segment as a stream of 2-bit codes (forward, or reverse complement when rc). The current
// 32-base word stays in a register: one load per 32 bases instead of one per base.
// Codes of a read segment in strand frame, one at a time: the current word sits pre-shifted so the
// next code is in its top bits (a reverse-complement word is reversed in 2-bit groups and
// complemented once, when loaded). Seven registers per stream.
struct PackedStream {
	static constexpr bool kTwoBit = true;
	const uint64_t *p; // next word to load
	uint64_t cur;      // codes not yet returned, next at the top
	int nleft;         // codes left in cur
	int remain;        // codes left in the segment
	int step;          // +1 forward, -1 reverse complement
	__device__ static __forceinline__ uint64_t rc_word(uint64_t x) {
		x = __builtin_bitreverse64(x);
		x = ((x >> 1) & 0x5555555555555555ull) | ((x & 0x5555555555555555ull) << 1); // back to 2-bit codes
		return x ^ 0xAAAAAAAAAAAAAAAAull;                                                 // complement: c ^ 2
	}
	// len codes from code `bit` of the word at q on (rc: the reverse complement, from that code
	// towards lower positions)
	__device__ static __forceinline__ PackedStream make(const uint64_t *q, int bit, int len, bool rc) {
		PackedStream s;
		s.remain = len, s.step = rc ? -1 : 1, s.p = q, s.cur = 0, s.nleft = 0;
		if (len <= 0) return s;
		if (!rc) s.cur = *q << (2 * bit), s.nleft = 32 - bit, s.p = q + 1;
		else s.cur = rc_word(*q) << (2 * (31 - bit)), s.nleft = bit + 1, s.p = q - 1; // read only if needed
		return s;
	}
	__device__ __forceinline__ int next() {
		if (remain <= 0) return kCodeNone;
		if (nleft == 0) {
			cur = step > 0 ? *p : rc_word(*p);
			p += step, nleft = 32;
		}
		const int c = (int) (cur >> 62);
		cur <<= 2, nleft--, remain--;
		return c;
	}
};

__device__ __forceinline__ void task_streams(const DpTask &t, const Packed &S, const Packed &Q, PackedStream &as,
		PackedStream &aq, int &lenS, int &lenQ) {
	const PackedStream subj = PackedStream::make(S.words + t.s_word, t.s_bit, t.s_len, (t.flags & kTaskSubjRC) != 0);
	const PackedStream qry = PackedStream::make(Q.words + t.q_word, t.q_bit, t.q_len, (t.flags & kTaskQryRC) != 0);
	const bool sw = (t.flags & kTaskSwapped) != 0;
	as = sw ? qry : subj;
	aq = sw ? subj : qry;
	lenS = sw ? t.q_len : t.s_len;
	lenQ = sw ? t.s_len : t.q_len;
}

// Pointer rows of one of SLOTS tasks per wavefront (two per lane in the packed form), interleaved
// so a row store is one coalesced write: word w of row i of slot s at ((i * NW) + w) * SLOTS + s.
template <int NW, int SLOTS>
struct SlotPtrs {
	uint32_t *base;
	int slot;
	__device__ __forceinline__ void put_row(int i, const uint32_t *pw) {
#pragma unroll
		for (int w = 0; w < NW; w++) base[((size_t) i * NW + w) * SLOTS + slot] = pw[w];
	}
	__device__ __forceinline__ uint32_t get(int i, int w) const { return base[((size_t) i * NW + w) * SLOTS + slot]; }
};

// Traceback of a forward-filled task into its edit words, step count and summary.
template <uint32_t MAP, class PS>
__device__ __forceinline__ uint64_t finish_task(const DpLaunch &L, uint32_t t, const DpTask &task, int lenS, int lenQ,
		const PS &ptrs) {
	const int cap = task.s_len + task.q_len;
	StepWriter wr{L.words + task.out_off, cap, -1, 0};
	const int ns = dp_traceback<MAP>(lenS, lenQ, task.band, ptrs, wr);
	wr.flush();
	L.nsteps[t] = ns < 0 ? 0u : (uint32_t) ns;
	L.sums[t] = wr.sb.finish();
	if (ns < 0) stat_add(L.stats, kStDpSkipped, 1ull);
	return band_cells(lenS, lenQ, task.band);
}

// A segment's codes staged in LDS by its lane, 16 per word (code p at bits 2 * (p & 15) of word
// p >> 4): three registers per stream instead of a packed-word reader's seven.
struct LdsStream {
	static constexpr bool kTwoBit = true;
	const uint32_t *w; // LDS
	int pos, len;
	__device__ __forceinline__ int next() {
		if (pos >= len) return kCodeNone;
		const int c = (int) ((w[pos >> 4] >> (2 * (pos & 15))) & 3u);
		pos++;
		return c;
	}
};
__device__ __forceinline__ LdsStream stage_codes(PackedStream s, int len, uint32_t *dst) {
	uint32_t acc = 0;
	for (int p = 0; p < len; p++) {
		acc |= (uint32_t) s.next() << (2 * (p & 15));
		if ((p & 15) == 15) dst[p >> 4] = acc, acc = 0;
	}
	if (len & 15) dst[len >> 4] = acc;
	return LdsStream{dst, 0, len};
}

// Batched form, two tasks per lane: persistent waves pulling 128 tasks at a time; lane l takes
// tasks 2l and 2l+1 of the batch (neighbours in the length-bucketed order) and runs them together
// in packed 16-bit arithmetic (dp_pk16.h), its four segments staged in LDS. The host selects this
// form only when every task of the class is within the 16-bit bound.
template <int BMAX, bool AFF, bool TAG = false>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(TAG ? 2 : BMAX <= 8 ? 3 : 1, 8))) void k_dp_batched_pk(
		DpLaunch L0) {
	static_assert(!TAG || AFF, "the packed tagged rows are affine");
	DpLaunch L = L0;
	if (L.range) L.order += L.range[0], L.ntasks = L.range[1] - L.range[0];
	constexpr int NW = (2 * BMAX + 1 + 15) / 16;
	const int lane = lane_id();
	uint32_t *wbase = L.scratch + (size_t) blockIdx.x * L.rows_max * NW * 128;
	SlotPtrs<NW, 128> pa{wbase, 2 * lane}, pb{wbase, 2 * lane + 1};
	// the lane's four segments (A's and B's aligner subject and query) in LDS
	extern __shared__ uint32_t lds_codes[];
	const int wps = pk_lds_words(L.rows_max);
	uint32_t *mine = lds_codes + (size_t) lane * 4 * wps;
	uint64_t cells = 0;
	for (;;) {
		uint32_t batch = 0;
		if (lane == 0) batch = atomicAdd(L.batch_counter, 1u);
		batch = __shfl(batch, 0, 64);
		const uint32_t first = batch * 128;
		if (first >= L.ntasks) break;
		const uint32_t ia = first + 2 * lane, ib = ia + 1;
		if (ia >= L.ntasks) continue;
		const uint32_t ta = L.order ? L.order[ia] : ia;
		const DpTask A = L.tasks[ta];
		PackedStream sa, qa;
		int lsa, lqa;
		task_streams(A, L.S, L.Q, sa, qa, lsa, lqa);
		const bool oka = A.band <= BMAX && lsa < L.rows_max;
		const bool has_b = ib < L.ntasks;
		uint32_t tb = 0;
		DpTask B;
		PackedStream sb = PackedStream::make(nullptr, 0, 0, false), qb = sb; // empty unless a task
		int lsb = 0, lqb = 0;
		bool okb = false;
		if (has_b) {
			tb = L.order ? L.order[ib] : ib;
			B = L.tasks[tb];
			task_streams(B, L.S, L.Q, sb, qb, lsb, lqb);
			okb = B.band <= BMAX && lsb < L.rows_max;
		}
		// the host selects this form only when every task of the class is within the 16-bit bound
		// (pk16_class_fits); a pair that is not (never expected) is counted as skipped
		const bool fit = oka && (!has_b || okb) &&
				(TAG ? pk16t_fits(L.o, lsa > lsb ? lsa : lsb, BMAX)
				     : pk16_fits(L.o, lsa, lqa) && (!has_b || pk16_fits(L.o, lsb, lqb)));
		if (!fit) {
			stat_add(L.stats, kStDpSkipped, has_b ? 2ull : 1ull);
			continue;
		}
		if (!has_b) lsb = 0, lqb = 0; // an empty second task: one row of sentinels, never read
		const Pk16Task PA{lsa, lqa, A.band, (A.flags & kTaskAlignLeft) != 0, (A.flags & kTaskAlignRight) != 0};
		const Pk16Task PB{lsb, lqb, has_b ? B.band : A.band, has_b && (B.flags & kTaskAlignLeft) != 0,
				has_b && (B.flags & kTaskAlignRight) != 0};
		LdsStream ls_sa = stage_codes(sa, lsa, mine), ls_qa = stage_codes(qa, lqa, mine + wps);
		LdsStream ls_sb = stage_codes(sb, lsb, mine + 2 * wps), ls_qb = stage_codes(qb, lqb, mine + 3 * wps);
		constexpr uint32_t kMap = TAG ? kPtrTaggedAffine : kPtrPlain;
		if constexpr (TAG) dp_forward_pk16_tagged<BMAX>(L.o, PA, PB, ls_sa, ls_qa, ls_sb, ls_qb, pa, pb);
		else dp_forward_pk16<BMAX, AFF>(L.o, PA, PB, ls_sa, ls_qa, ls_sb, ls_qb, pa, pb);
		cells += finish_task<kMap>(L, ta, A, lsa, lqa, pa);
		if (has_b) cells += finish_task<kMap>(L, tb, B, lsb, lqb, pb);
	}
	cells = wave_sum64(cells);
	if (lane == 0) stat_add(L.stats, kStCellsAll, (unsigned long long) cells);
}

// Occupancy the register allocator plans the one-task-per-lane DP for (waves per SIMD): the most
// that holds a band row in registers without spilling.
constexpr int dp_waves_per_simd(int bmax) {
	return bmax <= 2 ? 5 : bmax <= 4 ? 4 : bmax <= 8 ? 3 : bmax <= 12 ? 2 : 1;
}

// Batched form: one task per lane, persistent waves pulling 64 tasks at a time. TAG: affine rows
// in the tagged form (dp_forward_tagged), for launches with L.tagged set.
template <int BMAX, bool AFF, bool TAG = false>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(dp_waves_per_simd(BMAX), 8))) void k_dp_batched(
		DpLaunch L0) {
	static_assert(!TAG || BMAX <= kTaggedMaxBand, "tagged rows: band classes up to 12");
	DpLaunch L = L0;
	if (L.range) L.order += L.range[0], L.ntasks = L.range[1] - L.range[0];
	constexpr int NW = (2 * BMAX + 1 + 15) / 16;
	constexpr bool kSide = NW > 1 && (2 * BMAX + 1) % 16 == 1;
	using Ptrs = SplitRowPtrs<NW, kSide, 64>;
	const int lane = lane_id();
	extern __shared__ uint32_t ptr_lds[];
	Ptrs ptrs{ptr_lds, L.scratch + (size_t) blockIdx.x * L.rows_max * NW * 64, lane,
			Ptrs::rows_for((int) (L.ptr_lds_bytes / 256u)), 0u};
	uint64_t cells = 0;
	for (;;) {
		uint32_t batch = 0;
		if (lane == 0) batch = atomicAdd(L.batch_counter, 1u);
		batch = __shfl(batch, 0, 64);
		const uint32_t first = batch * 64;
		if (first >= L.ntasks) break;
		const uint32_t ti = first + lane;
		if (ti < L.ntasks) {
			const uint32_t t = L.order ? L.order[ti] : ti;
			const DpTask task = L.tasks[t];
			PackedStream as, aq;
			int lenS, lenQ;
			task_streams(task, L.S, L.Q, as, aq, lenS, lenQ);
			if (task.band <= BMAX && lenS < L.rows_max) {
				const bool la = (task.flags & kTaskAlignLeft) != 0, ra = (task.flags & kTaskAlignRight) != 0;
				if constexpr (TAG && AFF) dp_forward_tagged<BMAX>(L.o, lenS, lenQ, task.band, la, ra, as, aq, ptrs);
				else if constexpr (TAG) dp_forward_tagged_global<BMAX>(L.o, lenS, lenQ, task.band, la, ra, as, aq, ptrs);
				else dp_forward<BMAX, AFF>(L.o, lenS, lenQ, task.band, la, ra, as, aq, ptrs);
				const int cap = task.s_len + task.q_len;
				StepWriter wr{L.words + task.out_off, cap, -1, 0};
				const int ns = dp_traceback<!TAG ? kPtrPlain : AFF ? kPtrTaggedAffine : kPtrTaggedGlobal>(lenS, lenQ,
						task.band, ptrs, wr);
				wr.flush();
				L.nsteps[t] = ns < 0 ? 0u : (uint32_t) ns;
				L.sums[t] = wr.sb.finish();
				if (ns < 0) stat_add(L.stats, kStDpSkipped, 1ull);
				cells += band_cells(lenS, lenQ, task.band);
			}
		}
	}
	cells = wave_sum64(cells);
	if (lane == 0) stat_add(L.stats, kStCellsAll, (unsigned long long) cells);
}

// two tasks per lane in 16-bit halves: the tagged packed rows for tagged affine launches, the
// general packed rows otherwise; both within their 16-bit range for every call of the class
bool pk_form_fits(const Opts &o, int rows_max, int bmax, bool tagged) {
	if (bmax > kPk16MaxBand) return false;
	return tagged && o.aligner == 0 ? pk16t_class_fits(o, rows_max, bmax) : pk16_class_fits(o, rows_max);
}

// the tagged rows: the global aligner, or the affine one with equal gap opens
bool dp_tagged_form(const Opts &o, int rows_max) {
	return (o.aligner == 1 || o.dx == o.dy) && dp_tagged_fits(o, rows_max);
}

static int dp_batched_waves_per_cu_query(int bmax, bool aff, int per_lane, int rows_max, bool tagged) {
	int blocks = 0;
	hipError_t e = hipErrorInvalidValue;
#define HERO_OCC(BM)                                                                                            \
	if (bmax == BM && per_lane == 1)                                                                            \
		e = aff ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k_dp_batched<BM, true>, 64, 0)          \
		        : hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k_dp_batched<BM, false>, 64, 0);
#define HERO_OCCT(BM)                                                                                           \
	if (bmax == BM && per_lane == 1 && tagged)                                                                  \
		e = aff ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k_dp_batched<BM, true, true>, 64, 0)    \
		        : hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k_dp_batched<BM, false, true>, 64, 0);
#define HERO_OCC2(BM)                                                                                           \
	if (bmax == BM && per_lane == 2)                                                                            \
		e = aff && tagged ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k_dp_batched_pk<BM, true, true>,  \
		                  64, pk_lds_bytes(rows_max))                                                      \
		  : aff ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k_dp_batched_pk<BM, true>, 64,             \
		                  pk_lds_bytes(rows_max))                                                          \
		        : hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k_dp_batched_pk<BM, false>, 64,          \
		                  pk_lds_bytes(rows_max));
	HERO_OCC(4) HERO_OCC(6) HERO_OCC(8) HERO_OCC(10) HERO_OCC(12) HERO_OCC(16) HERO_OCC(24) HERO_OCC(32)
	HERO_OCCT(4) HERO_OCCT(6) HERO_OCCT(8) HERO_OCCT(10) HERO_OCCT(12)
	HERO_OCC2(4) HERO_OCC2(6) HERO_OCC2(8) HERO_OCC2(10) HERO_OCC2(12)
#undef HERO_OCC
#undef HERO_OCCT
#undef HERO_OCC2
	return (e == hipSuccess && blocks > 0) ? blocks : 4;
}
// Kept per (device, class, form, tasks per lane, rows): the DP stage of every chunk and serving step
// asks it for every class.
int dp_batched_waves_per_cu(int bmax, bool aff, int per_lane, int rows_max, bool tagged) {
	static std::mutex mu;
	static std::map<std::pair<uint64_t, int>, int> known;
	int dev = 0;
	if (hipGetDevice(&dev) != hipSuccess) return 4;
	const std::pair<uint64_t, int> key{((uint64_t) (uint32_t) dev << 32) | ((uint64_t) (bmax & 0xFFFF) << 16) |
			((uint64_t) (aff ? 1 : 0) << 15) | ((uint64_t) (tagged ? 1 : 0) << 14) | (uint64_t) (per_lane & 0xFF),
			per_lane == 2 ? rows_max : 0};
	{
		std::lock_guard<std::mutex> lk(mu);
		const auto it = known.find(key);
		if (it != known.end()) return it->second;
	}
	const int w = dp_batched_waves_per_cu_query(bmax, aff, per_lane, rows_max, tagged);
	std::lock_guard<std::mutex> lk(mu);
	known[key] = w;
	return w;
}

static uint32_t dp_batched_lds_bytes_query(int bmax, bool aff, bool tagged, int waves_per_cu) {
	int dev = 0, per_cu = 0, per_block = 0;
	if (waves_per_cu < 1 || hipGetDevice(&dev) != hipSuccess ||
			hipDeviceGetAttribute(&per_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) != hipSuccess ||
			hipDeviceGetAttribute(&per_block, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
		return 0;
	const void *fn = nullptr;
#define HERO_LDSF(BM)                                                                                           \
	if (bmax == BM) fn = tagged ? (aff ? (const void *) k_dp_batched<BM, true, true> : (const void *) k_dp_batched<BM, false, true>) \
	                            : (aff ? (const void *) k_dp_batched<BM, true> : (const void *) k_dp_batched<BM, false>);
	HERO_LDSF(4) HERO_LDSF(6) HERO_LDSF(8) HERO_LDSF(10) HERO_LDSF(12)
#undef HERO_LDSF
	if (!tagged) {
		if (bmax == 16) fn = aff ? (const void *) k_dp_batched<16, true> : (const void *) k_dp_batched<16, false>;
		if (bmax == 24) fn = aff ? (const void *) k_dp_batched<24, true> : (const void *) k_dp_batched<24, false>;
		if (bmax == 32) fn = aff ? (const void *) k_dp_batched<32, true> : (const void *) k_dp_batched<32, false>;
	}
	if (!fn) return 0;
	int without = 0;
	if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&without, fn, 64, 0) != hipSuccess || without < 1) return 0;
	const int want = waves_per_cu < without ? waves_per_cu : without;
	uint32_t bytes = (uint32_t) std::min(per_cu / want, per_block) & ~511u;
	for (; bytes > 0; bytes -= 512) { // LDS allocation granules may round a share up
		int blocks = 0;
		if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, fn, 64, bytes) == hipSuccess && blocks >= want) break;
	}
	return bytes;
}

// The answer depends only on the kernel, the device and the waves per CU: every DP stage asks it
// again (each chunk, each serving step), so it is kept per (device, class, form, waves per CU).
uint32_t dp_batched_lds_bytes(int bmax, bool aff, bool tagged, int waves_per_cu) {
	static std::mutex mu;
	static std::map<uint64_t, uint32_t> known;
	int dev = 0;
	if (hipGetDevice(&dev) != hipSuccess) return 0;
	const uint64_t key = ((uint64_t) (uint32_t) dev << 32) | ((uint64_t) (bmax & 0xFFFF) << 16) |
			((uint64_t) (aff ? 1 : 0) << 15) | ((uint64_t) (tagged ? 1 : 0) << 14) | (uint64_t) (waves_per_cu & 0x3FFF);
	{
		std::lock_guard<std::mutex> lk(mu);
		const auto it = known.find(key);
		if (it != known.end()) return it->second;
	}
	const uint32_t bytes = dp_batched_lds_bytes_query(bmax, aff, tagged, waves_per_cu);
	std::lock_guard<std::mutex> lk(mu);
	known[key] = bytes;
	return bytes;
}

__device__ __forceinline__ int task_bucket(const DpTask &t, int rows_max) {
	const int lenS = t.s_len > t.q_len ? t.s_len : t.q_len; // the aligner's subject is the longer side
	int cls = kNumBandClasses;
	if (lenS < rows_max)
		for (int c = kNumBandClasses - 1; c >= 0; c--)
			if (t.band <= kBandClass[c]) cls = c;
	// descending length, and within a length the aligned-left calls together: a wavefront of
	// aligned-left calls is the only kind whose last rows need the column-lenQ rule
	constexpr int kLenRows = kLenBins / 2;
	const int bin = 2 * (kLenRows - 1 - (lenS < kLenRows ? lenS : kLenRows - 1)) + ((t.flags & kTaskAlignLeft) ? 1 : 0);
	return cls * kLenBins + bin;
}

__global__ __launch_bounds__(1024) void k_task_histogram(const DpTask *tasks, uint32_t n0, const uint32_t *n_dev,
		int rows_max, uint32_t *hist) {
	const uint32_t n = n_dev ? *n_dev : n0;
	__shared__ uint32_t h[kNumBuckets];
	for (int i = threadIdx.x; i < kNumBuckets; i += blockDim.x) h[i] = 0;
	__syncthreads();
	for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x)
		atomicAdd(&h[task_bucket(tasks[t], rows_max)], 1u);
	__syncthreads();
	for (int i = threadIdx.x; i < kNumBuckets; i += blockDim.x)
		if (h[i]) atomicAdd(&hist[i], h[i]);
}
hipError_t launch_task_histogram(const DpTask *tasks, uint32_t ntasks, const uint32_t *ntasks_dev, int rows_max,
		uint32_t *hist, hipStream_t st) {
	if (ntasks == 0) return hipSuccess;
	const unsigned blocks = ntasks / (1024 * 16) + 1 < 512 ? ntasks / (1024 * 16) + 1 : 512;
	k_task_histogram<<<blocks, 1024, 0, st>>>(tasks, ntasks, ntasks_dev, rows_max, hist);
	return hipGetLastError();
}

// Scatter into the bucket order. A block ranks its tasks per bucket in LDS, reserves each
// bucket's range with one global atomic, then writes: few global atomics even when most tasks
// share a handful of buckets.
constexpr int kScatterThreads = 1024, kScatterPerThread = 8;
__global__ __launch_bounds__(1024) void k_task_scatter(const DpTask *tasks, uint32_t n0, const uint32_t *n_dev,
		int rows_max, uint32_t *cursor, uint32_t *order) {
	const uint32_t n = n_dev ? *n_dev : n0;
	__shared__ uint32_t cnt[kNumBuckets];
	for (int i = threadIdx.x; i < kNumBuckets; i += blockDim.x) cnt[i] = 0;
	__syncthreads();
	const uint32_t base = blockIdx.x * (kScatterThreads * kScatterPerThread);
	int bk[kScatterPerThread];
	uint32_t rk[kScatterPerThread];
#pragma unroll
	for (int i = 0; i < kScatterPerThread; i++) {
		const uint32_t t = base + i * kScatterThreads + threadIdx.x;
		bk[i] = -1;
		if (t < n) {
			bk[i] = task_bucket(tasks[t], rows_max);
			rk[i] = atomicAdd(&cnt[bk[i]], 1u);
		}
	}
	__syncthreads();
	for (int i = threadIdx.x; i < kNumBuckets; i += blockDim.x)
		if (cnt[i]) cnt[i] = atomicAdd(&cursor[i], cnt[i]);
	__syncthreads();
#pragma unroll
	for (int i = 0; i < kScatterPerThread; i++)
		if (bk[i] >= 0) order[cnt[bk[i]] + rk[i]] = base + i * kScatterThreads + threadIdx.x;
}
hipError_t launch_task_scatter(const DpTask *tasks, uint32_t ntasks, const uint32_t *ntasks_dev, int rows_max,
		uint32_t *cursor, uint32_t *order, hipStream_t st) {
	if (ntasks == 0) return hipSuccess;
	const uint32_t per_block = kScatterThreads * kScatterPerThread;
	k_task_scatter<<<(ntasks + per_block - 1) / per_block, kScatterThreads, 0, st>>>(tasks, ntasks, ntasks_dev, rows_max,
			cursor, order);
	return hipGetLastError();
}

__global__ void k_class_ranges(const uint32_t *bucket_off, uint32_t *ctl) {
	const int c = threadIdx.x;
	if (c <= kNumBandClasses) ctl[kCtlRange + c] = bucket_off[(size_t) c * kLenBins];
	if (c == kNumBandClasses + 1) ctl[kCtlRange + c] = ctl[0];
	if (c < kCtlWords - kCtlCounter) ctl[kCtlCounter + c] = 0;
}
hipError_t launch_class_ranges(const uint32_t *bucket_off, uint32_t *ctl, hipStream_t st) {
	k_class_ranges<<<1, 64, 0, st>>>(bucket_off, ctl);
	return hipGetLastError();
}
__global__ void k_fill_u32(uint32_t *p, uint32_t n, uint32_t v) {
	for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}
hipError_t launch_fill_u32(uint32_t *p, uint32_t n, uint32_t v, hipStream_t st) {
	if (n == 0) return hipSuccess;
	k_fill_u32<<<(n + 255) / 256 < 1024 ? (n + 255) / 256 : 1024, 256, 0, st>>>(p, n, v);
	return hipGetLastError();
}

hipError_t launch_dp_batched(const DpLaunch &L, int bmax, int per_lane, int nwaves, hipStream_t st) {
	if (L.ntasks == 0 && !L.range) return hipSuccess;
	const bool aff = L.o.aligner == 0;
#define HERO_DPB(BM)                                                                               \
	if (bmax == BM && per_lane == 1) {                                                             \
		if (aff) k_dp_batched<BM, true><<<nwaves, 64, L.ptr_lds_bytes, st>>>(L);                     \
		else k_dp_batched<BM, false><<<nwaves, 64, L.ptr_lds_bytes, st>>>(L);                        \
		return hipGetLastError();                                                                  \
	}
#define HERO_DPBT(BM)                                                                              \
	if (bmax == BM && per_lane == 1 && L.tagged) {                                                 \
		if (aff) k_dp_batched<BM, true, true><<<nwaves, 64, L.ptr_lds_bytes, st>>>(L);               \
		else k_dp_batched<BM, false, true><<<nwaves, 64, L.ptr_lds_bytes, st>>>(L);                  \
		return hipGetLastError();                                                                  \
	}
#define HERO_DPB2(BM)                                                                              \
	if (bmax == BM && per_lane == 2) {                                                             \
		if (aff && L.tagged) k_dp_batched_pk<BM, true, true><<<nwaves, 64, pk_lds_bytes(L.rows_max), st>>>(L); \
		else if (aff) k_dp_batched_pk<BM, true><<<nwaves, 64, pk_lds_bytes(L.rows_max), st>>>(L);    \
		else k_dp_batched_pk<BM, false><<<nwaves, 64, pk_lds_bytes(L.rows_max), st>>>(L);            \
		return hipGetLastError();                                                                  \
	}
	HERO_DPBT(4) HERO_DPBT(6) HERO_DPBT(8) HERO_DPBT(10) HERO_DPBT(12)
	HERO_DPB(4) HERO_DPB(6) HERO_DPB(8) HERO_DPB(10) HERO_DPB(12) HERO_DPB(16) HERO_DPB(24) HERO_DPB(32)
	HERO_DPB2(4) HERO_DPB2(6) HERO_DPB2(8) HERO_DPB2(10) HERO_DPB2(12)
#undef HERO_DPBT
#undef HERO_DPB
#undef HERO_DPB2
	return hipErrorInvalidValue;
}

// ---------------------------------------------------------------------------------------------
// Consensus (QueryDataset::callConsensus, src/QueryDataset.cpp:393-412)
// ---------------------------------------------------------------------------------------------
// Writes the corrected bases of good reads as ASCII at their base offsets.
// Consensus, one wavefront per read: lanes take 64 positions at a time, the matched-run
// difference array is decoded by a wave-wide inclusive scan (wrapping u32 sums) with a carry
// between rounds and added to the own base's counter, then every position decides independently
// (consensus_base).
__global__ __launch_bounds__(256) void k_consensus(Packed Q, Opts o, const uint64_t *q_boff, const uint64_t *t_off,
		const uint32_t *tally, const uint32_t *diff, int64_t q0, int64_t nq, uint8_t *codes_out, uint32_t *decoded_out,
		unsigned long long *stats) {
	const int64_t i = (int64_t) blockIdx.x * 4 + (threadIdx.x >> 6);
	const int lane = lane_id();
	if (i >= nq) return;
	const int64_t q = q0 + i;
	const uint64_t b = q_boff[q], tb = t_off[q];
	const int L = Q.len[q];
	const bool good = Q.good[q];
	const uint64_t *w = Q.words + Q.woff[q];
	uint32_t carry = 0, changed = 0;
	for (int p0 = 0; p0 < L; p0 += 64) {
		const int p = p0 + lane;
		uint32_t c[4] = {0, 0, 0, 0}, x = 0;
		int own = 0;
		if (p < L) {
			const uint4 t = *(const uint4 *) (tally + 4 * (tb + p));
			c[0] = t.x, c[1] = t.y, c[2] = t.z, c[3] = t.w;
			x = diff[tb + p];
			own = base_at(w, p);
		}
		const uint32_t run = carry + wave_excl_scan(x) + x;
		carry = __shfl(run, 63, 64);
		if (p >= L) continue;
		// records aligning a base to p, less the mismatch votes: the match votes for the own base
		const uint32_t m = run - (c[0] + c[1] + c[2] + c[3]);
		c[0] += own == 0 ? m : 0u;
		c[1] += own == 1 ? m : 0u;
		c[2] += own == 2 ? m : 0u;
		c[3] += own == 3 ? m : 0u;
		if (decoded_out) *(uint4 *) (decoded_out + 4 * (b + p)) = make_uint4(c[0], c[1], c[2], c[3]);
		if (!good) continue; // not corrected (src/QueryDataset.cpp:404-406); the host copies its bytes
		c[own] += 1;         // self vote (src/QueryRead.cpp:149)
		const int outc = consensus_base(o, c, own);
		changed += outc != own ? 1u : 0u;
		codes_out[b + p] = (uint8_t) code_ascii(outc);
	}
	const uint64_t m = __ballot(changed != 0);
	if (stats && good && lane == 0 && m) stat_add(stats, kStReadsChanged, 1ull);
}
hipError_t launch_consensus(const Packed &Q, const Opts &o, const uint64_t *q_boff, const uint64_t *t_off,
		const uint32_t *tally, const uint32_t *diff, int64_t q0, int64_t nq, uint8_t *codes_out, uint32_t *decoded_out,
		unsigned long long *stats, hipStream_t st) {
	if (nq == 0) return hipSuccess;
	k_consensus<<<(unsigned) ((nq + 3) / 4), 256, 0, st>>>(Q, o, q_boff, t_off, tally, diff, q0, nq, codes_out,
			decoded_out, stats);
	return hipGetLastError();
}

// Corrected bases (ASCII A/C/G/T) as 2-bit codes, 16 per u32: base i at bits 2 (i % 16) of word
// i / 16, i.e. at bits 2 (i % 4) of byte i / 4 (the code is (c >> 1) & 3, as the read packing's).
// The host copies a quarter of the bytes and unpacks them; bytes of reads that were not corrected
// pack to some code and the host puts their own bytes back.
__global__ __launch_bounds__(256) void k_pack_bases(const uint8_t *ascii, uint64_t nb, uint32_t *packed) {
	const uint64_t t = (uint64_t) blockIdx.x * 256 + threadIdx.x;
	const uint64_t b0 = t * 16;
	if (b0 >= nb) return;
	uint32_t w = 0;
	if (b0 + 16 <= nb) {
		const uint4 v = *(const uint4 *) (ascii + b0);
		const uint32_t x[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
		for (int k = 0; k < 4; k++) {
			const uint32_t y = (x[k] >> 1) & 0x03030303u; // one code in the low bits of each byte
			const uint32_t byte = (y & 3u) | ((y >> 6) & 0xCu) | ((y >> 12) & 0x30u) | ((y >> 18) & 0xC0u);
			w |= byte << (8 * k);
		}
	} else {
		for (uint64_t i = b0; i < nb; i++) w |= (((uint32_t) ascii[i] >> 1) & 3u) << (2 * (i - b0));
	}
	packed[t] = w;
}
hipError_t launch_pack_bases(const uint8_t *ascii, uint64_t nb, uint32_t *packed, hipStream_t st) {
	const uint64_t words = (nb + 15) / 16;
	if (words == 0) return hipSuccess;
	k_pack_bases<<<(unsigned) ((words + 255) / 256), 256, 0, st>>>(ascii, nb, packed);
	return hipGetLastError();
}

// ---------------------------------------------------------------------------------------------
// Exclusive scan of u32 (block-local scan, recursive scan of block sums, add-back)
// ---------------------------------------------------------------------------------------------
constexpr int kScanThreads = 256, kScanItems = 16, kScanBlock = kScanThreads * kScanItems;

__global__ __launch_bounds__(256) void k_scan_block(const uint32_t *in, uint32_t *out, int64_t n, uint32_t *block_sums) {
	__shared__ uint32_t sh[kScanThreads];
	const int64_t base = (int64_t) blockIdx.x * kScanBlock + (int64_t) threadIdx.x * kScanItems;
	uint32_t v[kScanItems];
	uint32_t sum = 0;
#pragma unroll
	for (int i = 0; i < kScanItems; i++) {
		v[i] = base + i < n ? in[base + i] : 0u;
		sum += v[i];
	}
	sh[threadIdx.x] = sum;
	__syncthreads();
	for (int off = 1; off < kScanThreads; off <<= 1) {
		const uint32_t t = threadIdx.x >= (unsigned) off ? sh[threadIdx.x - off] : 0u;
		__syncthreads();
		sh[threadIdx.x] += t;
		__syncthreads();
	}
	uint32_t run = sh[threadIdx.x] - sum;
#pragma unroll
	for (int i = 0; i < kScanItems; i++) {
		if (base + i < n) out[base + i] = run;
		run += v[i];
	}
	if (threadIdx.x == kScanThreads - 1) block_sums[blockIdx.x] = sh[threadIdx.x];
}
__global__ __launch_bounds__(256) void k_scan_add(uint32_t *out, int64_t n, const uint32_t *block_offsets) {
	const int64_t i = (int64_t) blockIdx.x * kScanThreads + threadIdx.x;
	if (i < n) out[i] += block_offsets[i / kScanBlock];
}
// 64-bit total of n words: a block's four wavefronts add their sums in LDS, then one atomic per
// block (atomics of every wavefront on the one address serialize: 70 us for 4096 of them)
__global__ __launch_bounds__(256) void k_sum64(const uint32_t *in, int64_t n, unsigned long long *total) {
	__shared__ uint64_t part[4];
	uint64_t s = 0;
	for (int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t) gridDim.x * blockDim.x) s += in[i];
	s = wave_sum64(s);
	if (lane_id() == 0) part[threadIdx.x >> 6] = s;
	__syncthreads();
	if (threadIdx.x == 0) {
		const uint64_t t = part[0] + part[1] + part[2] + part[3];
		if (t) atomicAdd(total, (unsigned long long) t);
	}
}

// Exclusive scans of K arrays of one length in the same launches (blockIdx.y = array): a block
// scans kScanBlock items of its array and writes its 64-bit sum; the block sums are scanned in 64
// bits (scan_u64), whose last entry is the array's exact total; a second launch adds each block's
// offset. The outputs are the scans modulo 2^32 (callers refuse totals of 2^32 or more).
struct ScanMulti {
	const uint32_t *in[kScanMulti];
	uint32_t *out[kScanMulti];
	uint64_t *bsum[kScanMulti], *boff[kScanMulti];
};
// The block's items pass through LDS (a padding word per 16 keeps a thread's 16 consecutive items in
// distinct banks), so the loads and stores are coalesced rows of the array.
__device__ __forceinline__ int scanm_pad(int i) { return i + (i >> 4); }
__global__ __launch_bounds__(256) void k_scanm_block(ScanMulti m, int64_t n) {
	const int k = (int) blockIdx.y;
	const uint32_t *in = m.in[k];
	uint32_t *out = m.out[k];
	__shared__ uint64_t sh[kScanThreads];
	__shared__ uint32_t tile[kScanBlock + kScanBlock / 16];
	const int64_t b0 = (int64_t) blockIdx.x * kScanBlock;
#pragma unroll
	for (int j = 0; j < kScanItems; j++) {
		const int idx = j * kScanThreads + (int) threadIdx.x;
		tile[scanm_pad(idx)] = b0 + idx < n ? in[b0 + idx] : 0u;
	}
	__syncthreads();
	const int t0 = (int) threadIdx.x * kScanItems;
	uint32_t v[kScanItems];
	uint64_t sum = 0;
#pragma unroll
	for (int i = 0; i < kScanItems; i++) {
		v[i] = tile[scanm_pad(t0 + i)];
		sum += v[i];
	}
	sh[threadIdx.x] = sum;
	__syncthreads();
	for (int off = 1; off < kScanThreads; off <<= 1) {
		const uint64_t t = threadIdx.x >= (unsigned) off ? sh[threadIdx.x - off] : 0ull;
		__syncthreads();
		sh[threadIdx.x] += t;
		__syncthreads();
	}
	uint64_t run = sh[threadIdx.x] - sum;
#pragma unroll
	for (int i = 0; i < kScanItems; i++) {
		tile[scanm_pad(t0 + i)] = (uint32_t) run; // a thread rewrites only its own items
		run += v[i];
	}
	if (threadIdx.x == kScanThreads - 1) m.bsum[k][blockIdx.x] = sh[threadIdx.x];
	__syncthreads();
#pragma unroll
	for (int j = 0; j < kScanItems; j++) {
		const int idx = j * kScanThreads + (int) threadIdx.x;
		if (b0 + idx < n) out[b0 + idx] = tile[scanm_pad(idx)];
	}
}
__global__ __launch_bounds__(256) void k_scanm_add(ScanMulti m, int64_t n) {
	const int k = (int) blockIdx.y;
	const int64_t i = (int64_t) blockIdx.x * kScanThreads + threadIdx.x;
	if (i < n) m.out[k][i] += (uint32_t) m.boff[k][i / kScanBlock];
}
// totals[k] = boff[k][nb]; bsum[k][nb] = 0 first (the scan's extra entry that becomes the total)
__global__ void k_scanm_edge(ScanMulti m, int K, int64_t nb, uint64_t *totals, int zero) {
	const int k = (int) threadIdx.x;
	if (k >= K) return;
	if (zero) m.bsum[k][nb] = 0ull;
	else totals[k] = m.boff[k][nb];
}

size_t scan_multi_tmp_bytes(int K, int64_t n) {
	const size_t nb1 = (size_t) ((n + kScanBlock - 1) / kScanBlock) + 1;
	return (size_t) K * 2 * ((nb1 * 8 + 255) & ~(size_t) 255) + scan_u64_tmp_bytes((int64_t) nb1);
}

hipError_t scan_u32_multi(int K, const uint32_t *const *in, uint32_t *const *out, int64_t n, uint64_t *totals_dev,
		void *tmp, size_t tmp_bytes, hipStream_t st) {
	if (K < 1 || K > kScanMulti || n < 0 || tmp_bytes < scan_multi_tmp_bytes(K, n)) return hipErrorInvalidValue;
	const int64_t nb = (n + kScanBlock - 1) / kScanBlock;
	const size_t part = (((size_t) nb + 1) * 8 + 255) & ~(size_t) 255;
	ScanMulti m;
	std::memset(&m, 0, sizeof m);
	uint8_t *t = (uint8_t *) tmp;
	for (int k = 0; k < K; k++) {
		m.in[k] = in[k], m.out[k] = out[k];
		m.bsum[k] = (uint64_t *) (t + (size_t) (2 * k) * part), m.boff[k] = (uint64_t *) (t + (size_t) (2 * k + 1) * part);
	}
	void *stmp = t + (size_t) (2 * K) * part;
	const size_t stmp_bytes = tmp_bytes - (size_t) (2 * K) * part;
	k_scanm_edge<<<1, 64, 0, st>>>(m, K, nb, totals_dev, 1);
	hipError_t e = hipGetLastError();
	if (e == hipSuccess && nb > 0) {
		k_scanm_block<<<dim3((unsigned) nb, (unsigned) K), kScanThreads, 0, st>>>(m, n);
		e = hipGetLastError();
	}
	for (int k = 0; k < K && e == hipSuccess; k++) e = scan_u64(m.bsum[k], m.boff[k], nb + 1, stmp, stmp_bytes, st);
	if (e == hipSuccess && nb > 1) {
		k_scanm_add<<<dim3((unsigned) ((n + kScanThreads - 1) / kScanThreads), (unsigned) K), kScanThreads, 0, st>>>(m, n);
		e = hipGetLastError();
	}
	if (e == hipSuccess) {
		k_scanm_edge<<<1, 64, 0, st>>>(m, K, nb, totals_dev, 0);
		e = hipGetLastError();
	}
	return e;
}

size_t scan_tmp_words(int64_t n) {
	size_t words = 0;
	int64_t m = n;
	while (m > 1) {
		m = (m + kScanBlock - 1) / kScanBlock;
		words += (size_t) m + 1;
	}
	return words + 1;
}

static hipError_t scan_rec(const uint32_t *in, uint32_t *out, int64_t n, uint32_t *tmp, hipStream_t st) {
	if (n <= 0) return hipSuccess;
	const int64_t nb = (n + kScanBlock - 1) / kScanBlock;
	uint32_t *sums = tmp;
	k_scan_block<<<(unsigned) nb, kScanThreads, 0, st>>>(in, out, n, sums);
	hipError_t e = hipGetLastError();
	if (e != hipSuccess || nb == 1) return e;
	e = scan_rec(sums, sums, nb, tmp + nb + 1, st);
	if (e != hipSuccess) return e;
	k_scan_add<<<(unsigned) ((n + kScanThreads - 1) / kScanThreads), kScanThreads, 0, st>>>(out, n, sums);
	return hipGetLastError();
}

hipError_t scan_u32(const uint32_t *in, uint32_t *out, int64_t n, unsigned long long *total_dev, uint32_t *tmp,
		size_t tmp_words, hipStream_t st) {
	if (tmp_words < scan_tmp_words(n)) return hipErrorInvalidValue;
	// zeroed by a kernel rather than a memset, so the scan can be captured in a graph
	hipError_t e = launch_fill_u32((uint32_t *) total_dev, 2, 0u, st);
	if (e != hipSuccess) return e;
	if (n > 0) {
		k_sum64<<<(unsigned) (n / 8192 + 1 < 512 ? n / 8192 + 1 : 512), 256, 0, st>>>(in, n, total_dev);
		e = hipGetLastError();
		if (e != hipSuccess) return e;
	}
	return scan_rec(in, out, n, tmp, st);
}

} // namespace hero
