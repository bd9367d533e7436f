// Locality order of a read set. Reads are processed as subjects, and their vote tallies laid out,
// in the order of a per-read key: the smallest 32-bit mix of the read's canonical 16-base windows
// (its minimizer). Reads that overlap on the genome share most windows and so, often, the
// minimizer, so a run of consecutive subjects in this order probes the same index slots, reads
// the same bucket entries and query reads, and votes into the same tally lines. Results do not
// depend on the order: votes are integer sums, and every decision is made per subject or per
// (subject, query) pair.
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace hero {
namespace {

constexpr int kOrderK = 16; // window length of the key: one 32-bit canonical code

// 32-bit finalizer (xorshift-multiply rounds): spreads canonical codes over the key range
__device__ __forceinline__ uint32_t mix32(uint32_t x) {
	x ^= x >> 16;
	x *= 0x7feb352du;
	x ^= x >> 15;
	x *= 0x846ca68bu;
	x ^= x >> 16;
	return x;
}

// key[i] = min over windows of mix32(canonical code); reads shorter than a window, or not active
// (their blocks return at once), sort last. val[i] = i.
__global__ __launch_bounds__(256) void k_order_keys(Packed R, const uint8_t *active, uint64_t *keys, uint64_t *vals) {
	const int64_t i = (int64_t) blockIdx.x * 256 + threadIdx.x;
	if (i >= R.n) return;
	const int L = R.len[i];
	uint32_t best = 0xFFFFFFFFu;
	if (active[i] && L >= kOrderK) {
		const uint64_t *w = R.words + R.woff[i];
		uint32_t fwd = 0, rc = 0;
		for (int p0 = 0; p0 < L; p0 += 32) {
			const uint64_t word = w[p0 >> 5];
			const int m = L - p0 < 32 ? L - p0 : 32;
			for (int t = 0; t < m; t++) {
				const uint32_t c = (uint32_t) (word >> (62 - 2 * t)) & 3u;
				fwd = (fwd << 2) | c;              // the last 16 bases
				rc = (rc >> 2) | ((c ^ 2u) << 30); // their reverse complement
				if (p0 + t >= kOrderK - 1) {
					const uint32_t h = mix32(fwd < rc ? fwd : rc);
					best = h < best ? h : best;
				}
			}
		}
	}
	keys[i] = best;
	vals[i] = (uint64_t) i;
}

// Read r of the ordered set is read perm[r] of R: length, uniqueId, flag, word count.
__global__ __launch_bounds__(256) void k_order_meta(Packed R, const uint8_t *flag, const uint64_t *perm, int32_t *len,
		int64_t *uid, uint8_t *good, uint32_t *nwords, uint32_t *nbases) {
	const int64_t r = (int64_t) blockIdx.x * 256 + threadIdx.x;
	if (r >= R.n) return;
	const int64_t i = (int64_t) perm[r];
	const int L = R.len[i];
	len[r] = L, uid[r] = R.uid[i], good[r] = flag[i];
	nwords[r] = (uint32_t) (L + 31) / 32u;
	nbases[r] = (uint32_t) L;
}

// The packed words of the ordered set, one wavefront per read; and, when t_off is given, the
// tally offset of each read of R: the base offset of its place in the order.
__global__ __launch_bounds__(256) void k_order_words(Packed R, const uint64_t *perm, const uint64_t *woff,
		uint64_t *words, const uint64_t *boff_ordered, uint64_t *t_off) {
	const int64_t r = (int64_t) blockIdx.x * 4 + (threadIdx.x >> 6);
	const int lane = (int) threadIdx.x & 63;
	if (r >= R.n) return;
	const int64_t i = (int64_t) perm[r];
	const uint64_t *src = R.words + R.woff[i];
	const uint32_t nw = (uint32_t) (R.len[i] + 31) / 32u;
	for (uint32_t t = (uint32_t) lane; t < nw; t += 64) words[woff[r] + t] = src[t];
	if (t_off && lane == 0) t_off[i] = boff_ordered[r];
}

// Exclusive scan of u32 values into u64 sums (block-local, recursive over block sums, add-back).
constexpr int kScanT = 256, kScanI = 8, kScanB = kScanT * kScanI;
template <class TI>
__global__ __launch_bounds__(256) void k_scan64_block(const TI *in, uint64_t *out, int64_t n, uint64_t *block_sums) {
	__shared__ uint64_t sh[kScanT];
	const int64_t base = (int64_t) blockIdx.x * kScanB + (int64_t) threadIdx.x * kScanI;
	uint64_t v[kScanI], sum = 0;
#pragma unroll
	for (int t = 0; t < kScanI; t++) {
		v[t] = base + t < n ? (uint64_t) in[base + t] : 0ull;
		sum += v[t];
	}
	sh[threadIdx.x] = sum;
	__syncthreads();
	for (int off = 1; off < kScanT; off <<= 1) {
		const uint64_t x = threadIdx.x >= (unsigned) off ? sh[threadIdx.x - off] : 0ull;
		__syncthreads();
		sh[threadIdx.x] += x;
		__syncthreads();
	}
	uint64_t run = sh[threadIdx.x] - sum;
#pragma unroll
	for (int t = 0; t < kScanI; t++) {
		if (base + t < n) out[base + t] = run;
		run += v[t];
	}
	if (threadIdx.x == kScanT - 1) block_sums[blockIdx.x] = sh[threadIdx.x];
}
__global__ __launch_bounds__(256) void k_scan64_add(uint64_t *out, int64_t n, const uint64_t *block_offsets) {
	const int64_t i = (int64_t) blockIdx.x * kScanT + threadIdx.x;
	if (i < n) out[i] += block_offsets[i / kScanB];
}
size_t scan64_tmp_words(int64_t n) {
	size_t words = 0;
	for (int64_t m = n; m > 1;) {
		m = (m + kScanB - 1) / kScanB;
		words += (size_t) m + 1;
	}
	return words + 1;
}
template <class TI>
hipError_t scan64_rec(const TI *in, uint64_t *out, int64_t n, uint64_t *tmp, hipStream_t st) {
	if (n <= 0) return hipSuccess;
	const int64_t nb = (n + kScanB - 1) / kScanB;
	uint64_t *sums = tmp;
	k_scan64_block<TI><<<(unsigned) nb, kScanT, 0, st>>>(in, out, n, sums);
	hipError_t e = hipGetLastError();
	if (e != hipSuccess || nb == 1) return e;
	if ((e = scan64_rec<uint64_t>(sums, sums, nb, tmp + nb + 1, st)) != hipSuccess) return e;
	k_scan64_add<<<(unsigned) ((n + kScanT - 1) / kScanT), kScanT, 0, st>>>(out, n, sums);
	return hipGetLastError();
}

size_t align256(size_t b) { return (b + 255) & ~(size_t) 255; }

// Index keys per read: the reference's key positions (0, k, 2k, ... and L - k) or every position.
__global__ __launch_bounds__(256) void k_key_counts(Packed R, int k, int all_positions, uint32_t *cnt) {
	const int64_t i = (int64_t) blockIdx.x * 256 + threadIdx.x;
	if (i >= R.n) return;
	const int L = R.len[i];
	cnt[i] = R.good[i] ? (uint32_t) (all_positions ? (L >= k ? L - k + 1 : 0) : num_keys(L, k)) : 0u;
}

} // namespace

size_t scan_u64_tmp_bytes(int64_t n) { return scan64_tmp_words(n) * 8; }
hipError_t scan_u64(const uint64_t *in, uint64_t *out, int64_t n, void *tmp, size_t tmp_bytes, hipStream_t st) {
	if (tmp_bytes < scan_u64_tmp_bytes(n)) return hipErrorInvalidValue;
	return scan64_rec<uint64_t>(in, out, n, (uint64_t *) tmp, st);
}

size_t key_offsets_tmp_bytes(int64_t n) { return align256(((size_t) n + 1) * 4) + align256(scan64_tmp_words(n + 1) * 8); }

// koff[i] = the index keys of reads before i (exclusive scan), koff[n] = all of them.
hipError_t launch_key_offsets(const Packed &R, int k, bool all_positions, uint64_t *koff, void *tmp, size_t tmp_bytes,
		hipStream_t st) {
	const int64_t n = R.n;
	if (tmp_bytes < key_offsets_tmp_bytes(n)) return hipErrorInvalidValue;
	uint32_t *cnt = (uint32_t *) tmp;
	uint64_t *scan_tmp = (uint64_t *) ((uint8_t *) tmp + align256(((size_t) n + 1) * 4));
	hipError_t e = hipMemsetAsync(cnt + n, 0, 4, st); // the scan's last element is the total
	if (e != hipSuccess) return e;
	if (n > 0) {
		k_key_counts<<<(unsigned) ((n + 255) / 256), 256, 0, st>>>(R, k, all_positions ? 1 : 0, cnt);
		if ((e = hipGetLastError()) != hipSuccess) return e;
	}
	return scan64_rec<uint32_t>(cnt, koff, n + 1, scan_tmp, st);
}

size_t order_tmp_bytes(int64_t n) {
	size_t sort_bytes = 0;
	(void) sort_pairs_u64(nullptr, nullptr, nullptr, nullptr, (uint32_t) n, 32, nullptr, sort_bytes, nullptr);
	const size_t n1 = (size_t) n + 1;
	// keys, values, sorted keys, permutation; word counts, base counts; ordered base offsets;
	// scan scratch; the sort's scratch
	return 4 * align256(n1 * 8) + 2 * align256(n1 * 4) + align256(n1 * 8) + align256(scan64_tmp_words(n) * 8) +
			align256(sort_bytes);
}

hipError_t launch_order(const Packed &R, const uint8_t *active, const uint8_t *flag, OrderOut &out, void *tmp,
		size_t tmp_bytes, hipStream_t st) {
	const int64_t n = R.n;
	if (tmp_bytes < order_tmp_bytes(n) || n >= ((int64_t) 1 << 32)) return hipErrorInvalidValue;
	if (n == 0) return hipSuccess;
	const size_t n1 = (size_t) n + 1;
	uint8_t *t = (uint8_t *) tmp;
	auto take = [&](size_t bytes) {
		uint8_t *p = t;
		t += align256(bytes);
		return p;
	};
	uint64_t *keys = (uint64_t *) take(n1 * 8), *vals = (uint64_t *) take(n1 * 8);
	uint64_t *skeys = (uint64_t *) take(n1 * 8), *perm = (uint64_t *) take(n1 * 8);
	uint32_t *nwords = (uint32_t *) take(n1 * 4), *nbases = (uint32_t *) take(n1 * 4);
	uint64_t *boff_ordered = (uint64_t *) take(n1 * 8);
	uint64_t *scan_tmp = (uint64_t *) take(scan64_tmp_words(n) * 8);
	size_t sort_bytes = 0;
	(void) sort_pairs_u64(nullptr, nullptr, nullptr, nullptr, (uint32_t) n, 32, nullptr, sort_bytes, nullptr);
	void *sort_tmp = take(sort_bytes);
	const unsigned blocks = (unsigned) ((n + 255) / 256);
	k_order_keys<<<blocks, 256, 0, st>>>(R, active, keys, vals);
	hipError_t e = hipGetLastError();
	if (e == hipSuccess) e = sort_pairs_u64(keys, skeys, vals, perm, (uint32_t) n, 32, sort_tmp, sort_bytes, st);
	if (e != hipSuccess) return e;
	k_order_meta<<<blocks, 256, 0, st>>>(R, flag, perm, out.len, out.uid, out.good, nwords, nbases);
	if ((e = hipGetLastError()) != hipSuccess) return e;
	if ((e = scan64_rec<uint32_t>(nwords, out.woff, n, scan_tmp, st)) != hipSuccess) return e;
	if ((e = scan64_rec<uint32_t>(nbases, boff_ordered, n, scan_tmp, st)) != hipSuccess) return e;
	k_order_words<<<(unsigned) ((n + 3) / 4), 256, 0, st>>>(R, perm, out.woff, out.words, boff_ordered, out.t_off);
	if ((e = hipGetLastError()) != hipSuccess) return e;
	if (out.perm) e = hipMemcpyAsync(out.perm, perm, (size_t) n * 8, hipMemcpyDeviceToDevice, st);
	return e;
}

} // namespace hero
