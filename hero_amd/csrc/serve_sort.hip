// Stable radix sort of the serving path's (key, value) u64 pairs on the low key_bits bits of the
// key: least-significant 8-bit digit first. Each pass counts the digits of every 2,048-item tile
// (a 256-bin LDS histogram), scans the digit-major counts into output offsets, and scatters: a tile
// places its items in eight rounds of 256, ranking equal digits inside each wavefront by ballot
// matching and across the block's four wavefronts through LDS, so every digit keeps the tile's
// order and the sort is stable.
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace hero {

namespace {

constexpr int kSortThreads = 256, kSortRounds = 8, kSortTile = kSortThreads * kSortRounds;

__device__ __forceinline__ uint32_t digit_of(uint64_t key, int shift, uint32_t mask) {
	return (uint32_t) (key >> shift) & mask;
}

__global__ __launch_bounds__(256) void k_sort_hist(const uint64_t *keys, uint32_t n, int shift, uint32_t mask,
		uint32_t ntiles, uint32_t *hist) {
	__shared__ uint32_t h[256];
	h[threadIdx.x] = 0;
	__syncthreads();
	const uint64_t base = (uint64_t) blockIdx.x * kSortTile;
	for (int r = 0; r < kSortRounds; r++) {
		const uint64_t i = base + (uint64_t) r * kSortThreads + threadIdx.x;
		if (i < n) atomicAdd(&h[digit_of(keys[i], shift, mask)], 1u);
	}
	__syncthreads();
	hist[(size_t) threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x]; // digit-major: [digit][tile]
}

__global__ __launch_bounds__(256) void k_sort_scatter(const uint64_t *kin, const uint64_t *vin, uint64_t *kout,
		uint64_t *vout, uint32_t n, int shift, uint32_t mask, uint32_t ntiles, const uint32_t *offs) {
	__shared__ uint32_t run[256];     // next output position of each digit in this tile
	__shared__ uint32_t wcnt[4][256]; // this digit pass: items of each digit per wavefront
	const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
	const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
	run[tid] = offs[(size_t) tid * ntiles + blockIdx.x];
	const uint64_t base = (uint64_t) blockIdx.x * kSortTile;
	for (int r = 0; r < kSortRounds; r++) {
		for (int w = 0; w < 4; w++) wcnt[w][tid] = 0;
		__syncthreads();
		const uint64_t i = base + (uint64_t) r * kSortThreads + tid;
		const bool valid = i < n;
		const uint64_t key = valid ? kin[i] : 0, val = valid ? vin[i] : 0;
		const uint32_t d = valid ? digit_of(key, shift, mask) : 0u;
		// the lanes of this wavefront holding an item with the same digit
		uint64_t peers = __ballot(valid);
#pragma unroll
		for (int b = 0; b < 8; b++) {
			const bool bit = (d >> b) & 1u;
			const uint64_t bal = __ballot(bit);
			peers &= bit ? bal : ~bal;
		}
		const uint32_t rank = (uint32_t) __popcll(peers & below);
		if (valid && rank == 0) wcnt[wv][d] = (uint32_t) __popcll(peers);
		__syncthreads();
		if (valid) {
			uint32_t pos = run[d] + rank;
			for (int w = 0; w < wv; w++) pos += wcnt[w][d];
			kout[pos] = key;
			vout[pos] = val;
		}
		__syncthreads();
		run[tid] += wcnt[0][tid] + wcnt[1][tid] + wcnt[2][tid] + wcnt[3][tid];
		__syncthreads();
	}
}

size_t align256(size_t b) { return (b + 255) & ~(size_t) 255; }

} // namespace

// tmp layout: a second key and value array (ping-pong), the tiles' digit counts, their scan, the
// scan's scratch and total.
hipError_t sort_pairs_u64(const uint64_t *keys_in, uint64_t *keys_out, const uint64_t *vals_in, uint64_t *vals_out,
		uint32_t n, int key_bits, void *tmp, size_t &tmp_bytes, hipStream_t st) {
	const uint32_t ntiles = (uint32_t) ((n + kSortTile - 1) / kSortTile);
	const size_t counts = (size_t) 256 * ntiles;
	const size_t b_keys = align256((size_t) n * 8), b_hist = align256(counts * 4);
	const size_t b_scan = align256(scan_tmp_words((int64_t) counts) * 4);
	const size_t need = 2 * b_keys + 2 * b_hist + b_scan + 256;
	if (!tmp) {
		tmp_bytes = need;
		return hipSuccess;
	}
	if (tmp_bytes < need) return hipErrorInvalidValue;
	if (n == 0) return hipSuccess;
	const int passes = (key_bits + 7) / 8;
	if (passes == 0) {
		hipError_t e = hipMemcpyAsync(keys_out, keys_in, (size_t) n * 8, hipMemcpyDeviceToDevice, st);
		if (e == hipSuccess) e = hipMemcpyAsync(vals_out, vals_in, (size_t) n * 8, hipMemcpyDeviceToDevice, st);
		return e;
	}
	uint8_t *t = (uint8_t *) tmp;
	uint64_t *pk = (uint64_t *) t, *pv = (uint64_t *) (t + b_keys);
	uint32_t *hist = (uint32_t *) (t + 2 * b_keys), *offs = (uint32_t *) (t + 2 * b_keys + b_hist);
	uint32_t *scan_tmp = (uint32_t *) (t + 2 * b_keys + 2 * b_hist);
	unsigned long long *total = (unsigned long long *) (t + 2 * b_keys + 2 * b_hist + b_scan);
	const uint64_t *sk = keys_in, *sv = vals_in;
	for (int p = 0; p < passes; p++) {
		// the last pass writes the caller's output
		const bool to_out = ((passes - 1 - p) % 2) == 0;
		uint64_t *dk = to_out ? keys_out : pk, *dv = to_out ? vals_out : pv;
		const int shift = 8 * p;
		const int bits = key_bits - shift < 8 ? key_bits - shift : 8;
		const uint32_t mask = (1u << bits) - 1u;
		k_sort_hist<<<ntiles, kSortThreads, 0, st>>>(sk, n, shift, mask, ntiles, hist);
		hipError_t e = hipGetLastError();
		if (e == hipSuccess) e = scan_u32(hist, offs, (int64_t) counts, total, scan_tmp, b_scan / 4, st);
		if (e != hipSuccess) return e;
		k_sort_scatter<<<ntiles, kSortThreads, 0, st>>>(sk, sv, dk, dv, n, shift, mask, ntiles, offs);
		if ((e = hipGetLastError()) != hipSuccess) return e;
		sk = dk, sv = dv;
	}
	return hipSuccess;
}

} // namespace hero
