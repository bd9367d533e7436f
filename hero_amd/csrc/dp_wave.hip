// Plain form of the banded DP (BandAligner::bandedAlign / bandedAlignAffineGap,
// src/BandAligner.cpp:56-239): one 64-lane wavefront per DP call.
//
// Band-diagonal coordinates as in dp_core.h: cell (i, j) sits on diagonal d = j - i + b, d in
// [0, 2b]. Anti-diagonal A = i + j = 2i + d - b. The wavefront sweeps A = 0 .. lenS + lenQ; on
// anti-diagonal A exactly the diagonals of parity (A + b) & 1 hold a cell, so each lane keeps two
// diagonals (d = 2·(64c + lane) + s for chunk c and slot s) and updates the one of the current
// parity. A cell reads its diagonal neighbour from its own slot (computed two steps before), its
// up neighbour (i-1, j) from diagonal d+1 and its left neighbour (i, j-1) from diagonal d-1, both
// computed on the previous step: one is the lane's other slot, the other comes from the adjacent
// lane by a cross-lane shuffle. H, X, Y stay in registers.
//
// Traceback pointers (2 bits per cell, the codes of cell_update) are collected per diagonal: a
// lane accumulates 16 consecutive rows of a diagonal in a register and stores the word at
// (i >> 4) * W + d, in LDS when the call's pointers fit the wave's LDS share, else in per-wave
// global scratch. The traceback then walks the pointers, rereading a word only when the path
// leaves it.
//
// Two kinds of calls run through the same sweep: DP tasks of the correction pipeline (2-bit
// packed reads; results as edit words and step summaries) and BandAligner calls over ASCII bytes
// (the C-ABI drop-in; results as CIGAR runs written on the device).
#include <hip/hip_runtime.h>

#include "../../include/hero_gpu.h"
#include "dp_core.h"
#include "kernels.h"

namespace hero {

namespace {

__device__ __forceinline__ int wlane() { return (int) threadIdx.x & 63; }

__device__ __forceinline__ void wstat_add(unsigned long long *stats, int slot, unsigned long long v) {
	atomicAdd(&stats[(blockIdx.x % kStatStripes) * kStatStride + slot], v);
}

// Pointer words of one call: one word per (16-row group, diagonal).
__host__ __device__ __forceinline__ size_t wave_ptr_words(int lenS, int b) {
	return (size_t) ((lenS >> 4) + 1) * (size_t) (2 * b + 1);
}

template <int S>
struct SlotTag {
	static constexpr int value = S;
};

// Forward sweep of one call. sc / qc: the aligner subject's and query's codes (LDS); ptr: the
// call's pointer words (LDS or global).
template <bool AFF, int C>
__device__ __forceinline__ void wave_forward(const Opts &o, int lenS, int lenQ, int b, bool la, bool ra,
		const uint8_t *sc, const uint8_t *qc, uint32_t *ptr) {
	const int lane = wlane();
	const int W = 2 * b + 1;
	int H[C][2], X[C][2], Y[C][2];
	uint32_t acc[C][2];
#pragma unroll
	for (int c = 0; c < C; c++)
#pragma unroll
		for (int s = 0; s < 2; s++) H[c][s] = kSent, X[c][s] = kSent, Y[c][s] = kSent, acc[c][s] = 0;
	// one anti-diagonal; S = the parity of the diagonals that hold its cells
	auto step = [&](int A, auto slot_tag) {
		constexpr int S = decltype(slot_tag)::value;
		constexpr int O = 1 - S;
		int nbH[C], nbV[C];
		{
			int rH[C], rV[C];
#pragma unroll
			for (int c = 0; c < C; c++) {
				// S == 0: left neighbour d - 1 = slot 1 of lane - 1 (lane 0: lane 63 of chunk c - 1)
				// S == 1: up neighbour d + 1 = slot 0 of lane + 1 (lane 63: lane 0 of chunk c + 1)
				const int src = S == 0 ? ((lane + 63) & 63) : ((lane + 1) & 63);
				rH[c] = __shfl(H[c][O], src, 64);
				rV[c] = AFF ? __shfl(S == 0 ? Y[c][O] : X[c][O], src, 64) : kSent;
			}
#pragma unroll
			for (int c = 0; c < C; c++) {
				if (S == 0) {
					nbH[c] = lane != 0 ? rH[c] : (c > 0 ? rH[c > 0 ? c - 1 : 0] : kSent);
					nbV[c] = lane != 0 ? rV[c] : (c > 0 ? rV[c > 0 ? c - 1 : 0] : kSent);
				} else {
					nbH[c] = lane != 63 ? rH[c] : (c + 1 < C ? rH[c + 1 < C ? c + 1 : c] : kSent);
					nbV[c] = lane != 63 ? rV[c] : (c + 1 < C ? rV[c + 1 < C ? c + 1 : c] : kSent);
				}
			}
		}
#pragma unroll
		for (int c = 0; c < C; c++) {
			const int d = 2 * (64 * c + lane) + S;
			const int i2 = A + b - d;
			const int i = i2 >> 1, j = A - i;
			if (d < W && i2 >= 0 && i <= lenS && j >= 0 && j <= lenQ) {
				int h, x, y, p;
				if (i == 0) {
					row0_cell<AFF>(o, j, b, lenQ, h, x, y); // src/BandAligner.cpp:63-111,140-171 at i = 0
					p = 1;
				} else {
					const bool dv = j >= 1;            // (i-1, j-1) is in the matrix (same diagonal)
					const bool uv = d + 1 < W;         // (i-1, j) is in the band
					const bool lv = j >= 1 && d >= 1;  // (i, j-1) is in the band
					const int hd = dv ? H[c][S] : kSent;
					const int xd = dv ? X[c][S] : kSent, yd = dv ? Y[c][S] : kSent;
					const int hu = uv ? (S == 0 ? H[c][O] : nbH[c]) : kSent;
					const int xu = uv ? (S == 0 ? X[c][O] : nbV[c]) : kSent;
					const int hl = lv ? (S == 1 ? H[c][O] : nbH[c]) : kSent;
					const int yl = lv ? (S == 1 ? Y[c][O] : nbV[c]) : kSent;
					cell_update<AFF>(o, sc[i - 1], j >= 1 ? qc[j - 1] : kCodeNone, j == 0, j == lenQ, la, ra, lenQ, hd,
							xd, yd, hu, xu, hl, yl, h, x, y, p);
				}
				H[c][S] = h;
				if (AFF) X[c][S] = x, Y[c][S] = y;
				acc[c][S] |= (uint32_t) p << (2 * (i & 15));
				const int imax = lenS < lenQ + b - d ? lenS : lenQ + b - d; // last row of diagonal d
				if ((i & 15) == 15 || i == imax) {
					ptr[(size_t) (i >> 4) * W + d] = acc[c][S];
					acc[c][S] = 0;
				}
			}
		}
	};
	const int last = lenS + lenQ;
	if ((b & 1) == 0) {
		for (int A = 0; A <= last; A += 2) {
			step(A, SlotTag<0>{});
			if (A + 1 <= last) step(A + 1, SlotTag<1>{});
		}
	} else {
		for (int A = 0; A <= last; A += 2) {
			step(A, SlotTag<1>{});
			if (A + 1 <= last) step(A + 1, SlotTag<0>{});
		}
	}
}

// Traceback (src/BandAligner.cpp:114-125,227-238) over the per-diagonal pointer words: from
// (lenS, lenQ) to the origin, emit(step, i, j) for each step last-to-first with (i, j) the cell it
// leaves. Returns the step count, or -1 if the path leaves the band (never for pointers written by
// wave_forward).
template <class EM>
__device__ __forceinline__ int wave_traceback(int lenS, int lenQ, int b, const uint32_t *ptr, EM &emit) {
	const int W = 2 * b + 1;
	int i = lenS, j = lenQ, n = 0;
	int cg = -1, cdg = -1;
	uint32_t word = 0;
	while (i > 0 || j > 0) {
		const int d = j - i + b;
		if (i < 0 || j < 0 || d < 0 || d >= W) return -1;
		const int g = i >> 4;
		if (g != cg || d != cdg) word = ptr[(size_t) g * W + d], cg = g, cdg = d;
		const int p = (int) ((word >> (2 * (i & 15))) & 3u);
		if (p >= 2) {
			emit(p == 2 ? kStepMatch : kStepMismatch, i, j);
			i--, j--;
		} else if (p == 1) {
			emit(kStepLeft, i, j);
			j--;
		} else {
			emit(kStepUp, i, j);
			i--;
		}
		n++;
	}
	return n;
}

// Where one call's pointer words go: the wave's LDS share when they fit, else its global scratch;
// null when neither holds them.
__device__ __forceinline__ uint32_t *wave_ptr_store(const WaveSizing &z, uint32_t *lds_ptr, size_t words) {
	if (words <= (size_t) z.ptr_lds_words) return lds_ptr;
	if (z.scratch && words <= z.scratch_words) return z.scratch + (size_t) blockIdx.x * z.scratch_words;
	return nullptr;
}

__device__ __forceinline__ uint32_t wave_next(uint32_t *counter) {
	uint32_t t = 0;
	if (wlane() == 0) t = atomicAdd(counter, 1u);
	return __shfl(t, 0, 64);
}

// ---- DP tasks of the correction pipeline ------------------------------------------------------
template <bool AFF, int C>
__global__ __launch_bounds__(64) void k_dp_wave_tasks(DpLaunch L0, WaveSizing z, int band_lo, int band_hi,
		uint32_t *over, uint32_t *over_n) {
	DpLaunch L = L0;
	if (L.range) L.order += L.range[0], L.ntasks = L.range[1] - L.range[0];
	extern __shared__ uint8_t wave_lds[];
	uint8_t *codes = wave_lds;
	uint32_t *lds_ptr = (uint32_t *) (wave_lds + z.seq_cap);
	const int lane = wlane();
	uint64_t cells = 0;
	for (;;) {
		const uint32_t ti = wave_next(L.batch_counter);
		if (ti >= L.ntasks) break;
		const uint32_t t = L.order ? L.order[ti] : ti;
		const DpTask task = L.tasks[t];
		const int b = task.band;
		if (b < band_lo) continue; // a narrower launch's task
		if (b > band_hi) {
			// a wider launch takes it: the first launch lists such tasks for the others; the last
			// launch counts those beyond every range
			if (over && lane == 0) over[atomicAdd(over_n, 1u)] = t;
			if (!over && band_hi >= kWaveMaxBand && lane == 0) wstat_add(L.stats, kStDpSkipped, 1ull);
			continue;
		}
		const bool sw = (task.flags & kTaskSwapped) != 0;
		const int lenS = sw ? task.q_len : task.s_len, lenQ = sw ? task.s_len : task.q_len;
		uint32_t *ptr = wave_ptr_store(z, lds_ptr, wave_ptr_words(lenS, b));
		if (!ptr || lenS + lenQ > z.seq_cap) {
			if (lane == 0) wstat_add(L.stats, kStDpSkipped, 1ull);
			continue;
		}
		// the aligner's subject and query codes in strand frame (reverse complement: code ^ 2)
		{
			const bool src = (task.flags & kTaskSubjRC) != 0, qrc = (task.flags & kTaskQryRC) != 0;
			uint8_t *cs = codes + (sw ? lenS : 0), *cq = codes + (sw ? 0 : lenS);
			for (int p = lane; p < task.s_len; p += 64)
				cs[p] = (uint8_t) segment_code(L.S.words, task.s_word, task.s_bit, p, src);
			for (int p = lane; p < task.q_len; p += 64)
				cq[p] = (uint8_t) segment_code(L.Q.words, task.q_word, task.q_bit, p, qrc);
		}
		__syncthreads();
		const uint8_t *sc = codes, *qc = codes + lenS;
		wave_forward<AFF, C>(L.o, lenS, lenQ, b, (task.flags & kTaskAlignLeft) != 0, (task.flags & kTaskAlignRight) != 0,
				sc, qc, ptr);
		__syncthreads();
		if (lane == 0) {
			StepWriter wr{L.words + task.out_off, task.s_len + task.q_len, -1, 0};
			auto emit = [&](uint32_t st, int, int) { wr(st); };
			const int ns = wave_traceback(lenS, lenQ, b, ptr, emit);
			wr.flush();
			L.nsteps[t] = ns < 0 ? 0u : (uint32_t) ns;
			L.sums[t] = wr.sb.finish();
			if (ns < 0) wstat_add(L.stats, kStDpSkipped, 1ull);
			cells += band_cells(lenS, lenQ, b);
		}
		__syncthreads();
	}
	if (lane == 0 && cells) wstat_add(L.stats, kStCellsAll, (unsigned long long) cells);
}

// ---- BandAligner calls over ASCII bytes (the C-ABI drop-in) -----------------------------------
__device__ __forceinline__ int byte_code(char ch) { return ch == 'N' ? kCodeN : ascii_code(ch); }

// Mismatch ops carry the original query's base letter (src/BandAligner.cpp:272-278,313-318).
__device__ __forceinline__ char code_letter(int c) { return "ACTGN"[c < 4 ? c : 4]; }

// CIGAR runs emitted last-to-first into a call's region [0, cap): written downward from cap, then
// moved to the front. Equal adjacent M / I / D steps merge; every mismatch is its own op of count
// 1 (BandAligner::calculateCigarAndPosition[Reversed], src/BandAligner.cpp:258-419).
struct CigarBackWriter {
	int32_t *cnt;
	char *typ;
	int pos;       // first op written so far
	int run;       // steps in the open op
	char cur;      // its type
	bool cur_mis;  // it is a mismatch (never extended)
	__device__ void push(char t, bool mis) {
		if (run > 0 && t == cur && !mis && !cur_mis) {
			run++;
			return;
		}
		flush();
		cur = t, cur_mis = mis, run = 1;
	}
	__device__ void flush() {
		if (run > 0) {
			pos--;
			cnt[pos] = run;
			typ[pos] = cur;
		}
		run = 0;
	}
};

template <bool AFF, int C>
__global__ __launch_bounds__(64) void k_dp_wave_pairs(WavePairs P, Opts o, WaveSizing z, int band_lo, int band_hi) {
	extern __shared__ uint8_t wave_lds[];
	uint8_t *codes = wave_lds;
	uint32_t *lds_ptr = (uint32_t *) (wave_lds + z.seq_cap);
	const int lane = wlane();
	for (;;) {
		const uint32_t t = wave_next(P.counter);
		if (t >= (uint32_t) P.n) break;
		const hero_band_pair pr = ((const hero_band_pair *) P.pairs)[t];
		// setSequences (src/BandAligner.cpp:559-573): the longer side is the aligner's subject (the
		// query side keeps it on ties); setBand + the widening of m_band (:59-62,144-148)
		int status = HERO_OK;
		if (pr.s_len < 0 || pr.q_len < 0 || pr.s_off < 0 || pr.q_off < 0) status = HERO_ERR_REF_ABORT;
		else if (pr.s_len == 0 && pr.q_len == 0) status = HERO_ERR_ARG;
		const bool sw = !(pr.q_len <= pr.s_len);
		const int lenS = sw ? pr.q_len : pr.s_len, lenQ = sw ? pr.s_len : pr.q_len;
		const int b = pr.band > lenS - lenQ ? pr.band : lenS - lenQ;
		if (status == HERO_OK && (lenS > P.max_len || b > P.max_band || b < 0)) status = HERO_ERR_LIMIT;
		uint32_t *ptr = nullptr;
		if (status == HERO_OK) {
			ptr = wave_ptr_store(z, lds_ptr, wave_ptr_words(lenS, b));
			if (!ptr || lenS + lenQ > z.seq_cap) status = HERO_ERR_LIMIT;
		}
		if (status != HERO_OK) {
			// reported once, by the launch of the narrowest band range
			if (band_lo == 0 && lane == 0) P.cigar_n[t] = status, P.band_out[t] = pr.band;
			continue;
		}
		if (b < band_lo || b > band_hi) continue; // another launch's band range
		const int64_t s_off = sw ? pr.q_off : pr.s_off, q_off = sw ? pr.s_off : pr.q_off;
		for (int p = lane; p < lenS; p += 64) codes[p] = (uint8_t) byte_code(P.seq[s_off + p]);
		for (int p = lane; p < lenQ; p += 64) codes[lenS + p] = (uint8_t) byte_code(P.seq[q_off + p]);
		__syncthreads();
		const uint8_t *sc = codes, *qc = codes + lenS;
		wave_forward<AFF, C>(o, lenS, lenQ, b, pr.flag == 1, pr.flag == 2, sc, qc, ptr);
		__syncthreads();
		if (lane == 0) {
			const int cap = pr.s_len + pr.q_len;
			int32_t *cnt = P.cigar_counts + P.cigar_off[t];
			char *typ = P.cigar_types + P.cigar_off[t];
			CigarBackWriter cw{cnt, typ, cap, 0, 0, false};
			auto emit = [&](uint32_t st, int i, int j) {
				if (st == kStepLeft) cw.push(sw ? 'D' : 'I', false);
				else if (st == kStepUp) cw.push(sw ? 'I' : 'D', false);
				else if (st == kStepMatch) cw.push('M', false);
				else cw.push(code_letter(sw ? sc[i - 1] : qc[j - 1]), true);
			};
			const int ns = wave_traceback(lenS, lenQ, b, ptr, emit);
			cw.flush();
			const int nops = cap - cw.pos;
			for (int k = 0; k < nops; k++) cnt[k] = cnt[cw.pos + k], typ[k] = typ[cw.pos + k];
			P.cigar_n[t] = ns < 0 ? HERO_ERR_HIP : nops;
			P.band_out[t] = b;
		}
		__syncthreads();
	}
}

} // namespace

// Diagonals per lane: 2·C, so one launch takes bands up to 64·C - 1.
constexpr int kWaveChunks[kWaveLaunches] = {1, 4, 16};

int wave_band_hi(int launch) { return 64 * kWaveChunks[launch] - 1; }

size_t wave_lds_bytes(const WaveSizing &z) { return (size_t) z.seq_cap + (size_t) z.ptr_lds_words * 4; }

hipError_t launch_dp_wave_tasks(const DpLaunch &L, const WaveSizing &z, int launch, uint32_t *ctl_wave, uint32_t *over,
		hipStream_t st) {
	if (L.ntasks == 0 && !L.range) return hipSuccess;
	const int lo = launch == 0 ? 0 : wave_band_hi(launch - 1) + 1, hi = wave_band_hi(launch);
	const size_t lds = wave_lds_bytes(z);
	const bool aff = L.o.aligner == 0;
	DpLaunch Lw = L;
	Lw.batch_counter = ctl_wave + kWaveCtlCounter + launch;
	uint32_t *ov = over, *ov_n = ctl_wave + kWaveCtlOver + 1;
	if (launch > 0) {
		// the tasks the first launch listed: range [0, count) of the overflow list
		Lw.order = over, Lw.range = ctl_wave + kWaveCtlOver, Lw.ntasks = 0;
		ov = nullptr, ov_n = nullptr;
	}
#define HERO_WT(I, CC)                                                                                     \
	if (launch == I) {                                                                                     \
		if (aff) k_dp_wave_tasks<true, CC><<<z.nwaves, 64, lds, st>>>(Lw, z, lo, hi, ov, ov_n);            \
		else k_dp_wave_tasks<false, CC><<<z.nwaves, 64, lds, st>>>(Lw, z, lo, hi, ov, ov_n);               \
		return hipGetLastError();                                                                          \
	}
	HERO_WT(0, 1) HERO_WT(1, 4) HERO_WT(2, 16)
#undef HERO_WT
	return hipErrorInvalidValue;
}

hipError_t launch_dp_wave_pairs(const WavePairs &P, const Opts &o, const WaveSizing &z, int launch, hipStream_t st) {
	if (P.n == 0) return hipSuccess;
	const int lo = launch == 0 ? 0 : wave_band_hi(launch - 1) + 1, hi = wave_band_hi(launch);
	const size_t lds = wave_lds_bytes(z);
	const bool aff = o.aligner == 0;
#define HERO_WP(I, CC)                                                                                     \
	if (launch == I) {                                                                                     \
		if (aff) k_dp_wave_pairs<true, CC><<<z.nwaves, 64, lds, st>>>(P, o, z, lo, hi);                    \
		else k_dp_wave_pairs<false, CC><<<z.nwaves, 64, lds, st>>>(P, o, z, lo, hi);                       \
		return hipGetLastError();                                                                          \
	}
	HERO_WP(0, 1) HERO_WP(1, 4) HERO_WP(2, 16)
#undef HERO_WP
	return hipErrorInvalidValue;
}

size_t wave_call_ptr_words(int lenS, int b) { return wave_ptr_words(lenS, b); }

} // namespace hero
