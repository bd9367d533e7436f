// Two banded DP tasks per lane in packed 16-bit arithmetic (the "two pairs per lane" tile of the
// batched form). Same recurrences, pointers and tie orders as dp_forward (dp_core.h); each 32-bit
// register holds task A's value in its low half and task B's in its high half, so one packed
// add / max / select does the work of two cells.
//
// Range. With s the largest step cost (|hit|, |miss|, |dx|, |dy|, |ex|, |sx|), every real value of
// a task lies in [-B, B] for B = s * (lenS + lenQ + 2), and every value derived from the sentinel
// lies in [kSent16 - B, kSent16 + B]. For B < kPk16Bound both ranges fit in int16 and stay apart,
// and every comparison has the outcome it has in dp_forward, where kSent = -2^28 (differences
// between two sentinel-derived values do not depend on the sentinel's value). Tasks beyond the
// bound run in dp_forward.
#pragma once
#include <type_traits>

#include "dp_core.h"

namespace hero {

constexpr int kSent16 = -16384;
constexpr int kPk16Bound = 8192;

// Largest |step cost| of the options: the per-cell bound of the range argument above.
HD int max_step_cost(const Opts &o) {
	int m = 0;
	const int c[7] = {o.hit, o.miss, o.dx, o.dy, o.ex, o.sx, o.unobserved};
	for (int i = 0; i < 7; i++) m = imax_(m, c[i] < 0 ? -c[i] : c[i]);
	return m;
}
HD bool pk16_fits(const Opts &o, int lenS, int lenQ) {
	return (int64_t) max_step_cost(o) * (int64_t) (lenS + lenQ + 2) < kPk16Bound;
}
// Every task of a batched band class fits: both sides are shorter than rows_max.
HD bool pk16_class_fits(const Opts &o, int rows_max) {
	return (int64_t) max_step_cost(o) * (int64_t) (2 * rows_max) < kPk16Bound;
}

struct I16x2 {
	uint32_t v;
};
HD I16x2 pk2(int a, int b) { return I16x2{(uint32_t) (uint16_t) (int16_t) a | ((uint32_t) (uint16_t) (int16_t) b << 16)}; }
HD int pk_lo(I16x2 a) { return (int) (int16_t) (uint16_t) (a.v & 0xFFFFu); }
HD int pk_hi(I16x2 a) { return (int) (int16_t) (uint16_t) (a.v >> 16); }

#if defined(__HIP_DEVICE_COMPILE__)
typedef short s16x2_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ s16x2_t pk_v(I16x2 a) { return __builtin_bit_cast(s16x2_t, a.v); }
__device__ __forceinline__ I16x2 pk_w(s16x2_t v) { return I16x2{__builtin_bit_cast(uint32_t, v)}; }
__device__ __forceinline__ I16x2 pk_add(I16x2 a, I16x2 b) { return pk_w(pk_v(a) + pk_v(b)); }
__device__ __forceinline__ I16x2 pk_sub(I16x2 a, I16x2 b) { return pk_w(pk_v(a) - pk_v(b)); }
__device__ __forceinline__ I16x2 pk_max(I16x2 a, I16x2 b) { return pk_w(__builtin_elementwise_max(pk_v(a), pk_v(b))); }
// 0xFFFF in each half that is negative
__device__ __forceinline__ uint32_t pk_neg(I16x2 a) {
	const s16x2_t s = {15, 15};
	return __builtin_bit_cast(uint32_t, pk_v(a) >> s);
}
#else
HD I16x2 pk_add(I16x2 a, I16x2 b) { return pk2(pk_lo(a) + pk_lo(b), pk_hi(a) + pk_hi(b)); }
HD I16x2 pk_sub(I16x2 a, I16x2 b) { return pk2(pk_lo(a) - pk_lo(b), pk_hi(a) - pk_hi(b)); }
HD I16x2 pk_max(I16x2 a, I16x2 b) { return pk2(imax_(pk_lo(a), pk_lo(b)), imax_(pk_hi(a), pk_hi(b))); }
HD uint32_t pk_neg(I16x2 a) { return (pk_lo(a) < 0 ? 0xFFFFu : 0u) | (pk_hi(a) < 0 ? 0xFFFF0000u : 0u); }
#endif
// a < b in each half (no overflow: both operands stay within the range above)
HD uint32_t pk_lt(I16x2 a, I16x2 b) { return pk_neg(pk_sub(a, b)); }
HD I16x2 pk_sel(uint32_t m, I16x2 a, I16x2 b) { return I16x2{(a.v & m) | (b.v & ~m)}; }
// per-half mask from bit d of two W-bit masks
template <class M>
HD uint32_t pk_bitmask(M a, M b, int d) {
	return ((uint32_t) -(int32_t) ((a >> d) & 1u) & 0xFFFFu) | ((uint32_t) -(int32_t) ((b >> d) & 1u) << 16);
}
template <class M>
HD M bits_range(int lo, int hi) { // bits lo..hi (empty when hi < lo), lo >= 0
	if (hi < lo) return 0;
	const M upto = (hi + 1 >= (int) (8 * sizeof(M))) ? ~(M) 0 : (((M) 1 << (hi + 1)) - 1);
	return upto & ~(((M) 1 << lo) - 1);
}
template <class M>
HD M bit_at(int d, int w) { return (d >= 0 && d < w) ? ((M) 1 << d) : 0; }

// One task of a lane pair: geometry and flags.
struct Pk16Task {
	int lenS, lenQ, b;
	bool left_al, right_al;
};

// Forward pass of tasks A (low halves) and B (high halves); rows 1..max(lenS) (a task past its
// last row computes values nobody reads). SA/QA/SB/QB stream the aligner subject / query codes
// (2-bit bases); pa / pb receive each task's pointer rows as dp_forward writes them for every
// cell inside the matrix.
//
// Rows follow dp_forward: every band cell takes the interior recurrence and nothing outside the
// matrix is cleared (no cell of the matrix reads one outside it, except column 0, whose rule
// reads only the cell above); column 0 takes its closed-form values, the aligned column lenQ its
// vertical cost, and cells past a band edge 2b < 2 * BMAX are cut from the left chain. Values
// beyond the matrix are never read, so only the cells of the matrix and the sentinel-derived
// cells past the band edge need the range argument above (sentinel-derived values drift by at
// most s per row: with both tasks within the bound they stay below every real value).
//
// Pointers are built two cells at a time: the 2-bit codes of both tasks sit in the halves of one
// word (8 cells per half), and each row is split into the tasks' own 16-cell words at its end.
template <int BMAX, bool AFFINE, class SS, class QS, class PA, class PB>
HD void dp_forward_pk16(const Opts &o, const Pk16Task &A, const Pk16Task &B, SS &SA, QS &QA, SS &SB, QS &QB, PA &pa,
		PB &pb) {
	static_assert(QS::kTwoBit && SS::kTwoBit, "packed DP reads 2-bit codes");
	constexpr int W = 2 * BMAX + 1;
	constexpr int NW = (W + 15) / 16; // pointer words per task row
	constexpr int NP = (W + 7) / 8;   // pointer words per row, both tasks
	static_assert(W <= 32, "packed query windows hold 32 cells");
	using M = uint32_t;
	const I16x2 K2 = pk2(kSent16, kSent16);
	const I16x2 hit2 = pk2(o.hit, o.hit), miss2 = pk2(o.miss, o.miss);
	const I16x2 dx2 = pk2(o.dx, o.dx), dy2 = pk2(o.dy, o.dy), ex2 = pk2(o.ex, o.ex), sx2 = pk2(o.sx, o.sx);
	const int g0a = (A.right_al || (A.lenQ == 0 && A.left_al)) ? o.sx : o.dx;
	const int g0b = (B.right_al || (B.lenQ == 0 && B.left_al)) ? o.sx : o.dx;
	const M band_a = bits_range<M>(0, 2 * A.b), band_b = bits_range<M>(0, 2 * B.b);
	I16x2 H[W + 1];
	I16x2 X[AFFINE ? W + 1 : 1], Y[AFFINE ? W + 1 : 1];
	// query windows of the current row, 2 bits per cell (QWin of dp_core.h): cell d reads
	// Q[i - b - 1 + d] of its task
	QWin<W, true> wa, wb;
	{ // row 0
		uint32_t pwa[NW], pwb[NW];
#pragma unroll
		for (int w = 0; w < NW; w++) pwa[w] = 0, pwb[w] = 0;
#pragma unroll
		for (int d = 0; d <= W; d++) {
			int ha, xa, ya, hb, xb, yb;
			const bool ina = row0_cell<AFFINE>(o, d - A.b, A.b, A.lenQ, ha, xa, ya) && d < W;
			const bool inb = row0_cell<AFFINE>(o, d - B.b, B.b, B.lenQ, hb, xb, yb) && d < W;
			H[d] = pk2(ina ? ha : kSent16, inb ? hb : kSent16);
			if (AFFINE) {
				X[d] = pk2(ina ? xa : kSent16, inb ? xb : kSent16);
				Y[d] = pk2(ina ? ya : kSent16, inb ? yb : kSent16);
			}
			if (d < W && ina && d - A.b >= 1) pwa[d >> 4] |= 1u << (2 * (d & 15));
			if (d < W && inb && d - B.b >= 1) pwb[d >> 4] |= 1u << (2 * (d & 15));
		}
		pa.put_row(0, pwa);
		pb.put_row(0, pwb);
	}
#pragma unroll
	for (int d = 0; d < W; d++) {
		wa.set(d, d < A.b ? kCodeNone : QA.next());
		wb.set(d, d < B.b ? kCodeNone : QB.next());
	}
	constexpr int kC0 = 1, kCQ = 2, kCut = 4;
	// MODE is wave-uniform and known per row only: one copy of the row code, its edge parts behind
	// uniform branches (separate copies per variant would cost registers, not instructions)
	auto row = [&](int i, const int MODE) {
		const uint64_t ma = wa.differ_bits(SA.next()), mb = wb.differ_bits(SB.next());
		M c0a = 0, c0b = 0, cqa = 0, cqb = 0;
		I16x2 x0 = K2, y0 = K2;
		if (MODE & kC0) {
			c0a = bit_at<M>(A.b - i, W), c0b = bit_at<M>(B.b - i, W);
			const int xa = AFFINE ? (A.right_al ? i * o.sx : o.dx + (i - 1) * o.ex) : i * g0a;
			const int xb = AFFINE ? (B.right_al ? i * o.sx : o.dx + (i - 1) * o.ex) : i * g0b;
			x0 = pk2(xa, xb);
			y0 = pk2(A.right_al ? o.dy : xa, B.right_al ? o.dy : xb);
		}
		if (MODE & kCQ) {
			cqa = A.left_al ? bit_at<M>(A.lenQ - i + A.b, W) : 0;
			cqb = B.left_al ? bit_at<M>(B.lenQ - i + B.b, W) : 0;
		}
		uint32_t pp[NP];
#pragma unroll
		for (int w = 0; w < NP; w++) pp[w] = 0;
		I16x2 hl = K2, yl = K2;
#pragma unroll
		for (int d = 0; d < W; d++) {
			// 1 in each half whose codes differ
			const I16x2 m01{((uint32_t) (ma >> (2 * d)) & 1u) | (((uint32_t) (mb >> (2 * d)) & 1u) << 16)};
			const I16x2 s = pk_sel(m01.v * 0xFFFFu, miss2, hit2); // 0xFFFF in each differing half
			const uint32_t pdiag = 0x00020002u | m01.v;
			I16x2 h;
			uint32_t p;
			if (!AFFINE) { // src/BandAligner.cpp:63-111
				I16x2 g = dx2;
				if (MODE & kCQ) g = pk_sel(pk_bitmask(cqa, cqb, d), sx2, dx2);
				const I16x2 up = pk_add(H[d + 1], g), lf = pk_add(hl, dy2), dg = pk_add(H[d], s);
				h = pk_max(pk_max(dg, lf), up);
				const uint32_t ndiag = pk_lt(dg, lf) | pk_lt(dg, up);
				const uint32_t pleft = ~pk_lt(lf, up) & 0x00010001u;
				p = (ndiag & pleft) | (~ndiag & pdiag);
				if (MODE & kC0) { // column 0: the vertical move only
					const uint32_t c0 = pk_bitmask(c0a, c0b, d);
					h = pk_sel(c0, x0, h);
					p &= ~c0;
				}
				if (MODE & kCut) h = pk_sel(pk_bitmask(band_a, band_b, d), h, K2);
				H[d] = h;
				hl = h;
			} else { // src/BandAligner.cpp:150-222
				I16x2 y = pk_max(pk_add(hl, dy2), pk_add(yl, ex2));
				if (MODE & kCut) y = pk_sel(pk_bitmask(band_a, band_b, d), y, K2);
				I16x2 x = pk_max(pk_add(H[d + 1], dx2), pk_add(X[d + 1], ex2));
				if (MODE & kCQ) x = pk_sel(pk_bitmask(cqa, cqb, d), pk_add(pk_max(H[d + 1], X[d + 1]), sx2), x);
				h = pk_add(pk_max(pk_max(H[d], X[d]), Y[d]), s);
				const uint32_t diag = pk_lt(x, h) & ~pk_lt(h, y); // h > x && h >= y
				p = (diag & pdiag) | (~diag & pk_lt(x, y) & 0x00010001u);
				if (MODE & kC0) { // column 0 (src/BandAligner.cpp:150-163)
					const uint32_t c0 = pk_bitmask(c0a, c0b, d);
					x = pk_sel(c0, x0, x);
					y = pk_sel(c0, y0, y);
					h = pk_sel(c0, x0, h);
					p &= ~c0;
				}
				H[d] = h, X[d] = x, Y[d] = y;
				hl = h, yl = y;
			}
			pp[d >> 3] |= p << (2 * (d & 7));
		}
		// split into each task's words: cells 16w..16w+15 from the halves of pp[2w], pp[2w+1]
		uint32_t pwa[NW], pwb[NW];
#pragma unroll
		for (int w = 0; w < NW; w++) {
			const uint32_t lo = pp[2 * w], hi = 2 * w + 1 < NP ? pp[2 * w + 1] : 0u;
			pwa[w] = (lo & 0xFFFFu) | (hi << 16);
			pwb[w] = (lo >> 16) | (hi & 0xFFFF0000u);
		}
		pa.put_row(i, pwa);
		pb.put_row(i, pwb);
		// slide the query windows to row i+1
		wa.slide((i - A.b + W - 1 >= 0) ? QA.next() : kCodeNone);
		wb.slide((i - B.b + W - 1 >= 0) ? QB.next() : kCodeNone);
	};
	const int rows = imax_(A.lenS, B.lenS);
	const int cut = wave_any(A.b < BMAX || B.b < BMAX) ? kCut : 0;
	for (int i = 1; i <= rows; i++) {
		const bool c0 = i <= A.b || i <= B.b;
		const bool cq = (A.left_al && A.lenQ - i <= A.b) || (B.left_al && B.lenQ - i <= B.b);
		const int mode = cut | ((wave_any(c0) || wave_any(cq)) ? kC0 | kCQ : 0);
		row(i, mode);
	}
}

// ---- tagged rows, two tasks per lane ---------------------------------------------------------
// dp_forward_tagged's recurrence (dp_core.h) in packed 16-bit halves: task A in the low half, B in
// the high half, one packed add / max / mad / bit operation per cell for both tasks.
//
// Range. The scaled value of a cell is V = 8 v + t - 4 hit (i + j). With hit >= 0, miss <= hit and
// every gap cost (dx == dy, ex, sx) <= 0, a cell's score is at most hit * min(i, j), so V <= 3.
// Every state of an in-band cell of the matrix is at least the score of a simple path (down its
// diagonal from row 0 or column 0, whose closed forms are >= -s (b + 1), s the largest |cost|)
// plus one gap move: v >= min(miss, 0) (rows + 1) - s (BMAX + 3) for rows >= max(i, j). So
// V >= L(rows) = 8 (min(miss, 0) (rows + 1) - s (BMAX + 3)) - 8 hit (rows + BMAX) - 4. Values
// derived from the sentinel kSent16T never rise above kSent16T + 3 (every constant added is <= 0)
// and never fall below kSent16T - D, D = |C| + |E| + |DM| + 8: a cut cell's y is the sentinel and
// its max3 takes it, so its values restart near the sentinel every row instead of drifting. With
// L(rows) > kSent16T + 3 and kSent16T - D > -32768 every comparison that involves a cell of the
// matrix has the outcome it has in 32 bits. Cells outside the matrix may wrap: nothing reads them.
constexpr int kSent16T = -24576;

HD bool pk16t_fits(const Opts &o, int rows, int bmax) {
	if (o.hit < 0 || o.miss > o.hit || o.dx != o.dy || o.dx > 0 || o.ex > 0 || o.sx > 0) return false;
	int s = o.hit;
	const int c[5] = {o.dx, o.ex, o.sx, o.miss, o.dy};
	for (int k = 0; k < 5; k++) s = imax_(s, c[k] < 0 ? -c[k] : c[k]);
	const int64_t miss0 = o.miss < 0 ? o.miss : 0;
	const int64_t L = 8 * (miss0 * (rows + 1) - (int64_t) s * (bmax + 3)) - 8 * (int64_t) o.hit * (rows + bmax) - 4;
	const int64_t D = (int64_t) 8 * -o.dx + 4 * o.hit + 8 * -o.ex + 4 * o.hit + 8 * (o.hit - o.miss) + 1 + 8;
	return L > kSent16T + 3 && kSent16T - D > -32768;
}
// every call of a band class fits: both sides are shorter than rows_max
HD bool pk16t_class_fits(const Opts &o, int rows_max, int bmax) { return pk16t_fits(o, rows_max, bmax); }

// Bit planes of a task's query window: bit d of lo / hi = the low / high bit of cell d's code.
template <int W>
struct QPlanes {
	static_assert(W <= 32, "bit planes hold 32 cells");
	uint32_t lo = 0, hi = 0;
	HD void set(int d, int c) { lo |= (uint32_t) (c & 1) << d, hi |= (uint32_t) ((c >> 1) & 1) << d; }
	HD void slide(int next) {
		lo = (lo >> 1) | ((uint32_t) (next & 1) << (W - 1));
		hi = (hi >> 1) | ((uint32_t) ((next >> 1) & 1) << (W - 1));
	}
	// bit d set where cell d's code differs from sc
	HD uint32_t differ(int sc) const {
		return (lo ^ (0u - (uint32_t) (sc & 1))) | (hi ^ (0u - (uint32_t) ((sc >> 1) & 1)));
	}
};

// c + m * a in each half, m in {0, 1}
#if defined(__HIP_DEVICE_COMPILE__)
__device__ __forceinline__ I16x2 pk_mad01(I16x2 m, I16x2 a, I16x2 c) { return pk_w(pk_v(m) * pk_v(a) + pk_v(c)); }
#else
HD I16x2 pk_mad01(I16x2 m, I16x2 a, I16x2 c) { return pk2(pk_lo(c) + pk_lo(m) * pk_lo(a), pk_hi(c) + pk_hi(m) * pk_hi(a)); }
#endif
HD I16x2 pk_max3(I16x2 a, I16x2 b, I16x2 c) { return pk_max(pk_max(a, b), c); }
HD I16x2 pk_or(I16x2 a, uint32_t m) { return I16x2{a.v | m}; }

// Forward pass of tasks A and B in tagged packed rows; the pointers each task receives are
// dp_forward_tagged's (kPtrTaggedAffine). mode_or: row variants taken on every row.
template <int BMAX, class SS, class QS, class PA, class PB>
HD void dp_forward_pk16_tagged(const Opts &o, const Pk16Task &A, const Pk16Task &B, SS &SA, QS &QA, SS &SB, QS &QB,
		PA &pa, PB &pb, int mode_or = 0) {
	static_assert(SS::kTwoBit && QS::kTwoBit, "the tagged form reads 2-bit codes");
	constexpr int W = 2 * BMAX + 1;
	static_assert(W <= 32, "bit planes hold 32 cells");
	constexpr int NW = (W + 15) / 16;
	constexpr int kLast = W - 16 * (NW - 1);
	const int step = 4 * o.hit;
	const int C = 8 * o.dx - step, E = 8 * o.ex - step, DM = 8 * (o.miss - o.hit) - 1;
	const int CQx = 8 * (o.sx - o.dx), CQe = 8 * (o.sx - o.ex);
	const I16x2 S2 = pk2(kSent16T, kSent16T);
	const I16x2 Cm2 = pk2(C - 2, C - 2), E2 = pk2(E, E), DM2 = pk2(DM, DM);
	const I16x2 CQx2 = pk2(CQx, CQx), CQe2 = pk2(CQe, CQe);
	I16x2 MK[W + 1], U[W + 1];
	I16x2 Gq = S2, XEq = S2; // G and XE of the cell above the next row's column lenQ, per task
	QPlanes<W> wa, wb;
	{ // row 0 (src/BandAligner.cpp:140-171): pointers 0 (left) everywhere
		uint32_t pw[NW];
#pragma unroll
		for (int w = 0; w < NW; w++) pw[w] = 0;
		int gqa = kSent16T, xqa = kSent16T, gqb = kSent16T, xqb = kSent16T;
#pragma unroll
		for (int d = 0; d <= W; d++) {
			int mk[2], u[2];
			const Pk16Task *T[2] = {&A, &B};
			for (int k = 0; k < 2; k++) {
				int h, x, y;
				const int j = d - T[k]->b;
				const bool in = row0_cell<true>(o, j, T[k]->b, T[k]->lenQ, h, x, y) && d < W;
				const int off = step * j;
				const int g = in ? 8 * h - off + C : kSent16T, xe = in ? 8 * x + 3 - off + E : kSent16T;
				mk[k] = in ? 8 * max3_(h, x, y) - off + 2 : kSent16T;
				u[k] = in ? imax_(g, xe) | 3 : kSent16T;
				if (T[k]->left_al && d == T[k]->lenQ + T[k]->b) (k ? gqb : gqa) = g, (k ? xqb : xqa) = xe;
			}
			MK[d] = pk2(mk[0], mk[1]);
			U[d] = pk2(u[0], u[1]);
		}
		Gq = pk2(gqa, gqb), XEq = pk2(xqa, xqb);
		pa.put_row(0, pw);
		pb.put_row(0, pw);
	}
#pragma unroll
	for (int d = 0; d < W; d++) {
		wa.set(d, d < A.b ? kCodeNone : QA.next());
		wb.set(d, d < B.b ? kCodeNone : QB.next());
	}
	constexpr int kC0 = 1, kCQ = 2, kCut = 4;
	using M = uint32_t;
	const M band_a = bits_range<M>(0, 2 * A.b), band_b = bits_range<M>(0, 2 * B.b);
	auto row = [&](int i, auto mode_tag) {
		constexpr int MODE = decltype(mode_tag)::value;
		// mismatch bits of both tasks: cells 0..15 in lo16 (A at bit d, B at bit 16 + d), 16.. in hi16
		const uint32_t da = wa.differ(SA.next()), db = wb.differ(SB.next());
		const uint32_t lo16 = (da & 0xFFFFu) | (db << 16), hi16 = (da >> 16) | (db & 0xFFFF0000u);
		M c0a = 0, c0b = 0, cqa = 0, cqb = 0;
		I16x2 h0 = S2, x0 = S2, y0 = S2, g0 = S2;
		if (MODE & kC0) {
			c0a = bit_at<M>(A.b - i, W), c0b = bit_at<M>(B.b - i, W);
			const int off = step * i;
			const int xa = A.right_al ? i * o.sx : o.dx + (i - 1) * o.ex;
			const int xb = B.right_al ? i * o.sx : o.dx + (i - 1) * o.ex;
			const int ya = A.right_al ? o.dy : xa, yb = B.right_al ? o.dy : xb;
			h0 = pk2(8 * xa + 2 - off, 8 * xb + 2 - off), x0 = pk2(8 * xa + 3 - off, 8 * xb + 3 - off);
			y0 = pk2(8 * ya - off, 8 * yb - off), g0 = pk2(8 * xa - off + C, 8 * xb - off + C);
		}
		if (MODE & kCQ) {
			cqa = A.left_al ? bit_at<M>(A.lenQ - i + A.b, W) : 0;
			cqb = B.left_al ? bit_at<M>(B.lenQ - i + B.b, W) : 0;
		}
		uint32_t pwa[NW], pwb[NW];
#pragma unroll
		for (int w = 0; w < NW; w++) pwa[w] = 0, pwb[w] = 0;
		I16x2 gl = S2, yel = S2, gq = Gq, xeq = XEq;
#pragma unroll
		for (int d = 0; d < W; d++) {
			const I16x2 m{((d < 16 ? lo16 : hi16) >> (d & 15)) & 0x00010001u};
			I16x2 h = pk_mad01(m, DM2, MK[d]);
			I16x2 x = U[d + 1];
			I16x2 y = pk_max(gl, yel);
			if (MODE & kCQ) {
				const uint32_t cq = pk_bitmask(cqa, cqb, d);
				x = pk_sel(cq, pk_or(pk_max(pk_add(Gq, CQx2), pk_add(XEq, CQe2)), 0x00030003u), x);
			}
			if (MODE & kCut) y = pk_sel(pk_bitmask(band_a, band_b, d), y, S2);
			I16x2 g = pk_add(pk_add(h, m), Cm2);
			uint32_t tagv;
			I16x2 mx;
			if (MODE & kC0) {
				const uint32_t c0 = pk_bitmask(c0a, c0b, d);
				h = pk_sel(c0, h0, h), x = pk_sel(c0, x0, x), y = pk_sel(c0, y0, y), g = pk_sel(c0, g0, g);
				mx = pk_max3(h, x, y);
				tagv = mx.v | (c0 & 0x00030003u); // column 0 points up
			} else {
				mx = pk_max3(h, x, y);
				tagv = mx.v;
			}
			pwa[d >> 4] = (pwa[d >> 4] >> 2) | (tagv << 30);
			pwb[d >> 4] = (pwb[d >> 4] >> 2) | ((tagv >> 16) << 30);
			MK[d] = I16x2{(mx.v & 0xFFFCFFFCu) | 0x00020002u};
			const I16x2 xe = pk_add(x, E2);
			U[d] = pk_or(pk_max(g, xe), 0x00030003u);
			if (MODE & kCQ) {
				const uint32_t cq = pk_bitmask(cqa, cqb, d);
				gq = pk_sel(cq, g, gq), xeq = pk_sel(cq, xe, xeq);
			}
			gl = g, yel = pk_add(y, E2);
		}
		if (MODE & kCQ) Gq = gq, XEq = xeq;
		if (kLast < 16) pwa[NW - 1] >>= 2 * (16 - kLast), pwb[NW - 1] >>= 2 * (16 - kLast);
		pa.put_row(i, pwa);
		pb.put_row(i, pwb);
		wa.slide((i - A.b + W - 1 >= 0) ? QA.next() : kCodeNone);
		wb.slide((i - B.b + W - 1 >= 0) ? QB.next() : kCodeNone);
	};
	const int rows = imax_(A.lenS, B.lenS);
	// Head rows 1..BMAX when both tasks of every lane have band BMAX and no aligned column lenQ in
	// them: row I's column 0 is band cell BMAX - I in both halves, so the row starts there with
	// the column-0 values and takes the interior rule after it (as dp_forward_tagged's head rows).
	auto head = [&](auto itag) {
		constexpr int I = decltype(itag)::value + 1;
		constexpr int D0 = BMAX - I;
		if (I > rows) return;
		const uint32_t da = wa.differ(SA.next()), db = wb.differ(SB.next());
		const uint32_t lo16 = (da & 0xFFFFu) | (db << 16), hi16 = (da >> 16) | (db & 0xFFFF0000u);
		const int off = step * I;
		const int xa = A.right_al ? I * o.sx : o.dx + (I - 1) * o.ex;
		const int xb = B.right_al ? I * o.sx : o.dx + (I - 1) * o.ex;
		const int ya = A.right_al ? o.dy : xa, yb = B.right_al ? o.dy : xb;
		uint32_t pwa[NW], pwb[NW];
#pragma unroll
		for (int w = 0; w < NW; w++) pwa[w] = 0, pwb[w] = 0;
		I16x2 gl, yel;
		{ // column 0: a vertical gap from the origin, pointing up
			const I16x2 h0 = pk2(8 * xa + 2 - off, 8 * xb + 2 - off), x0 = pk2(8 * xa + 3 - off, 8 * xb + 3 - off);
			const I16x2 y0 = pk2(8 * ya - off, 8 * yb - off), g0 = pk2(8 * xa - off + C, 8 * xb - off + C);
			const I16x2 mx = pk_max3(h0, x0, y0);
			pwa[D0 >> 4] = (pwa[D0 >> 4] >> 2) | (3u << 30);
			pwb[D0 >> 4] = (pwb[D0 >> 4] >> 2) | (3u << 30);
			MK[D0] = I16x2{(mx.v & 0xFFFCFFFCu) | 0x00020002u};
			U[D0] = pk_or(pk_max(g0, pk_add(x0, E2)), 0x00030003u);
			gl = g0, yel = pk_add(y0, E2);
		}
#pragma unroll
		for (int d = D0 + 1; d < W; d++) {
			const I16x2 m{((d < 16 ? lo16 : hi16) >> (d & 15)) & 0x00010001u};
			const I16x2 h = pk_mad01(m, DM2, MK[d]);
			const I16x2 x = U[d + 1];
			const I16x2 y = pk_max(gl, yel);
			const I16x2 g = pk_add(pk_add(h, m), Cm2);
			const I16x2 mx = pk_max3(h, x, y);
			pwa[d >> 4] = (pwa[d >> 4] >> 2) | (mx.v << 30);
			pwb[d >> 4] = (pwb[d >> 4] >> 2) | ((mx.v >> 16) << 30);
			MK[d] = I16x2{(mx.v & 0xFFFCFFFCu) | 0x00020002u};
			U[d] = pk_or(pk_max(g, pk_add(x, E2)), 0x00030003u);
			gl = g, yel = pk_add(y, E2);
		}
		if (kLast < 16) pwa[NW - 1] >>= 2 * (16 - kLast), pwb[NW - 1] >>= 2 * (16 - kLast);
		pa.put_row(I, pwa);
		pb.put_row(I, pwb);
		wa.slide(QA.next()); // I - b + W - 1 = I + BMAX >= 0
		wb.slide(QB.next());
	};
	int i0 = 1;
	if (mode_or == 0 && wave_all(A.b == BMAX && B.b == BMAX && !(A.left_al && A.lenQ <= 2 * BMAX) &&
	                             !(B.left_al && B.lenQ <= 2 * BMAX))) {
		for_each_index(head, std::make_integer_sequence<int, BMAX>());
		i0 = BMAX + 1;
	}
	const int cut = mode_or | (wave_any(A.b < BMAX || B.b < BMAX) ? kCut : 0);
	for (int i = i0; i <= rows; i++) {
		const bool c0 = i <= A.b || i <= B.b;
		const bool cq = (A.left_al && A.lenQ - i <= A.b) || (B.left_al && B.lenQ - i <= B.b);
		const int mode = cut | (wave_any(c0) ? kC0 : 0) | (wave_any(cq) ? kCQ : 0);
		switch (mode) {
		case 0: row(i, std::integral_constant<int, 0>()); break;
		case kCut: row(i, std::integral_constant<int, kCut>()); break;
		case kCQ: row(i, std::integral_constant<int, kCQ>()); break;
		default: row(i, std::integral_constant<int, kC0 | kCQ | kCut>()); break;
		}
	}
}

} // namespace hero
