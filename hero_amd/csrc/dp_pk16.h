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
// last row computes values nobody reads). SA/QA/SB/QB stream the aligner subject / query codes;
// pa / pb receive each task's pointer rows exactly as dp_forward writes them.
template <int BMAX, bool AFFINE, class SS, class QS, class PA, class PB>
HD void dp_forward_pk16(const Opts &o, const Pk16Task &A, const Pk16Task &B, SS &SA, QS &QA, SS &SB, QS &QB, PA &pa,
		PB &pb) {
	constexpr int W = 2 * BMAX + 1;
	constexpr int NW = (W + 15) / 16;
	static_assert(W <= 64, "row masks hold one bit per band cell");
	using M = typename std::conditional<(W <= 32), uint32_t, uint64_t>::type;
	const I16x2 K2 = pk2(kSent16, kSent16);
	const I16x2 hit2 = pk2(o.hit, o.hit), miss2 = pk2(o.miss, o.miss);
	const I16x2 dx2 = pk2(o.dx, o.dx), dy2 = pk2(o.dy, o.dy), ex2 = pk2(o.ex, o.ex), sx2 = pk2(o.sx, o.sx);
	// per-task gap constants and flag masks
	const I16x2 gcol0 = pk2((A.right_al || (A.lenQ == 0 && A.left_al)) ? o.sx : o.dx,
			(B.right_al || (B.lenQ == 0 && B.left_al)) ? o.sx : o.dx);
	const I16x2 gcolQ = pk2(A.left_al ? o.sx : o.dx, B.left_al ? o.sx : o.dx);
	const uint32_t right_m = (A.right_al ? 0xFFFFu : 0u) | (B.right_al ? 0xFFFF0000u : 0u);
	const uint32_t left_m = (A.left_al ? 0xFFFFu : 0u) | (B.left_al ? 0xFFFF0000u : 0u);
	const M band_a = bits_range<M>(0, 2 * A.b), band_b = bits_range<M>(0, 2 * B.b);
	I16x2 H[W + 1];
	I16x2 X[AFFINE ? W + 1 : 1], Y[AFFINE ? W + 1 : 1];
	uint32_t qw[W]; // packed query codes of the current row: low A, high B
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
		const int qa = d < A.b ? kCodeNone : QA.next();
		const int qb = d < B.b ? kCodeNone : QB.next();
		qw[d] = (uint32_t) qa | ((uint32_t) qb << 16);
	}
	const int rows = imax_(A.lenS, B.lenS);
	// rows without a column-0 or column-lenQ cell in either task need no edge handling
	const int in_lo = imax_(A.b, B.b) + 1;
	const int in_hi = (A.lenQ - A.b < B.lenQ - B.b ? A.lenQ - A.b : B.lenQ - B.b) - 1;
	auto row = [&](int i, auto edge_tag) {
		constexpr bool EDGE = decltype(edge_tag)::value;
		const int sa = SA.next(), sb = SB.next();
		const uint32_t sc2 = (uint32_t) sa | ((uint32_t) sb << 16);
		M va = band_a, vb = band_b, c0a = 0, c0b = 0, cqa = 0, cqb = 0;
		if (EDGE) {
			const int d0a = A.b - i, dQa = A.lenQ - i + A.b, d0b = B.b - i, dQb = B.lenQ - i + B.b;
			va = bits_range<M>(d0a > 0 ? d0a : 0, dQa < 2 * A.b ? dQa : 2 * A.b);
			vb = bits_range<M>(d0b > 0 ? d0b : 0, dQb < 2 * B.b ? dQb : 2 * B.b);
			c0a = bit_at<M>(d0a, W), c0b = bit_at<M>(d0b, W), cqa = bit_at<M>(dQa, W), cqb = bit_at<M>(dQb, W);
		}
		uint32_t pwa[NW], pwb[NW];
#pragma unroll
		for (int w = 0; w < NW; w++) pwa[w] = 0, pwb[w] = 0;
		I16x2 hl = K2, yl = K2;
#pragma unroll
		for (int d = 0; d < W; d++) {
			const uint32_t vm = pk_bitmask(va, vb, d);
			const uint32_t mism = pk_neg(pk_sub(I16x2{0u}, I16x2{sc2 ^ qw[d]})); // codes differ
			const I16x2 s = pk_sel(mism, miss2, hit2);
			const uint32_t pdiag = 0x00020002u | (mism & 0x00010001u);
			I16x2 h, x = K2, y = K2;
			uint32_t p;
			if (!AFFINE) { // src/BandAligner.cpp:63-111
				I16x2 g = dx2;
				if (EDGE) g = pk_sel(pk_bitmask(c0a, c0b, d), gcol0, pk_sel(pk_bitmask(cqa, cqb, d), gcolQ, dx2));
				const I16x2 up = pk_add(H[d + 1], g), lf = pk_add(hl, dy2), dg = pk_add(H[d], s);
				h = pk_max(pk_max(dg, lf), up);
				const uint32_t ndiag = pk_lt(dg, lf) | pk_lt(dg, up);
				const uint32_t pleft = ~pk_lt(lf, up) & 0x00010001u;
				p = (ndiag & pleft) | (~ndiag & pdiag);
			} else { // src/BandAligner.cpp:150-222
				y = pk_max(pk_add(hl, dy2), pk_add(yl, ex2));
				x = pk_max(pk_add(H[d + 1], dx2), pk_add(X[d + 1], ex2));
				if (EDGE) {
					const uint32_t cq = pk_bitmask(cqa, cqb, d) & left_m;
					x = pk_sel(cq, pk_max(pk_add(H[d + 1], sx2), pk_add(X[d + 1], sx2)), x);
				}
				h = pk_add(pk_max(pk_max(H[d], X[d]), Y[d]), s);
				const uint32_t diag = pk_lt(x, h) & ~pk_lt(h, y); // h > x && h >= y
				p = (diag & pdiag) | (~diag & pk_lt(x, y) & 0x00010001u);
				if (EDGE) { // column 0
					const uint32_t c0 = pk_bitmask(c0a, c0b, d);
					const I16x2 x0 = pk_sel(right_m, pk_add(H[d + 1], sx2), pk_add(X[d + 1], ex2));
					const I16x2 y0 = pk_sel(right_m, dy2, x0);
					x = pk_sel(c0, x0, x);
					y = pk_sel(c0, y0, y);
					h = pk_sel(c0, x0, h);
					p &= ~c0;
				}
			}
			h = pk_sel(vm, h, K2);
			H[d] = h;
			hl = h;
			if (AFFINE) {
				X[d] = pk_sel(vm, x, K2);
				Y[d] = pk_sel(vm, y, K2);
				yl = Y[d];
			}
			p &= vm;
			pwa[d >> 4] |= (p & 3u) << (2 * (d & 15));
			pwb[d >> 4] |= ((p >> 16) & 3u) << (2 * (d & 15));
		}
		pa.put_row(i, pwa);
		pb.put_row(i, pwb);
		// slide the query windows to row i+1
#pragma unroll
		for (int d = 0; d + 1 < W; d++) qw[d] = qw[d + 1];
		const int na = (i - A.b + W - 1 >= 0) ? QA.next() : kCodeNone;
		const int nb = (i - B.b + W - 1 >= 0) ? QB.next() : kCodeNone;
		qw[W - 1] = (uint32_t) na | ((uint32_t) nb << 16);
	};
	for (int i = 1; i <= rows; i++) {
		if (i >= in_lo && i <= in_hi) row(i, std::false_type());
		else row(i, std::true_type());
	}
}

} // namespace hero
