// Per-lane banded DP of BandAligner (src/BandAligner.cpp:56-239): one task per lane, band row in
// registers, 2-bit traceback pointers stored row by row through a PtrStore.
//
// Band-diagonal coordinates: cell (i, j) of the (lenS+1) x (lenQ+1) matrix sits at row i,
// column d = j - i + b, d in [0, 2b]. Neighbours: diagonal (i-1, j-1) -> previous row, d;
// up (i-1, j) -> previous row, d+1; left (i, j-1) -> this row, d-1. Cells outside the matrix
// (j < 0 or j > lenQ) and outside the band (d > 2b) hold kSent. A neighbour the reference treats
// as absent (its vals[] = -MAXDUL, or IX/IY := -MAXDUL at the band edges) is exactly a kSent
// cell here, and sentinel-derived values only ever compete against real values (which stay far
// above -MAXDUL), so every comparison the reference makes has the same outcome.
#pragma once
#include <type_traits>
#include <utility>

#include "hd.h"

namespace hero {

constexpr int kSent = -(1 << 28);
constexpr int kCodeN = 4;    // 'N' (BandAligner::CWILD) in the standalone aligner
constexpr int kCodeNone = 7; // beyond the end of a sequence (never equal to a base)

HD int imax_(int a, int b) { return a > b ? a : b; }

// f(integral_constant<int, K>) for each K of the sequence, in order
template <class F, int... K>
HD void for_each_index(F &f, std::integer_sequence<int, K...>) {
	(f(std::integral_constant<int, K>()), ...);
}

// Wave-uniform row decisions of the per-lane DP (one task per lane): true when the predicate
// holds in any / every active lane. A single task (host code) is its own wave.
HD bool wave_any(bool v) {
#if defined(__HIP_DEVICE_COMPILE__)
	return __ballot(v) != 0;
#else
	return v;
#endif
}
HD bool wave_all(bool v) {
#if defined(__HIP_DEVICE_COMPILE__)
	return __ballot(!v) == 0;
#else
	return v;
#endif
}

// Diagonal score of the global aligner (src/BandAligner.cpp:89-97): a mismatch involving N
// scores unobserved_match; equal characters (N==N included) score hit.
HD int diag_score_global(const Opts &o, int a, int q) {
	return a != q ? ((a == kCodeN || q == kCodeN) ? o.unobserved : o.miss) : o.hit;
}

// One cell of the recurrence. d0 / dQ flag the cell as column 0 / column lenQ; *_diag, *_up,
// *_left are the neighbours (kSent when absent). Writes the cell's values and pointer code:
// the reference's BM1 (0 up, 1 left, 2 diagonal) with the diagonal split into 2 (the two
// characters are equal, a CIGAR 'M') and 3 (they differ, a mismatch op), so the traceback needs
// no sequence access.
template <bool AFFINE>
HD void cell_update(const Opts &o, int sc, int qc, bool col0, bool colQ, bool left_al, bool right_al, int lenQ,
		int h_diag, int x_diag, int y_diag, int h_up, int x_up, int h_left, int y_left, int &h, int &x, int &y, int &p) {
	if (!AFFINE) { // src/BandAligner.cpp:63-111
		int g = o.dx;
		if (col0) g = right_al ? o.sx : ((lenQ == 0 && left_al) ? o.sx : o.dx);
		else if (colQ) g = left_al ? o.sx : o.dx;
		const int up = h_up + g;
		const int lf = h_left + o.dy;
		const int dg = h_diag + diag_score_global(o, sc, qc);
		if (dg >= lf && dg >= up) h = dg, p = 2 + (sc != qc);
		else if (lf >= up) h = lf, p = 1;
		else h = up, p = 0;
		x = kSent, y = kSent;
	} else if (col0) { // column 0 (src/BandAligner.cpp:150-163)
		if (right_al) x = h_up + o.sx, y = o.dy;
		else x = x_up + o.ex, y = x;
		h = x;
		p = 0;
	} else { // src/BandAligner.cpp:176-222
		y = imax_(h_left + o.dy, y_left + o.ex);
		if (left_al && colQ) x = imax_(h_up + o.sx, x_up + o.sx);
		else x = imax_(h_up + o.dx, x_up + o.ex);
		const int s = sc != qc ? o.miss : o.hit;
		h = imax_(imax_(h_diag, x_diag), y_diag) + s; // = max(h_diag + s, x_diag + s, y_diag + s)
		p = (h > x && h >= y) ? 2 + (sc != qc) : (y > x ? 1 : 0);
	}
}

// Row 0 of the matrix (global: src/BandAligner.cpp:63-111 at i = 0; affine: :140-171).
// Returns whether cell j of row 0 exists and its values; its pointer is 1 (left) for j >= 1.
template <bool AFFINE>
HD bool row0_cell(const Opts &o, int j, int b, int lenQ, int &h, int &x, int &y) {
	const bool in = j >= 0 && j <= b && j <= lenQ;
	if (AFFINE) {
		const int v = j == 0 ? 0 : (o.dx - o.ex) + j * o.ex;
		h = in ? v : kSent;
		x = in ? (j == 0 ? o.dx - o.ex : v) : kSent;
		y = x;
	} else {
		h = in ? o.dy * j : kSent;
		x = kSent, y = kSent;
	}
	return in;
}

// Query window of a row: the code of cell d is Q[i - b - 1 + d]. Streams whose codes are 2-bit
// bases (Stream::kTwoBit) keep the window packed, 2 bits per cell, in one 64-bit value (band
// width up to 32 cells); others keep one register per cell. A code past either end of the query
// is read only by cells outside the matrix, so the packed form may keep it as any base.
template <int W, bool TWO_BIT>
struct QWin {
	int q[W];
	HD int at(int d) const { return q[d]; }
	HD void set(int d, int c) { q[d] = c; }
	HD void slide(int next) {
#pragma unroll
		for (int d = 0; d + 1 < W; d++) q[d] = q[d + 1];
		q[W - 1] = next;
	}
	HD uint32_t differs(int d, int sc) const { return q[d] != sc ? 1u : 0u; }
	HD uint64_t differ_bits(int) const { return 0; }
};
template <int W>
struct QWin<W, true> {
	static_assert(W <= 32, "packed window holds 32 cells");
	uint64_t v = 0; // cell d at bits 2d, 2d+1
	HD int at(int d) const { return (int) ((v >> (2 * d)) & 3u); }
	HD void set(int d, int c) { v |= (uint64_t) (c & 3) << (2 * d); }
	HD void slide(int next) { v = (v >> 2) | ((uint64_t) (next & 3) << (2 * (W - 1))); }
	// bit 2d set where cell d's code differs from sc
	HD uint64_t differ_bits(int sc) const {
		const uint64_t x = v ^ (0x5555555555555555ull * (uint64_t) sc);
		return (x | (x >> 1)) & 0x5555555555555555ull;
	}
};

// Forward pass. S.next() / Q.next() yield the aligner subject / query codes in order and
// kCodeNone past the end. ptrs.put_row(i, words) receives row i's pointers, 16 cells per word,
// cell d at bits 2*(d%16); codes are the reference's BM1 values: 2 diagonal, 1 left, 0 up.
template <int BMAX, bool AFFINE, class SS, class QS, class PS>
HD void dp_forward(const Opts &o, int lenS, int lenQ, int b, bool left_al, bool right_al, SS &S, QS &Q, PS &ptrs) {
	constexpr int W = 2 * BMAX + 1;
	constexpr int NW = (W + 15) / 16;
	int H[W + 1];
	int X[AFFINE ? W + 1 : 1], Y[AFFINE ? W + 1 : 1];
	constexpr bool TWO_BIT = QS::kTwoBit && W <= 32;
	QWin<W, TWO_BIT> qw; // qw.at(d) = Q[i - b - 1 + d] for the current row i
	// Row 0 (global: src/BandAligner.cpp:63-111 at i = 0; affine: :140-171).
	{
		uint32_t pw[NW];
#pragma unroll
		for (int w = 0; w < NW; w++) pw[w] = 0;
#pragma unroll
		for (int d = 0; d <= W; d++) {
			int h, x, y;
			const bool in = row0_cell<AFFINE>(o, d - b, b, lenQ, h, x, y) && d < W;
			H[d] = in ? h : kSent;
			if (AFFINE) X[d] = in ? x : kSent, Y[d] = in ? y : kSent;
			if (d < W && in && d - b >= 1) pw[d >> 4] |= 1u << (2 * (d & 15));
		}
		ptrs.put_row(0, pw);
	}
#pragma unroll
	for (int d = 0; d < W; d++) qw.set(d, d < b ? kCodeNone : Q.next()); // row 1: Q[d - b]
	// Every row computes all band cells by the interior recurrence and clears nothing outside the
	// matrix: a cell of the matrix reads only cells of the matrix (columns j - 1 and j of rows
	// i - 1 and i), except the column-0 cell, whose own rule reads only the cell above it. So
	// values computed left of column 0 or right of column lenQ are never read. The edge rules are
	// applied where they can occur, by per-cell selects:
	//   kC0 - column 0 is in the band (i <= b): that cell takes the column-0 rule, whose values
	//         have a closed form in i (a vertical gap from the origin: the affine x is i * sx
	//         after setAlignRight, else dx + (i - 1) * ex, with y = dy or x; the global h is i * g0);
	//   kCQ - an aligned-left call has column lenQ in the band (i >= lenQ - b): that cell's
	//         vertical move costs sx (src/BandAligner.cpp:89-97,196-200);
	//   kCut - some band is narrower than BMAX: cells past the band edge 2b are cut from the left
	//         chain, so they stay sentinel-derived (with b = BMAX the edge is the register row's end).
	// Row variants are chosen per wavefront (any lane that needs a rule), so lanes stay together;
	// rows with an edge rule take all three (fewer variants keep the register allocation lean).
	constexpr int kC0 = 1, kCQ = 2, kCut = 4;
	const int d0base = b, dQbase = lenQ + b; // d0 = b - i, dQ = lenQ - i + b
	const int edge_d = 2 * b;                // last band cell
	const int g0 = (right_al || (lenQ == 0 && left_al)) ? o.sx : o.dx; // global: column-0 vertical cost
	auto row = [&](int i, int sc, uint32_t *pw, auto mode_tag) {
		constexpr int MODE = decltype(mode_tag)::value;
#pragma unroll
		for (int w = 0; w < NW; w++) pw[w] = 0;
		int hl = kSent, yl = kSent; // this row's cell at d-1
		const int d0 = d0base - i, dQ = left_al ? dQbase - i : -1;
		const int x0 = AFFINE ? (right_al ? i * o.sx : o.dx + (i - 1) * o.ex) : i * g0; // column-0 values
		const int y0 = right_al ? o.dy : x0;
		const uint64_t mbits = qw.differ_bits(sc);
#pragma unroll
		for (int d = 0; d < W; d++) {
			// mismatch flag of the cell: from the packed window's bits, or by comparing codes
			const uint32_t m1 = TWO_BIT ? (uint32_t) (mbits >> (2 * d)) & 1u : (sc != qw.at(d) ? 1u : 0u);
			const bool mm = m1 != 0;
			const bool c0 = (MODE & kC0) && d == d0;
			const bool cq = (MODE & kCQ) && d == dQ;
			uint32_t p;
			if (!AFFINE) { // src/BandAligner.cpp:63-111
				const int g = (MODE & kCQ) ? (cq ? o.sx : o.dx) : o.dx;
				const int up = H[d + 1] + g;
				const int lf = hl + o.dy;
				const int dg = H[d] + (TWO_BIT ? (mm ? o.miss : o.hit) : diag_score_global(o, sc, qw.at(d)));
				int h;
				if (dg >= lf && dg >= up) h = dg, p = 2u + m1;
				else if (lf >= up) h = lf, p = 1u;
				else h = up, p = 0u;
				if (MODE & kC0) { // column 0: the vertical move only
					h = c0 ? x0 : h;
					p = c0 ? 0u : p;
				}
				if (MODE & kCut) h = d <= edge_d ? h : kSent;
				H[d] = h;
				hl = h;
			} else { // src/BandAligner.cpp:150-222
				int y = imax_(hl + o.dy, yl + o.ex);
				if (MODE & kCut) y = d <= edge_d ? y : kSent;
				int x = imax_(H[d + 1] + o.dx, X[d + 1] + o.ex);
				if (MODE & kCQ) x = cq ? imax_(H[d + 1], X[d + 1]) + o.sx : x;
				int h = imax_(imax_(H[d], X[d]), Y[d]) + (mm ? o.miss : o.hit);
				p = (h > x && h >= y) ? 2u + m1 : (y > x ? 1u : 0u);
				if (MODE & kC0) { // column 0 (src/BandAligner.cpp:150-163)
					x = c0 ? x0 : x;
					y = c0 ? y0 : y;
					h = c0 ? x0 : h;
					p = c0 ? 0u : p;
				}
				H[d] = h, X[d] = x, Y[d] = y;
				hl = h, yl = y;
			}
			pw[d >> 4] |= p << (2 * (d & 15));
		}
	};
	const int cut = wave_all(b == BMAX) ? 0 : kCut;
	for (int i = 1; i <= lenS; i++) {
		const int sc = S.next();
		uint32_t pw[NW];
		const int mode = cut | (wave_any(i <= b) ? kC0 : 0) | (wave_any(left_al && lenQ - i <= b) ? kCQ : 0);
		switch (mode) {
		case 0: row(i, sc, pw, std::integral_constant<int, 0>()); break;
		case kCut: row(i, sc, pw, std::integral_constant<int, kCut>()); break;
		default: row(i, sc, pw, std::integral_constant<int, kC0 | kCQ | kCut>()); break;
		}
		ptrs.put_row(i, pw);
		// slide the query window to row i+1: cell d reads Q[i - b + d]
		qw.slide((i - b + W - 1 >= 0) ? Q.next() : kCodeNone);
	}
}

// Tagged affine forward pass: the same recurrence, results and pointers as dp_forward<BMAX, true>
// for gap opens of equal cost (o.dx == o.dy) and 2-bit codes, in fewer operations per cell.
//
// A value v of cell (i, j) is kept as 8 v + t - 8 ex (i + j): scaled by 8, a tag t in the two low
// bits, and an offset that moves by 8 ex per up or left step and by 16 ex per diagonal step, so a
// gap extension adds nothing (the two gap chains are plain maxima along the row and down the band)
// and a diagonal move adds K + 8 (s - hit), K = 8 hit - 16 ex. Values of one cell share the offset,
// so comparing them compares the values, and equal values compare by tag. The tags rank the three
// states as the reference breaks ties (src/BandAligner.cpp:176-222: the diagonal needs h > x and
// h >= y, then the left gap needs y > x): 3 the vertical gap (up), 2 / 1 the diagonal (match /
// mismatch), 0 the horizontal gap (left). The max3 of a cell's three states is therefore at once
// the next row's diagonal input and, in its low bits, the cell's pointer, stored as it is.
//
// A cell's h plus a gap open, G = 8 h - off + 8 (dx - ex) (untagged), is its up neighbour's and
// its left neighbour's gap-open candidate at once (dx == dy; both moves add 8 ex to the offset);
// the extended vertical gap is XE = x (tag 3 kept), the horizontal one y (tag 0). Per band cell
// the rows carry two values to the next row:
//   MK - the diagonal base: 8 max(h, x, y) - off + 2 + K (the match tag preset), so the next row's
//        h is MK on a match and MK + 8 (miss - hit) - 1 on a mismatch;
//   U  - the vertical gap of the cell below: max(G, XE) | 3.
// G and y + E of the cell to the left run along the row. The aligned column lenQ's rule needs G
// and XE of the cell above separately: each lane keeps them for its next row's column lenQ.
// Sentinel cells hold kSentT; the caller keeps every real value within the scaled range
// (dp_tagged_fits). Pointer codes: 3 up, 2 match, 1 mismatch, 0 left (kPtrTaggedAffine).
constexpr int kSentT = -(1 << 29);
constexpr int kTagRange = 1 << 27;
HD int max3_(int a, int b, int c) { return imax_(imax_(a, b), c); }
// bit p of w (0 or 1)
HD uint32_t bit_u32(uint32_t w, int p) {
#if defined(__HIP_DEVICE_COMPILE__)
	return __builtin_amdgcn_ubfe(w, (uint32_t) p, 1u);
#else
	return (w >> p) & 1u;
#endif
}
// The mismatch bit of cell d as one extracted value that the cell's two uses share (left to itself,
// the compiler extracts it once for the product and again, by a shift and an or, for the sum).
HD int cell_bit(uint32_t lo, uint32_t hi, int d) {
	int m = (int) bit_u32(d < 16 ? lo : hi, 2 * (d & 15));
#if defined(__HIP_DEVICE_COMPILE__)
	asm("" : "+v"(m));
#endif
	return m;
}
// c + m * a for m in {0, 1} and |a| < 2^23: one v_mad_i32_i24 (written out, because the compiler
// turns a product with a one-bit value into a bit test and a select: three instructions)
HD int mad01(int m, int a, int c) {
#if defined(__HIP_DEVICE_COMPILE__)
	int r;
	asm("v_mad_i32_i24 %0, %1, %2, %3" : "=v"(r) : "v"(m), "v"(a), "v"(c));
	return r;
#else
	return c + m * a;
#endif
}
// Every scaled value of a task with rows_max rows stays within +-kTagRange: |v| <= s (lenS + lenQ)
// for the largest step cost s, times 8, plus the offset 8 ex (lenS + lenQ) and the tag.
HD bool dp_tagged_fits(const Opts &o, int rows_max) {
	int s = o.hit < 0 ? -o.hit : o.hit;
	const int c[5] = {o.dx, o.ex, o.sx, o.miss, o.dy};
	for (int k = 0; k < 5; k++) s = imax_(s, c[k] < 0 ? -c[k] : c[k]);
	return (int64_t) (16 * (int64_t) s + 4) * (int64_t) (2 * rows_max + 2) < (int64_t) kTagRange;
}

// mode_or: row variants taken in addition to those the task needs (a lane in a wavefront with
// other tasks takes theirs too); the kernels pass 0.
template <int BMAX, class SS, class QS, class PS>
HD void dp_forward_tagged(const Opts &o, int lenS, int lenQ, int b, bool left_al, bool right_al, SS &S, QS &Q,
		PS &ptrs, int mode_or = 0) {
	static_assert(SS::kTwoBit && QS::kTwoBit, "the tagged form reads 2-bit codes");
	constexpr int W = 2 * BMAX + 1;
	static_assert(W <= 32, "the packed query window holds 32 cells");
	constexpr int NW = (W + 15) / 16;
	constexpr int kLast = W - 16 * (NW - 1); // cells in the last pointer word
	const int step = 8 * o.ex;                  // offset of one up or left move
	const int C = 8 * o.dx - step;              // G = 8 h - off + C
	const int E = 8 * o.ex - step;              // XE = x + E, and y + E along the row: 0
	const int K = 8 * o.hit - 2 * step;         // a diagonal match
	const int DM = 8 * (o.miss - o.hit) - 1;    // h on a mismatch: MK + DM (tag 2 -> 1)
	const int CQx = 8 * (o.sx - o.dx), CQe = 8 * (o.sx - o.ex); // aligned column lenQ: vertical cost sx
	int MK[W + 1], U[W + 1];
	int Gq = kSentT, XEq = kSentT; // G and XE of the cell above the next row's column lenQ
	QWin<W, true> qw;              // qw.at(d) = Q[i - b - 1 + d] for the current row i
	{ // row 0 (src/BandAligner.cpp:140-171): pointers 0 (left) everywhere
		uint32_t pw[NW];
#pragma unroll
		for (int w = 0; w < NW; w++) pw[w] = 0;
#pragma unroll
		for (int d = 0; d <= W; d++) {
			int h, x, y;
			const int j = d - b;
			const bool in = row0_cell<true>(o, j, b, lenQ, h, x, y) && d < W;
			const int off = step * j;
			const int g = in ? 8 * h - off + C : kSentT, xe = in ? 8 * x + 3 - off + E : kSentT;
			MK[d] = in ? 8 * max3_(h, x, y) - off + 2 + K : kSentT;
			U[d] = in ? imax_(g, xe) | 3 : kSentT;
			if (left_al && d == lenQ + b) Gq = g, XEq = xe; // above row 1's column lenQ
		}
		ptrs.put_row(0, pw);
	}
#pragma unroll
	for (int d = 0; d < W; d++) qw.set(d, d < b ? kCodeNone : Q.next()); // row 1: Q[d - b]
	// Rows as in dp_forward: every band cell takes the interior recurrence, and the edge rules
	// (column 0, the aligned column lenQ, the cut past a narrower band's edge) are selects in
	// the rows where some lane needs them.
	constexpr int kC0 = 1, kCQ = 2, kCut = 4;
	const int d0base = b, dQbase = lenQ + b;
	const int edge_d = 2 * b;
	auto row = [&](int i, int sc, uint32_t *pw, auto mode_tag) {
		constexpr int MODE = decltype(mode_tag)::value;
		const int d0 = d0base - i, dQ = left_al ? dQbase - i : -1;
		// column 0 (src/BandAligner.cpp:150-163): a vertical gap from the origin; off = step * i
		int h0 = 0, x0 = 0, y0 = 0, g0 = 0;
		if (MODE & kC0) {
			const int xv = right_al ? i * o.sx : o.dx + (i - 1) * o.ex;
			const int yv = right_al ? o.dy : xv;
			const int off = step * i;
			h0 = 8 * xv + 2 - off, x0 = 8 * xv + 3 - off, y0 = 8 * yv - off, g0 = 8 * xv - off + C;
		}
		const uint64_t mbits = qw.differ_bits(sc);
		const uint32_t mlo = (uint32_t) mbits, mhi = (uint32_t) (mbits >> 32);
		int gl = kSentT, yel = kSentT; // G and y + E of this row's cell d - 1
		int gq = Gq, xeq = XEq;         // of this row's column lenQ, for the next row
#pragma unroll
		for (int d = 0; d < W; d++) {
			const int m = cell_bit(mlo, mhi, d);
			int h = mad01(m, DM, MK[d]);
			int x = U[d + 1];
			int y = imax_(gl, yel);
			if (MODE & kCQ) x = d == dQ ? imax_(Gq + CQx, XEq + CQe) | 3 : x;
			if (MODE & kCut) y = d <= edge_d ? y : kSentT;
			int g = h + m + (C - 2); // 8 h - off + C: the tag 2 - m taken off
			int mx;
			if (MODE & kC0) {
				const bool c0 = d == d0;
				h = c0 ? h0 : h, x = c0 ? x0 : x, y = c0 ? y0 : y, g = c0 ? g0 : g;
				mx = max3_(h, x, y);
				const int pv = c0 ? 3 : mx; // column 0 points up
				pw[d >> 4] = (pw[d >> 4] >> 2) | ((uint32_t) pv << 30);
			} else {
				mx = max3_(h, x, y);
				pw[d >> 4] = (pw[d >> 4] >> 2) | ((uint32_t) mx << 30);
			}
			MK[d] = ((mx & ~3) | 2) + K;
			const int xe = x + E;
			U[d] = imax_(g, xe) | 3;
			if (MODE & kCQ) gq = d == dQ ? g : gq, xeq = d == dQ ? xe : xeq;
			gl = g, yel = y + E;
		}
		if (MODE & kCQ) Gq = gq, XEq = xeq;
		if (kLast < 16) pw[NW - 1] >>= 2 * (16 - kLast);
	};
	// Head rows 1..BMAX when every call of the wavefront has band BMAX and no aligned column lenQ
	// among them (lenQ > 2 BMAX where aligned left): row I's column 0 is band cell BMAX - I for all
	// lanes, a compile-time index, so the row starts there with the column-0 values and takes the
	// interior rule for the cells after it. The cells before it lie left of column 0: no cell of
	// the matrix reads them and the traceback never visits them, so they are not computed (their
	// pointer bits stay 0; the shift-in below still puts every computed cell at bit 2 (d % 16)).
	auto row_head = [&](int sc, uint32_t *pw, auto itag) {
		constexpr int I = decltype(itag)::value;
		constexpr int D0 = BMAX - I;
		const int xv = right_al ? I * o.sx : o.dx + (I - 1) * o.ex;
		const int yv = right_al ? o.dy : xv;
		const int off = step * I;
		const int h0 = 8 * xv + 2 - off, x0 = 8 * xv + 3 - off, y0 = 8 * yv - off, g0 = 8 * xv - off + C;
		const uint64_t mbits = qw.differ_bits(sc);
		const uint32_t mlo = (uint32_t) mbits, mhi = (uint32_t) (mbits >> 32);
		int gl, yel;
		{ // column 0 (src/BandAligner.cpp:150-163): a vertical gap from the origin, pointing up
			const int mx = max3_(h0, x0, y0);
			pw[D0 >> 4] = (pw[D0 >> 4] >> 2) | (3u << 30);
			MK[D0] = ((mx & ~3) | 2) + K;
			U[D0] = imax_(g0, x0 + E) | 3;
			gl = g0, yel = y0 + E;
		}
#pragma unroll
		for (int d = D0 + 1; d < W; d++) {
			const int m = cell_bit(mlo, mhi, d);
			const int h = mad01(m, DM, MK[d]);
			const int x = U[d + 1];
			const int y = imax_(gl, yel);
			const int g = h + m + (C - 2);
			const int mx = max3_(h, x, y);
			pw[d >> 4] = (pw[d >> 4] >> 2) | ((uint32_t) mx << 30);
			MK[d] = ((mx & ~3) | 2) + K;
			U[d] = imax_(g, x + E) | 3;
			gl = g, yel = y + E;
		}
		if (kLast < 16) pw[NW - 1] >>= 2 * (16 - kLast);
	};
	auto head = [&](auto itag) {
		constexpr int I = decltype(itag)::value + 1;
		if (I <= lenS) {
			const int sc = S.next();
			uint32_t pw[NW];
#pragma unroll
			for (int w = 0; w < NW; w++) pw[w] = 0;
			row_head(sc, pw, std::integral_constant<int, I>());
			ptrs.put_row(I, pw);
			qw.slide(Q.next()); // I - b + W - 1 = I + BMAX >= 0
		}
	};
	int i0 = 1;
	if (mode_or == 0 && wave_all(b == BMAX && !(left_al && lenQ <= 2 * BMAX))) {
		for_each_index(head, std::make_integer_sequence<int, BMAX>());
		i0 = BMAX + 1;
	}
	const int cut = mode_or | (wave_all(b == BMAX) ? 0 : kCut);
	for (int i = i0; i <= lenS; i++) {
		const int sc = S.next();
		uint32_t pw[NW];
#pragma unroll
		for (int w = 0; w < NW; w++) pw[w] = 0;
		const int mode = cut | (wave_any(i <= b) ? kC0 : 0) | (wave_any(left_al && lenQ - i <= b) ? kCQ : 0);
		switch (mode) {
		case 0: row(i, sc, pw, std::integral_constant<int, 0>()); break;
		case kCut: row(i, sc, pw, std::integral_constant<int, kCut>()); break;
		case kCQ: row(i, sc, pw, std::integral_constant<int, kCQ>()); break;
		default: row(i, sc, pw, std::integral_constant<int, kC0 | kCQ | kCut>()); break;
		}
		ptrs.put_row(i, pw);
		qw.slide((i - b + W - 1 >= 0) ? Q.next() : kCodeNone);
	}
}

// Tagged global forward pass: the results and pointers of dp_forward<BMAX, false> (the global
// aligner, src/BandAligner.cpp:63-111) for 2-bit codes, in fewer operations per cell.
//
// A cell's value h is kept as V = 8 h - 4 hit (i + j) + t, the tag t in the two low bits ranking the
// moves as the reference breaks ties (the diagonal when dg >= lf and dg >= up, then the left move
// when lf >= up): 3 diagonal match, 2 diagonal mismatch, 1 left, 0 up. A diagonal move adds
// 8 (s - hit), nothing on a match; an up move adds 8 g - 4 hit, a left move 8 dy - 4 hit, both
// multiples of 4, so the tag of a candidate is set by one or. The max3 of the three candidates is
// the cell's value and, in its low bits, its pointer. Each row carries two values per band cell to
// the next: MK, the cell's value with the match tag (the next row's diagonal candidate on a match;
// a mismatch adds 8 (miss - hit) - 1), and UP, its value plus the vertical move (the up candidate
// of the cell below). The aligned column lenQ's vertical move costs sx instead of dx: its up
// candidate adds 8 (sx - dx). Values outside the matrix are never read (as in dp_forward).
// Pointer codes: 0 up, 1 left, 2 mismatch, 3 match (kPtrTaggedGlobal).
template <int BMAX, class SS, class QS, class PS>
HD void dp_forward_tagged_global(const Opts &o, int lenS, int lenQ, int b, bool left_al, bool right_al, SS &S, QS &Q,
		PS &ptrs, int mode_or = 0) {
	static_assert(SS::kTwoBit && QS::kTwoBit, "the tagged form reads 2-bit codes");
	constexpr int W = 2 * BMAX + 1;
	static_assert(W <= 32, "the packed query window holds 32 cells");
	constexpr int NW = (W + 15) / 16;
	constexpr int kLast = W - 16 * (NW - 1);
	const int step = 4 * o.hit;
	const int CU = 8 * o.dx - step;          // up move (dx)
	const int CL1 = 8 * o.dy - step + 1;     // left move (dy), tag 1
	const int DM = 8 * (o.miss - o.hit) - 1; // diagonal on a mismatch: match tag 3 -> 2
	const int CQ = 8 * (o.sx - o.dx);        // the aligned column lenQ's vertical move costs sx
	const int g0 = (right_al || (lenQ == 0 && left_al)) ? o.sx : o.dx; // column-0 vertical cost
	int MK[W + 1], UP[W + 1];
	QWin<W, true> qw;
	{ // row 0 (src/BandAligner.cpp:63-111 at i = 0): h = j dy, pointers 1 (left) for j >= 1
		uint32_t pw[NW];
#pragma unroll
		for (int w = 0; w < NW; w++) pw[w] = 0;
#pragma unroll
		for (int d = 0; d <= W; d++) {
			int h, x, y;
			const int j = d - b;
			const bool in = row0_cell<false>(o, j, b, lenQ, h, x, y) && d < W;
			const int v = 8 * h - step * j; // untagged
			MK[d] = in ? v | 3 : kSentT;
			UP[d] = in ? v + CU : kSentT;
			if (d < W && in && j >= 1) pw[d >> 4] |= 1u << (2 * (d & 15));
		}
		ptrs.put_row(0, pw);
	}
#pragma unroll
	for (int d = 0; d < W; d++) qw.set(d, d < b ? kCodeNone : Q.next()); // row 1: Q[d - b]
	constexpr int kC0 = 1, kCQ = 2, kCut = 4;
	const int d0base = b, dQbase = lenQ + b;
	const int edge_d = 2 * b;
	auto row = [&](int i, int sc, uint32_t *pw, auto mode_tag) {
		constexpr int MODE = decltype(mode_tag)::value;
		const int d0 = d0base - i, dQ = left_al ? dQbase - i : -1;
		const int v0 = 8 * i * g0 - step * i; // column 0 (a vertical gap from the origin), pointing up
		const uint64_t mbits = qw.differ_bits(sc);
		const uint32_t mlo = (uint32_t) mbits, mhi = (uint32_t) (mbits >> 32);
		int hl = kSentT; // this row's cell d - 1, untagged
#pragma unroll
		for (int d = 0; d < W; d++) {
			const int m = (int) bit_u32(d < 16 ? mlo : mhi, 2 * (d & 15));
			const int dg = mad01(m, DM, MK[d]);
			int up = UP[d + 1];
			if (MODE & kCQ) up = d == dQ ? up + CQ : up;
			const int lf = hl + CL1;
			int mx = max3_(dg, lf, up);
			if (MODE & kC0) mx = d == d0 ? v0 : mx;
			if (MODE & kCut) mx = d <= edge_d ? mx : kSentT;
			pw[d >> 4] = (pw[d >> 4] >> 2) | ((uint32_t) mx << 30);
			const int v = mx & ~3;
			MK[d] = v | 3;
			UP[d] = v + CU;
			hl = v;
		}
		if (kLast < 16) pw[NW - 1] >>= 2 * (16 - kLast);
	};
	const int cut = mode_or | (wave_all(b == BMAX) ? 0 : kCut);
	for (int i = 1; i <= lenS; i++) {
		const int sc = S.next();
		uint32_t pw[NW];
#pragma unroll
		for (int w = 0; w < NW; w++) pw[w] = 0;
		const int mode = cut | (wave_any(i <= b) ? kC0 : 0) | (wave_any(left_al && lenQ - i <= b) ? kCQ : 0);
		switch (mode) {
		case 0: row(i, sc, pw, std::integral_constant<int, 0>()); break;
		case kCut: row(i, sc, pw, std::integral_constant<int, kCut>()); break;
		case kCQ: row(i, sc, pw, std::integral_constant<int, kCQ>()); break;
		default: row(i, sc, pw, std::integral_constant<int, kC0 | kCQ | kCut>()); break;
		}
		ptrs.put_row(i, pw);
		qw.slide((i - b + W - 1 >= 0) ? Q.next() : kCodeNone);
	}
}

// Pointer rows of one lane among LANES, split between two stores: rows i < near_rows in `near`
// (the wavefront's LDS on the device) and the others in `far` (global scratch: word w of row i at
// ((i * NW) + w) * LANES + lane). Both are lane-interleaved, so a wavefront's row store is one
// coalesced write and lanes reading different rows never share an LDS bank. When a row's last word
// holds a single cell (2 BMAX + 1 = 16 (NW - 1) + 1: band classes 8 and 16) the near store keeps
// only the NW - 1 full words of a row and puts that cell's 2-bit code into a side word shared by
// 16 consecutive rows, so it holds almost twice the rows. Rows are put in order from row 0 (as the
// forward passes do): a side word restarts at every 16th row.
template <int NW, bool SIDE, int LANES>
struct SplitRowPtrs {
	static_assert(!SIDE || NW > 1, "a side word takes the single cell of a row's last word");
	static constexpr int kNear = SIDE ? NW - 1 : NW; // words per row in the near store
	uint32_t *near, *far;
	int lane, near_rows;
	uint32_t side;
	// the rows whose near-store words (and side words) fit in `words` words per lane
	static HD int rows_for(int words) {
		if (words <= 0) return 0;
		int r = SIDE ? (16 * words) / (16 * kNear + 1) : words / kNear;
		while (r > 0 && r * kNear + (SIDE ? (r + 15) / 16 : 0) > words) r--;
		return r;
	}
	HD void put_row(int i, const uint32_t *pw) {
		if (i < near_rows) {
#pragma unroll
			for (int w = 0; w < kNear; w++) near[(i * kNear + w) * LANES + lane] = pw[w];
			if constexpr (SIDE) {
				const uint32_t c = (pw[NW - 1] & 3u) << (2 * (i & 15));
				side = (i & 15) ? (side | c) : c;
				near[(near_rows * kNear + (i >> 4)) * LANES + lane] = side;
			}
		} else {
#pragma unroll
			for (int w = 0; w < NW; w++) far[((size_t) i * NW + w) * LANES + lane] = pw[w];
		}
	}
	// word w of row i; of a side word, the row's code in the low two bits (the traceback reads the
	// last word's single cell at bit 0)
	HD uint32_t get(int i, int w) const {
		if (i < near_rows) {
			if (SIDE && w == NW - 1) return near[(near_rows * kNear + (i >> 4)) * LANES + lane] >> (2 * (i & 15));
			return near[(i * kNear + w) * LANES + lane];
		}
		return far[((size_t) i * NW + w) * LANES + lane];
	}
};

// Pointer codes of the forward passes as maps from a stored 2-bit code t to the plain code
// p = (MAP >> 2 t) & 3 (0 up, 1 left, 2 match, 3 mismatch).
constexpr uint32_t kPtrPlain = 0xE4u;         // dp_forward: stored as is
constexpr uint32_t kPtrTaggedAffine = 0x2Du;  // dp_forward_tagged: 0 left, 1 mismatch, 2 match, 3 up
constexpr uint32_t kPtrTaggedGlobal = 0xB4u;  // dp_forward_tagged_global: 0 up, 1 left, 2 mismatch, 3 match

// Traceback (src/BandAligner.cpp:114-125,227-238): from (lenS, lenQ) follow the pointers to the
// origin. Steps come last-to-first through emit(step): kStepMatch / kStepMismatch for a diagonal,
// kStepLeft (j-1), kStepUp (i-1). Returns the step count, or -1 if the path leaves the band
// (never for pointers written by the forward passes). MAP: the forward pass's pointer codes.
template <uint32_t MAP = kPtrPlain, class PS, class EM>
HD int dp_traceback(int lenS, int lenQ, int b, const PS &ptrs, EM &emit) {
	int i = lenS, j = lenQ, n = 0;
	while (!(i == 0 && j == 0)) {
		const int d = j - i + b;
		if (i < 0 || j < 0 || d < 0 || d > 2 * b) return -1; // left the band: corrupt pointers
		const uint32_t word = ptrs.get(i, d >> 4);
		const int t = (int) ((word >> (2 * (d & 15))) & 3u);
		const int p = (int) ((MAP >> (2 * t)) & 3u);
		if (p >= 2) {
			emit(p == 2 ? kStepMatch : kStepMismatch);
			i--, j--;
		} else if (p == 1) {
			emit(kStepLeft);
			j--;
		} else {
			emit(kStepUp);
			i--;
		}
		n++;
	}
	return n;
}

} // namespace hero
