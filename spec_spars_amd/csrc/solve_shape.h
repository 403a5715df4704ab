// Summation shape of one row of a level-scheduled triangular solve.
//
// A row's entries are sorted by the level of their column.  A forward row
// of a graph's trailing run of small levels has two segments: the leading
// entries whose columns sit in the graph's head levels, then the rest;
// every other row is one segment.  The boundary depends on the graph alone
// (its own small-level run), not on other graphs sharing the launches.
//
// Each segment [a, b) updates the running value acc (the row's start value
// first): with P = seg_parts(b - a) == 1, acc -= val[q] * x[col[q]] for q in
// order; otherwise part s (0 <= s < P) is a running subtraction from zero
// over q = a + s, a + s + P, ..., the P parts are folded as a balanced
// pairwise tree (fold(lo..hi) = fold(left half) + fold(right half)) and
// acc += that fold.
//
// A shape that depends only on the row means any kernel may finish it --
// a wave holding every part in registers, a block of 4 waves, or the waves
// of the small-level kernel sharing a row -- with the same bits.
#pragma once

namespace ss {

constexpr int kLongRow = 32;      // rows longer than this get a whole block per level launch
constexpr int kSmallRows = 16;    // levels with at most this many rows ...
constexpr int kSmallNnz = 2048;   // ... and entries go to the single-workgroup kernel

static_assert(kSmallRows == 16, "the small-level kernel runs 16 waves, at most one row each");

constexpr int seg_parts(int len) { return len <= 8 ? 1 : (len <= 32 ? 4 : (len <= 128 ? 8 : 16)); }
constexpr int kMaxParts = 16;
constexpr int kCoreBlock = 64;     // rows per block of the dense-core substitution
constexpr int kCoreMaxRows = 16384; // largest dense core per graph (its trailing small levels beyond stay sparse)
constexpr int kLevelWaves = 16;
constexpr int kXcds = 8;           // MI355X: workgroups are dealt to the 8 XCDs round robin   // wavefronts per block of a per-level launch (one long row, or 16 short rows)

}  // namespace ss
