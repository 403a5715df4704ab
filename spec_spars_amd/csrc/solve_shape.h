// Summation shape of one row of a level-scheduled triangular solve.
//
// A row with start value acc and entries [a, b) in schedule order is
// computed as follows.  With P = seg_parts(b - a) == 1, acc -= val[q] *
// x[col[q]] for q in order; otherwise part s (0 <= s < P) is a running
// subtraction from zero over q = a + s, a + s + P, ..., the P parts are
// folded as a balanced pairwise tree (fold(lo..hi) = fold(left half) +
// fold(right half)) and acc += that fold.  Every accumulate is one fused
// multiply-add.
//
// A shape that depends only on the row means any kernel may finish it --
// one wave holding every part in registers, a block of waves sharing the
// parts, several workgroups each folding 16 parts (rows of more than 1024
// entries), the small-level kernel -- with the same bits, whatever other
// rows or graphs share the launch.
#pragma once

namespace ss {

constexpr int seg_parts(int len) {
    return len <= 8 ? 1 : (len <= 32 ? 4 : (len <= 128 ? 8 : (len <= 1024 ? 16 : (len <= 4096 ? 64 : 128))));
}
constexpr int kWgParts = 16;      // parts one workgroup folds (16 waves, a part each at most)
constexpr int kWgRowMax = 1024;   // the longest row of kWgParts parts
constexpr int kMaxParts = 128;    // rows past kWgRowMax: kMaxParts / kWgParts workgroups at most
constexpr int kMaxWg = kMaxParts / kWgParts;   // workgroups of one row, at most
// Workgroups that fold a row of len entries.
constexpr int row_wgs(int len) { return seg_parts(len) > kWgParts ? seg_parts(len) / kWgParts : 1; }
static_assert(row_wgs(1 << 30) == kMaxWg, "the longest rows take kMaxWg workgroups");
static_assert(seg_parts(kWgRowMax) == kWgParts && seg_parts(kWgRowMax + 1) > kWgParts, "one workgroup per row up to kWgRowMax");

// Row classes of a per-level launch.  In a level of at least kWaveLevelRows
// rows, rows of up to kWaveRowMax entries take one wave each; in smaller
// levels only rows of up to kLongRow entries do.  Longer rows get a block of
// waves sharing their parts.  Rows of at most kTinyRow entries (one part) go
// several to a wave, kTinyGroup at a time.
constexpr int kLongRow = 32;
constexpr int kWaveRowMax = 128;
constexpr int kWaveLevelRows = 2048;
constexpr int kTinyRow = 8;
constexpr int kTinyGroup = 4;
static_assert(seg_parts(kTinyRow) == 1 && seg_parts(kTinyRow + 1) > 1, "tiny rows are the one-part rows");
static_assert(seg_parts(kWaveRowMax) <= 8 && kWaveRowMax <= 128, "a wave row has at most 8 parts and 2 entries per lane");

constexpr int kMaxK = 256;        // sketch columns / right-hand sides per call, at most

constexpr int kSmallRows = 16;    // levels with at most this many rows ...
constexpr int kSmallNnz = 2048;   // ... and entries go to the single-workgroup kernel
static_assert(kSmallRows == 16, "the small-level kernel runs 16 waves, at most one row each");

constexpr int kCoreTile = 64;       // rows / columns of a dense-core tile
constexpr int kCoreSeg = 2048;      // inner rows of one dense-core product piece, at most
constexpr int kCoreMaxRows = 4096;  // largest dense core per graph taken from trailing small levels
                                    // (a factor's dense tail can be larger)
constexpr int kLevelWaves = 16;     // wavefronts per block of a per-level launch (at most)
constexpr int kXcds = 8;            // MI355X: workgroups are dealt to the 8 XCDs round robin

}  // namespace ss
