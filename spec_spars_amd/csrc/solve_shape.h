// Summation shape of one row of a level-scheduled triangular solve.
//
// A row's entries are summed in P parts: entry q of a row whose entries
// start at q0 goes to part (q - q0) mod P, each part is a running
// subtraction from zero in entry order, the parts are added left to right
// and the result is added to the row's start value.  P = 1 is a single
// running subtraction from the start value.  P depends on the row's length
// alone, so the per-level kernels and the single-workgroup small-level
// kernels give a row the same bits whichever of them finishes it, and a
// graph's solve does not change when other graphs share its launches.
#pragma once

namespace ss {

constexpr int kLongRow = 64;   // rows longer than this are summed in parts

constexpr int row_parts(int len) { return len <= kLongRow ? 1 : (len <= 16 * kLongRow ? 4 : 16); }

}  // namespace ss
