// Batched complex matvec y = (I +) A x over many systems, S RHS each.
//
// The system matrix is streamed from HBM exactly once per call with 16-byte
// nontemporal loads (one c128 or two c64 per lane; the matrix is not reused
// within a call, so it is kept out of the caches that hold x).  Each wave of
// a workgroup owns R consecutive rows of one system and runs along the full
// rows with U chunks per lane in flight; the x chunk loaded for a column is
// reused for the R rows and for up to SG right-hand sides held in registers.
// Measured on MI355X at n = 30000 this tile streams c128 at 6.9 TB/s and c64
// at 6.5 TB/s, above a plain streaming-read kernel.  Jobs whose system has
// converged for all right-hand sides return before touching the matrix.
#pragma once

#include <algorithm>
#include <vector>

#include "krylov.hpp"

namespace mpa {

// Tile classes: systems are grouped by RHS count so that every launch uses
// one kernel instance: class 0 (S = 1), 1 (S = 2), 2 (S = 3..4), 3 (S >= 5).
// Each class has a list of tile variants (vector-ALU row tiles of R rows per
// wave x SG RHS per pass x U chunks in flight, or the matrix-core GEMM); the
// variant in use is per (element type, class), set by the autotuner
// (mpa_autotune_matvec) and defaulting to the measured best.
constexpr int kTileClasses = 4;
int gemv_class(int S);
// rows per workgroup of class c for element size es (8: c128, 4: c64)
int gemv_rows_of_class(int c, int es);
int gemv_row_blocks(int n, int S, int es);

// Batched matvec: one launch per non-empty class over its (job, row block)
// work list.  `sys` may be null when no job has skip >= 0.
int gemv_launch(int es, const MvDev* jobs, const SysDev* sys, int2* const* work, const int* nwork, hipStream_t st);

// A batched GEMM's work list (job, row block) with the longest rows first
// (len(job) = its row length): workgroups over big matrices start early and
// those over small ones fill the tail.  Row blocks of one job keep their order.
// (batch256's FP64 batch: the c128 GEMM 396 -> 371-384 ms per step.)
template <class F>
inline void order_work_longest_first(std::vector<int2>& mw, F len) {
    std::stable_sort(mw.begin(), mw.end(), [&](const int2& a, const int2& b) { return len(a.x) > len(b.x); });
}
// One job without device-side descriptors.
int gemv_launch1(int es, const MvDev& J, hipStream_t st);

}  // namespace mpa
