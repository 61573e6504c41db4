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

#include "krylov.hpp"

namespace mpa {

typedef double dvec2 __attribute__((ext_vector_type(2)));
typedef float fvec4 __attribute__((ext_vector_type(4)));

template <typename T> struct Chunk;
template <> struct Chunk<double> {
    using v = dvec2;
    static constexpr int C = 1;  // complex entries per 16-byte chunk
};
template <> struct Chunk<float> {
    using v = fvec4;
    static constexpr int C = 2;
};

__device__ inline void chunk_fma(double& ar, double& ai, dvec2 a, dvec2 x) {
    ar = fma(a.x, x.x, ar);
    ar = fma(-a.y, x.y, ar);
    ai = fma(a.x, x.y, ai);
    ai = fma(a.y, x.x, ai);
}
__device__ inline void chunk_fma(float& ar, float& ai, fvec4 a, fvec4 x) {
    ar = fmaf(a.x, x.x, ar);
    ar = fmaf(-a.y, x.y, ar);
    ai = fmaf(a.x, x.y, ai);
    ai = fmaf(a.y, x.x, ai);
    ar = fmaf(a.z, x.z, ar);
    ar = fmaf(-a.w, x.w, ar);
    ai = fmaf(a.z, x.w, ai);
    ai = fmaf(a.w, x.z, ai);
}

__device__ inline bool all_done(const SysDev& d) {
    for (int s = 0; s < d.S; s++)
        if (!d.scal[s].done) return false;
    return true;
}

// One workgroup's share of a matvec job: its four waves own R rows each,
// starting at row 4 R rb; SG right-hand sides per pass, U chunks in flight.
template <typename T, int R, int SG, int U>
__device__ inline void gemv_rows(const MvDev& J, int rb) {
    using V = typename Chunk<T>::v;
    constexpr int C = Chunk<T>::C;
    const int n = J.n;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int row0 = (rb * 4 + w) * R;
    if (row0 >= n) return;
    const int nv = n / C;  // 16-byte chunks per row (n is even)
    const T* A = (const T*)J.A;
    const V* Ar[R];
#pragma unroll
    for (int r = 0; r < R; r++) Ar[r] = (const V*)(A + (size_t)2 * n * min(row0 + r, n - 1));
    for (int s0 = 0; s0 < J.S; s0 += SG) {
        T ar[R][SG], ai[R][SG];
#pragma unroll
        for (int r = 0; r < R; r++)
#pragma unroll
            for (int s = 0; s < SG; s++) ar[r][s] = ai[r][s] = 0;
        const V* xs[SG];
#pragma unroll
        for (int s = 0; s < SG; s++) xs[s] = (const V*)((const T*)J.x + (size_t)2 * n * min(s0 + s, J.S - 1));
        int c = lane;
        for (; c + 64 * (U - 1) < nv; c += 64 * U) {
            V xv[SG][U];
#pragma unroll
            for (int s = 0; s < SG; s++)
#pragma unroll
                for (int u = 0; u < U; u++) xv[s][u] = xs[s][c + 64 * u];
            V av[R][U];
#pragma unroll
            for (int r = 0; r < R; r++)
#pragma unroll
                for (int u = 0; u < U; u++) av[r][u] = __builtin_nontemporal_load(Ar[r] + c + 64 * u);
#pragma unroll
            for (int r = 0; r < R; r++)
#pragma unroll
                for (int s = 0; s < SG; s++)
#pragma unroll
                    for (int u = 0; u < U; u++) chunk_fma(ar[r][s], ai[r][s], av[r][u], xv[s][u]);
        }
        for (; c < nv; c += 64) {
            V xv[SG];
#pragma unroll
            for (int s = 0; s < SG; s++) xv[s] = xs[s][c];
#pragma unroll
            for (int r = 0; r < R; r++) {
                const V a = __builtin_nontemporal_load(Ar[r] + c);
#pragma unroll
                for (int s = 0; s < SG; s++) chunk_fma(ar[r][s], ai[r][s], a, xv[s]);
            }
        }
#pragma unroll
        for (int r = 0; r < R; r++)
#pragma unroll
            for (int s = 0; s < SG; s++) {
                T re = ar[r][s], im = ai[r][s];
                for (int off = 32; off > 0; off >>= 1) {
                    re += __shfl_xor(re, off);
                    im += __shfl_xor(im, off);
                }
                const int row = row0 + r;
                if (lane == 0 && row < n && s0 + s < J.S) {
                    const size_t o = (size_t)(s0 + s) * n + row;
                    cplx_t<T> yv = cmake<T>(re, im);
                    if (J.add_identity) yv = cadd(yv, ((const cplx_t<T>*)J.x)[o]);
                    ((cplx_t<T>*)J.y)[o] = yv;
                }
            }
    }
}

// batched: one workgroup per (job, row block) of the work list
template <typename T, int RB, int SG, int U>
__global__ __launch_bounds__(kThreads) void k_gemv(const MvDev* __restrict__ jobs, const SysDev* __restrict__ sys,
                                                   const int2* __restrict__ work) {
    const int2 wk = work[blockIdx.x];
    const MvDev J = jobs[wk.x];
    if (J.skip >= 0 && all_done(sys[J.skip])) return;
    gemv_rows<T, RB, SG, U>(J, wk.y);
}

// one job passed by value: workgroup b owns row block b
template <typename T, int RB, int SG, int U>
__global__ __launch_bounds__(kThreads) void k_gemv1(MvDev J) {
    gemv_rows<T, RB, SG, U>(J, blockIdx.x);
}

// Tile classes: systems are grouped by RHS count so that every launch uses
// one template instance.  Rows per wave x RHS per pass: class 0 (S = 1)
// 8 x 1, class 1 (S = 2) 4 x 2, class 2 (S = 3..4) 2 x 4, class 3 (S >= 5)
// 1 x 8 (one pass over the matrix for up to 8 RHS); a workgroup is four
// waves, so it covers 4 R rows.
constexpr int kTileClasses = 4;
inline int gemv_class(int S) { return S == 1 ? 0 : (S == 2 ? 1 : (S <= 4 ? 2 : 3)); }
inline int gemv_rows_of_class(int c) { return 4 * (c == 0 ? 8 : (c == 1 ? 4 : (c == 2 ? 2 : 1))); }
inline int gemv_row_blocks(int n, int S) {
    const int rb = gemv_rows_of_class(gemv_class(S));
    return (n + rb - 1) / rb;
}

// Launch the batched matvec: one kernel per non-empty tile class.
template <typename T>
int gemv_launch(const MvDev* jobs, const SysDev* sys, int2* const* work, const int* nwork, hipStream_t st) {
    ProfScope prof(sizeof(T) == 8 ? 0 : 1, st);
    if (nwork[0]) hipLaunchKernelGGL((k_gemv<T, 8, 1, 2>), dim3(nwork[0]), dim3(kThreads), 0, st, jobs, sys, work[0]);
    if (nwork[1]) hipLaunchKernelGGL((k_gemv<T, 4, 2, 2>), dim3(nwork[1]), dim3(kThreads), 0, st, jobs, sys, work[1]);
    if (nwork[2]) hipLaunchKernelGGL((k_gemv<T, 2, 4, 2>), dim3(nwork[2]), dim3(kThreads), 0, st, jobs, sys, work[2]);
    if (nwork[3]) hipLaunchKernelGGL((k_gemv<T, 1, 8, 2>), dim3(nwork[3]), dim3(kThreads), 0, st, jobs, sys, work[3]);
    MPA_LAUNCH_CHECK();
    return MPA_OK;
}

// One job without device-side descriptors (no workspace needed).
template <typename T>
int gemv_launch1(const MvDev& J, hipStream_t st) {
    ProfScope prof(sizeof(T) == 8 ? 0 : 1, st);
    const unsigned nb = (unsigned)gemv_row_blocks(J.n, J.S);
    const int c = gemv_class(J.S);
    if (c == 0) hipLaunchKernelGGL((k_gemv1<T, 8, 1, 2>), dim3(nb), dim3(kThreads), 0, st, J);
    else if (c == 1) hipLaunchKernelGGL((k_gemv1<T, 4, 2, 2>), dim3(nb), dim3(kThreads), 0, st, J);
    else if (c == 2) hipLaunchKernelGGL((k_gemv1<T, 2, 4, 2>), dim3(nb), dim3(kThreads), 0, st, J);
    else hipLaunchKernelGGL((k_gemv1<T, 1, 8, 2>), dim3(nb), dim3(kThreads), 0, st, J);
    MPA_LAUNCH_CHECK();
    return MPA_OK;
}

}  // namespace mpa
