// Device-side state and helpers shared by the batched Krylov kernels.
#pragma once

#include "common.hpp"

namespace mpa {

constexpr int kThreads = 256;
// partial-sum slots per (system, RHS, workgroup):
//   0 <r0,v>, 1 <t,s>, 2 <t,t>, 3 ||r||^2 (also ||b||^2 at start), 4 <r0,r>
constexpr int kSlots = 5;
constexpr int kMaxG = 240;  // most workgroups per (system, RHS) in vector kernels

// Per-RHS BiCGSTAB scalars, device resident.
struct Scal {
    double2 rho, alpha, w;  // committed values of the current iteration
    double2 rho_next, alpha_next, w_next;
    double r0_sq, rhs_sq, rnorm2, tol2;
    int it, done, status, restarts;
    int maxiter, restart_pending, init_pending, pad1;  // init_pending: x0 = b, r = b - A b not yet formed
};

// One linear system (I + T) x = b with S right-hand sides, as the kernels see it.
struct SysDev {
    const void* A;  // [n, n] row-major, dtype of the solve
    const void* b;  // [S, n]
    void* x;        // [S, n]
    void* r;
    void* r0;
    void* p;
    void* v;
    void* s;
    void* t;        // [S, n] each
    Scal* scal;     // [S]
    double2* part;  // [S][G][kSlots]
    void* y;        // M^-1 p (the matvec operand of the first half); aliases p without a preconditioner
    void* z;        // M^-1 s (second half); aliases s without a preconditioner
    const void* minv;  // [n, bs] stack of block inverses of the right preconditioner, or null
    double tol;
    int n, S, G, add_identity;
    int maxiter, flags;  // MPA_SOLVE_* bits
    int bs, pad;
};

// One y = (I +) A x job of a batched matvec.
struct MvDev {
    const void* A;  // [n, n]
    const void* x;  // [S, n]
    void* y;        // [S, n]
    int n, S, add_identity, skip;  // skip: index of a SysDev whose all-done flag skips the job, or -1
};

inline size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

// Workgroups per RHS for the vector kernels of an n-long system: about 1024
// elements (4 per thread) each, so a single system of moderate n still
// spreads over tens of CUs and each thread's loads overlap.
inline int groups_for(int n) {
    int g = (n + 1023) / 1024;
    return g < 1 ? 1 : (g > kMaxG ? kMaxG : g);
}

__device__ inline double2 to_d(double2 z) { return z; }
__device__ inline double2 to_d(float2 z) { return make_double2(z.x, z.y); }
template <typename T> __device__ inline cplx_t<T> from_d(double2 z);
template <> __device__ inline double2 from_d<double>(double2 z) { return z; }
template <> __device__ inline float2 from_d<float>(double2 z) { return make_float2((float)z.x, (float)z.y); }

// block-wide sum of K complex values; the result is valid in thread 0
template <int K>
__device__ inline void block_sum(double2 (&v)[K]) {
    __shared__ double2 red[K][kThreads / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < K; k++) {
        double re = v[k].x, im = v[k].y;
        for (int off = 32; off > 0; off >>= 1) {
            re += __shfl_xor(re, off);
            im += __shfl_xor(im, off);
        }
        if (lane == 0) red[k][wid] = make_double2(re, im);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; k++) {
            double2 s = red[k][0];
            for (int w = 1; w < kThreads / 64; w++) s = cadd(s, red[k][w]);
            v[k] = s;
        }
    }
    __syncthreads();
}

// Fixed-order sum of the G partials of `slot` for RHS s of a system; the
// value is returned to every thread of the block.
__device__ inline double2 reduce_partials(const double2* part, int G, int s, int slot) {
    __shared__ double2 tmp[kThreads / 64];
    __shared__ double2 result;
    double2 acc = make_double2(0, 0);
    for (int g = threadIdx.x; g < G; g += kThreads) acc = cadd(acc, part[((size_t)s * G + g) * kSlots + slot]);
    double re = acc.x, im = acc.y;
    for (int off = 32; off > 0; off >>= 1) {
        re += __shfl_xor(re, off);
        im += __shfl_xor(im, off);
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) tmp[wid] = make_double2(re, im);
    __syncthreads();
    if (threadIdx.x == 0) {
        double2 t = tmp[0];
        for (int w = 1; w < kThreads / 64; w++) t = cadd(t, tmp[w]);
        result = t;
    }
    __syncthreads();
    return result;
}

}  // namespace mpa
