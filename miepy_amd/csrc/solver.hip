// Dense complex matvec and BiCGSTAB for the cluster system (I + T) x = b.
//
// Replaces the solve of cpp/src/interactions.cpp:347-368: Eigen::BiCGSTAB with
// the identity preconditioner over an OpenMP row-block GEMV (:80-91).  The
// iteration is Eigen's (IterativeLinearSolvers/BiCGSTAB.h): zero initial
// guess, stop when ||r||^2 <= tol^2 ||b||^2 or after maxiter iterations,
// restart when rho falls below eps^2 ||r0||^2.  The reference hard-codes
// tol 1e-5 / maxiter 1000 (:362-363); here both are arguments.
//
// Layout: the matrix is row-major [n, n] (c128 or c64); the S right-hand
// sides are RHS-major [S, n].  All iteration scalars live on the device, one
// record per RHS, so the host can enqueue many iterations without reading
// anything back: every kernel returns early for an RHS whose `done` flag is
// set.  Dot products are two-stage and deterministic (per-workgroup partials
// written to the workspace, re-reduced in a fixed order by the consumer).

#include "common.hpp"

#include <cmath>
#include <cstring>
#include <vector>

namespace {

constexpr int kThreads = 256;
constexpr int kVecBlocks = 240;  // vector-kernel workgroups per RHS (partials per RHS)
// partial-sum slots: 0 <r0,v>, 1 <t,s>, 2 <t,t>, 3 ||r||^2, 4 <r0,r>
constexpr int kSlots = 5;

// --------------------------------------------------------------------------
// matvec: y = (add_identity ? x : 0) + A x, for S RHS sharing one pass over A
// --------------------------------------------------------------------------

template <typename T> struct Acc { using type = T; };

// Each wave owns R consecutive rows; its lanes stride the columns.  The
// x values of a column chunk are loaded once per wave and reused for the R
// rows; A is streamed with 16-byte loads (one c128 or two c64 per lane).
template <typename T, int R, int S>
__global__ __launch_bounds__(kThreads) void gemv_kernel(int n, const cplx_t<T>* __restrict__ A,
                                                         const cplx_t<T>* __restrict__ x, cplx_t<T>* __restrict__ y,
                                                         int add_identity) {
    using C = cplx_t<T>;
    const int lane = threadIdx.x & 63;
    const int wave = (blockIdx.x * kThreads + threadIdx.x) >> 6;
    const int row0 = wave * R;
    if (row0 >= n) return;
    C acc[R][S];
#pragma unroll
    for (int r = 0; r < R; r++)
#pragma unroll
        for (int s = 0; s < S; s++) acc[r][s] = cmake<T>(0, 0);

    constexpr int V = (sizeof(T) == 8) ? 1 : 2;  // complex entries per 16-byte load
    const int step = 64 * V;
    const int nfull = (n / step) * step;
#pragma unroll 4
    for (int c = lane * V; c < nfull; c += step) {
        C xv[S][V];
#pragma unroll
        for (int s = 0; s < S; s++)
#pragma unroll
            for (int u = 0; u < V; u++) xv[s][u] = x[(size_t)s * n + c + u];
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int row = row0 + r;
            if (row < n) {
                const C* Ar = A + (size_t)row * n + c;
                C av[V];
#pragma unroll
                for (int u = 0; u < V; u++) av[u] = Ar[u];
#pragma unroll
                for (int s = 0; s < S; s++)
#pragma unroll
                    for (int u = 0; u < V; u++) cfma(acc[r][s], av[u], xv[s][u]);
            }
        }
    }
    for (int c = nfull + lane; c < n; c += 64) {
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int row = row0 + r;
            if (row < n) {
                const C a = A[(size_t)row * n + c];
#pragma unroll
                for (int s = 0; s < S; s++) cfma(acc[r][s], a, x[(size_t)s * n + c]);
            }
        }
    }
#pragma unroll
    for (int r = 0; r < R; r++)
#pragma unroll
        for (int s = 0; s < S; s++) {
            T re = acc[r][s].x, im = acc[r][s].y;
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                re += __shfl_xor(re, off);
                im += __shfl_xor(im, off);
            }
            acc[r][s].x = re;
            acc[r][s].y = im;
        }
    if (lane == 0) {
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int row = row0 + r;
            if (row < n) {
#pragma unroll
                for (int s = 0; s < S; s++) {
                    C v = acc[r][s];
                    if (add_identity) v = cadd(v, x[(size_t)s * n + row]);
                    y[(size_t)s * n + row] = v;
                }
            }
        }
    }
}

template <typename T, int R, int S>
void launch_gemv(int n, const void* A, const void* x, void* y, int add_identity, hipStream_t st) {
    const int waves = (n + R - 1) / R;
    const int blocks = (waves + 3) / 4;
    hipLaunchKernelGGL((gemv_kernel<T, R, S>), dim3(blocks), dim3(kThreads), 0, st, n, (const cplx_t<T>*)A,
                       (const cplx_t<T>*)x, (cplx_t<T>*)y, add_identity);
}

// y = op(A) x for S RHS; S is processed in groups of 1, 2, 4 or 8.
template <typename T>
int matvec(int n, int S, const void* A, int add_identity, const void* x, void* y, hipStream_t st) {
    using C = cplx_t<T>;
    ProfScope prof(sizeof(T) == 8 ? 0 : 1, st);
    const C* xs = (const C*)x;
    C* ys = (C*)y;
    int s0 = 0;
    while (s0 < S) {
        const int left = S - s0;
        const void* xp = xs + (size_t)s0 * n;
        void* yp = ys + (size_t)s0 * n;
        if (left >= 8) {
            launch_gemv<T, 1, 8>(n, A, xp, yp, add_identity, st);
            s0 += 8;
        } else if (left >= 4) {
            launch_gemv<T, 1, 4>(n, A, xp, yp, add_identity, st);
            s0 += 4;
        } else if (left >= 2) {
            launch_gemv<T, 2, 2>(n, A, xp, yp, add_identity, st);
            s0 += 2;
        } else {
            launch_gemv<T, 4, 1>(n, A, xp, yp, add_identity, st);
            s0 += 1;
        }
    }
    MPA_LAUNCH_CHECK();
    return MPA_OK;
}

// --------------------------------------------------------------------------
// BiCGSTAB state
// --------------------------------------------------------------------------

struct Scal {
    double2 rho, alpha, w;        // committed values of the current iteration
    double2 rho_next, alpha_next, w_next;
    double r0_sq, rhs_sq, rnorm2, tol2;
    int it, done, status, restarts;
    int maxiter, restart_pending, pad0, pad1;
};

template <typename T>
struct Work {
    cplx_t<T>*r, *r0, *p, *v, *s, *t;
    double2* part;  // [S][kVecBlocks][kSlots] partial sums (complex, or real in .x)
    Scal* scal;     // [S]
};

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

template <typename T>
size_t work_bytes(int n, int S) {
    return 6 * align256(sizeof(cplx_t<T>) * (size_t)n * S) + align256(sizeof(double2) * (size_t)S * kVecBlocks * kSlots) +
           align256(sizeof(Scal) * (size_t)S);
}

template <typename T>
Work<T> carve(void* ws, int n, int S) {
    Work<T> w;
    char* p = (char*)ws;
    const size_t vb = align256(sizeof(cplx_t<T>) * (size_t)n * S);
    w.r = (cplx_t<T>*)p; p += vb;
    w.r0 = (cplx_t<T>*)p; p += vb;
    w.p = (cplx_t<T>*)p; p += vb;
    w.v = (cplx_t<T>*)p; p += vb;
    w.s = (cplx_t<T>*)p; p += vb;
    w.t = (cplx_t<T>*)p; p += vb;
    w.part = (double2*)p; p += align256(sizeof(double2) * (size_t)S * kVecBlocks * kSlots);
    w.scal = (Scal*)p;
    return w;
}

__device__ inline double2 to_d(double2 z) { return z; }
__device__ inline double2 to_d(float2 z) { return make_double2(z.x, z.y); }
template <typename T> __device__ inline cplx_t<T> from_d(double2 z);
template <> __device__ inline double2 from_d<double>(double2 z) { return z; }
template <> __device__ inline float2 from_d<float>(double2 z) { return make_float2((float)z.x, (float)z.y); }

// block-wide sum of K complex values; result valid in thread 0
template <int K>
__device__ void block_sum(double2 (&v)[K]) {
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

// Fixed-order reduction of the kVecBlocks partials of slot `slot` for RHS s.
__device__ double2 reduce_partials(const double2* part, int s, int slot) {
    __shared__ double2 tmp[kThreads / 64];
    __shared__ double2 result;
    double2 acc = make_double2(0, 0);
    for (int b = threadIdx.x; b < kVecBlocks; b += kThreads) acc = cadd(acc, part[((size_t)s * kVecBlocks + b) * kSlots + slot]);
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

// init: x = 0, r = r0 = b, p = v = 0; partials of ||b||^2
template <typename T>
__global__ __launch_bounds__(kThreads) void k_init(int n, const cplx_t<T>* __restrict__ b, cplx_t<T>* __restrict__ x,
                                                    Work<T> w) {
    const int s = blockIdx.y;
    double2 acc[1] = {make_double2(0, 0)};
    for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += kVecBlocks * kThreads) {
        const size_t k = (size_t)s * n + i;
        const cplx_t<T> bv = b[k];
        w.r[k] = bv;
        w.r0[k] = bv;
        const cplx_t<T> z = cmake<T>(0, 0);
        w.p[k] = z;
        w.v[k] = z;
        x[k] = z;
        const double2 bd = to_d(bv);
        acc[0].x += bd.x * bd.x + bd.y * bd.y;
    }
    block_sum<1>(acc);
    if (threadIdx.x == 0) w.part[((size_t)s * kVecBlocks + blockIdx.x) * kSlots + 3] = acc[0];
}

__global__ void k_init_scalars(int S, double tol, int maxiter, Scal* scal, const double2* part) {
    const int s = blockIdx.x;
    double2 bb = reduce_partials(part, s, 3);
    if (threadIdx.x == 0) {
        Scal& c = scal[s];
        c.rho = c.alpha = c.w = make_double2(1, 0);
        c.rho_next = c.alpha_next = c.w_next = make_double2(1, 0);
        c.rhs_sq = bb.x;
        c.r0_sq = bb.x;
        c.rnorm2 = bb.x;
        c.tol2 = tol * tol * bb.x;
        c.it = 0;
        c.restarts = 0;
        c.maxiter = maxiter;
        c.status = MPA_OK;
        c.restart_pending = 0;
        // rho for the first iteration is <r0, r> = ||b||^2 (x0 = 0)
        c.rho_next = make_double2(bb.x, 0);
        c.done = (bb.x == 0.0 || !(bb.x > c.tol2) || maxiter <= 0) ? 1 : 0;
        if (bb.x != bb.x) {
            c.done = 1;
            c.status = MPA_NONFINITE;
        }
    }
}

// B: rho_new = <r0, r> (from the previous H, or init); beta; p = r + beta (p - w v)
template <typename T>
__global__ __launch_bounds__(kThreads) void k_update_p(int n, Work<T> w) {
    const int s = blockIdx.y;
    const Scal& c = w.scal[s];
    if (c.done || c.restart_pending) return;
    const double2 rho_new = c.rho_next;
    const double2 beta = cmul(cdiv(rho_new, c.rho), cdiv(c.alpha, c.w));
    const cplx_t<T> bt = from_d<T>(beta), wt = from_d<T>(c.w);
    for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += kVecBlocks * kThreads) {
        const size_t k = (size_t)s * n + i;
        cplx_t<T> q = csub(w.p[k], cmul(wt, w.v[k]));
        w.p[k] = cadd(w.r[k], cmul(bt, q));
    }
}

// D: partial <r0, v>
template <typename T>
__global__ __launch_bounds__(kThreads) void k_dot_r0v(int n, Work<T> w) {
    const int s = blockIdx.y;
    if (w.scal[s].done || w.scal[s].restart_pending) return;
    double2 acc[1] = {make_double2(0, 0)};
    for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += kVecBlocks * kThreads) {
        const size_t k = (size_t)s * n + i;
        acc[0] = cadd(acc[0], cmulc(to_d(w.r0[k]), to_d(w.v[k])));
    }
    block_sum<1>(acc);
    if (threadIdx.x == 0) w.part[((size_t)s * kVecBlocks + blockIdx.x) * kSlots + 0] = acc[0];
}

// E: alpha = rho / <r0, v>; s = r - alpha v
template <typename T>
__global__ __launch_bounds__(kThreads) void k_update_s(int n, Work<T> w) {
    const int s = blockIdx.y;
    if (w.scal[s].done || w.scal[s].restart_pending) return;
    const double2 d = reduce_partials(w.part, s, 0);
    const double2 alpha = cdiv(w.scal[s].rho_next, d);
    const cplx_t<T> at = from_d<T>(alpha);
    for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += kVecBlocks * kThreads) {
        const size_t k = (size_t)s * n + i;
        w.s[k] = csub(w.r[k], cmul(at, w.v[k]));
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) w.scal[s].alpha_next = alpha;
}

// G: partials <t, s>, <t, t>
template <typename T>
__global__ __launch_bounds__(kThreads) void k_dot_ts(int n, Work<T> w) {
    const int s = blockIdx.y;
    if (w.scal[s].done || w.scal[s].restart_pending) return;
    double2 acc[2] = {make_double2(0, 0), make_double2(0, 0)};
    for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += kVecBlocks * kThreads) {
        const size_t k = (size_t)s * n + i;
        const double2 tv = to_d(w.t[k]);
        acc[0] = cadd(acc[0], cmulc(tv, to_d(w.s[k])));
        acc[1].x += tv.x * tv.x + tv.y * tv.y;
    }
    block_sum<2>(acc);
    if (threadIdx.x == 0) {
        w.part[((size_t)s * kVecBlocks + blockIdx.x) * kSlots + 1] = acc[0];
        w.part[((size_t)s * kVecBlocks + blockIdx.x) * kSlots + 2] = acc[1];
    }
}

// H: w = <t,s>/<t,t>; x += alpha p + w s; r = s - w t; partials ||r||^2, <r0, r>
template <typename T>
__global__ __launch_bounds__(kThreads) void k_update_xr(int n, cplx_t<T>* __restrict__ x, Work<T> w) {
    const int s = blockIdx.y;
    if (w.scal[s].done || w.scal[s].restart_pending) return;
    const double2 ts = reduce_partials(w.part, s, 1);
    const double2 tt = reduce_partials(w.part, s, 2);
    const double2 om = tt.x > 0 ? cscale(ts, 1.0 / tt.x) : make_double2(0, 0);
    const cplx_t<T> at = from_d<T>(w.scal[s].alpha_next), wt = from_d<T>(om);
    double2 acc[2] = {make_double2(0, 0), make_double2(0, 0)};
    for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += kVecBlocks * kThreads) {
        const size_t k = (size_t)s * n + i;
        const cplx_t<T> sv = w.s[k];
        x[k] = cadd(x[k], cadd(cmul(at, w.p[k]), cmul(wt, sv)));
        const cplx_t<T> rv = csub(sv, cmul(wt, w.t[k]));
        w.r[k] = rv;
        const double2 rd = to_d(rv);
        acc[0].x += rd.x * rd.x + rd.y * rd.y;
        acc[1] = cadd(acc[1], cmulc(to_d(w.r0[k]), rd));
    }
    block_sum<2>(acc);
    if (threadIdx.x == 0) {
        w.part[((size_t)s * kVecBlocks + blockIdx.x) * kSlots + 3] = acc[0];
        w.part[((size_t)s * kVecBlocks + blockIdx.x) * kSlots + 4] = acc[1];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) w.scal[s].w_next = om;
}

// I: commit the iteration's scalars, convergence / restart checks (one block per RHS)
__global__ void k_commit(Scal* scal, const double2* part) {
    const int s = blockIdx.x;
    if (scal[s].done || scal[s].restart_pending) return;
    const double2 rr = reduce_partials(part, s, 3);
    const double2 r0r = reduce_partials(part, s, 4);
    if (threadIdx.x == 0) {
        Scal& c = scal[s];
        c.rho = c.rho_next;
        c.alpha = c.alpha_next;
        c.w = c.w_next;
        c.rnorm2 = rr.x;
        c.it += 1;
        c.rho_next = r0r;
        const double eps = 2.220446049250313e-16;  // NumTraits<double>::epsilon()
        if (!(rr.x == rr.x) || !(r0r.x == r0r.x) || !(r0r.y == r0r.y)) {
            c.done = 1;
            c.status = MPA_NONFINITE;
        } else if (!(rr.x > c.tol2)) {
            c.done = 1;
        } else if (c.it >= c.maxiter) {
            c.done = 1;
            c.status = MPA_NOT_CONVERGED;
        } else if (sqrt(r0r.x * r0r.x + r0r.y * r0r.y) < eps * eps * c.r0_sq) {
            c.restart_pending = 1;  // the host re-seeds r0 = r = b - A x
        }
    }
}

// restart support: r = b - y (y = A x already computed), r0 = r, partial ||r||^2
template <typename T>
__global__ __launch_bounds__(kThreads) void k_restart(int n, const cplx_t<T>* __restrict__ b,
                                                       const cplx_t<T>* __restrict__ y, Work<T> w) {
    const int s = blockIdx.y;
    if (!w.scal[s].restart_pending) return;
    double2 acc[1] = {make_double2(0, 0)};
    for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += kVecBlocks * kThreads) {
        const size_t k = (size_t)s * n + i;
        const cplx_t<T> rv = csub(b[k], y[k]);
        w.r[k] = rv;
        w.r0[k] = rv;
        const double2 rd = to_d(rv);
        acc[0].x += rd.x * rd.x + rd.y * rd.y;
    }
    block_sum<1>(acc);
    if (threadIdx.x == 0) w.part[((size_t)s * kVecBlocks + blockIdx.x) * kSlots + 3] = acc[0];
}

__global__ void k_restart_commit(Scal* scal, const double2* part) {
    const int s = blockIdx.x;
    if (!scal[s].restart_pending) return;
    const double2 rr = reduce_partials(part, s, 3);
    if (threadIdx.x == 0) {
        Scal& c = scal[s];
        c.r0_sq = rr.x;
        // Eigen: rho = r0_sqnorm = ||r||^2 replaces this iteration's rho; the
        // committed rho stays as rho_old for beta = (rho/rho_old)(alpha/w).
        c.rho_next = make_double2(rr.x, 0);
        if (c.restarts++ == 0) c.it = 0;
        c.restart_pending = 0;
    }
}

template <typename T>
int bicgstab_impl(int n, int S, const void* A, int add_identity, const void* b, void* x, double tol, int maxiter,
                  void* ws, int* iters_out, double* resid_out, int* status_out, int sync_every, hipStream_t st) {
    using C = cplx_t<T>;
    Work<T> w = carve<T>(ws, n, S);
    const dim3 grid(kVecBlocks, S), block(kThreads);
    hipLaunchKernelGGL(k_init<T>, grid, block, 0, st, n, (const C*)b, (C*)x, w);
    hipLaunchKernelGGL(k_init_scalars, dim3(S), block, 0, st, S, tol, maxiter, w.scal, w.part);
    MPA_LAUNCH_CHECK();
    std::vector<Scal> host(S);
    int launched = 0;
    const int sync_k = sync_every > 0 ? sync_every : 1;
    for (;;) {
        // one enqueued batch of sync_k iterations
        for (int k = 0; k < sync_k; k++) {
            hipLaunchKernelGGL(k_update_p<T>, grid, block, 0, st, n, w);
            int rc = matvec<T>(n, S, A, add_identity, w.p, w.v, st);
            if (rc) return rc;
            hipLaunchKernelGGL(k_dot_r0v<T>, grid, block, 0, st, n, w);
            hipLaunchKernelGGL(k_update_s<T>, grid, block, 0, st, n, w);
            rc = matvec<T>(n, S, A, add_identity, w.s, w.t, st);
            if (rc) return rc;
            hipLaunchKernelGGL(k_dot_ts<T>, grid, block, 0, st, n, w);
            hipLaunchKernelGGL(k_update_xr<T>, grid, block, 0, st, n, (C*)x, w);
            hipLaunchKernelGGL(k_commit, dim3(S), block, 0, st, w.scal, w.part);
            MPA_LAUNCH_CHECK();
        }
        launched += sync_k;
        MPA_HIP_CHECK(hipMemcpyAsync(host.data(), w.scal, sizeof(Scal) * S, hipMemcpyDeviceToHost, st));
        MPA_HIP_CHECK(hipStreamSynchronize(st));
        bool all_done = true, any_restart = false;
        for (int s = 0; s < S; s++) {
            all_done &= host[s].done != 0;
            any_restart |= host[s].restart_pending != 0;
        }
        if (any_restart) {
            // r = b - A x for the RHS that asked for it (Eigen's restart)
            int rc = matvec<T>(n, S, A, add_identity, x, w.t, st);
            if (rc) return rc;
            hipLaunchKernelGGL(k_restart<T>, grid, block, 0, st, n, (const C*)b, (const C*)w.t, w);
            hipLaunchKernelGGL(k_restart_commit, dim3(S), block, 0, st, w.scal, w.part);
            MPA_LAUNCH_CHECK();
            continue;
        }
        if (all_done) break;
        if (launched > 4 * (maxiter + 8)) {  // restarts reset the count once; never loop forever
            mpa_set_error("bicgstab: iteration budget exhausted");
            return MPA_NOT_CONVERGED;
        }
    }
    int worst = MPA_OK;
    for (int s = 0; s < S; s++) {
        if (iters_out) iters_out[s] = host[s].it;
        if (resid_out) resid_out[s] = host[s].rhs_sq > 0 ? std::sqrt(host[s].rnorm2 / host[s].rhs_sq) : 0.0;
        if (status_out) status_out[s] = host[s].status;
        if (host[s].status != MPA_OK && worst == MPA_OK) worst = host[s].status;
    }
    return worst;
}


// ---------------------------------------------------------------------------
// Mixed precision: complex64 BiCGSTAB corrections + FP64 iterative refinement
// ---------------------------------------------------------------------------

// r = b - y (FP64), partial ||r||^2 per RHS
__global__ __launch_bounds__(kThreads) void k_residual64(int n, const double2* __restrict__ b,
                                                        const double2* __restrict__ y, double2* __restrict__ r,
                                                        double2* __restrict__ part) {
    const int s = blockIdx.y;
    double2 acc[1] = {make_double2(0, 0)};
    for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += kVecBlocks * kThreads) {
        const size_t k = (size_t)s * n + i;
        const double2 rv = csub(b[k], y[k]);
        r[k] = rv;
        acc[0].x += rv.x * rv.x + rv.y * rv.y;
    }
    block_sum<1>(acc);
    if (threadIdx.x == 0) part[(size_t)s * kVecBlocks + blockIdx.x] = acc[0];
}

// partial ||v||^2 per RHS
__global__ __launch_bounds__(kThreads) void k_norm2(int n, const double2* __restrict__ v, double2* __restrict__ part) {
    const int s = blockIdx.y;
    double2 acc[1] = {make_double2(0, 0)};
    for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += kVecBlocks * kThreads) {
        const double2 z = v[(size_t)s * n + i];
        acc[0].x += z.x * z.x + z.y * z.y;
    }
    block_sum<1>(acc);
    if (threadIdx.x == 0) part[(size_t)s * kVecBlocks + blockIdx.x] = acc[0];
}

__global__ void k_norm_final(const double2* __restrict__ part, double* __restrict__ out) {
    const int s = blockIdx.x;
    __shared__ double red[kThreads];
    double a = 0;
    for (int b = threadIdx.x; b < kVecBlocks; b += kThreads) a += part[(size_t)s * kVecBlocks + b].x;
    red[threadIdx.x] = a;
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[s] = red[0];
}

// r32 = r * scale[s] rounded to complex64
__global__ __launch_bounds__(kThreads) void k_to32(int n, const double2* __restrict__ r, const double* __restrict__ scale,
                                                  float2* __restrict__ r32) {
    const int s = blockIdx.y;
    const double f = scale[s];
    for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += kVecBlocks * kThreads) {
        const size_t k = (size_t)s * n + i;
        const double2 z = r[k];
        r32[k] = make_float2((float)(z.x * f), (float)(z.y * f));
    }
}

// x += d32 * unscale[s] in FP64
__global__ __launch_bounds__(kThreads) void k_accum(int n, const float2* __restrict__ d, const double* __restrict__ unscale,
                                                   double2* __restrict__ x) {
    const int s = blockIdx.y;
    const double f = unscale[s];
    if (f == 0.0) return;
    for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += kVecBlocks * kThreads) {
        const size_t k = (size_t)s * n + i;
        const float2 z = d[k];
        double2 xv = x[k];
        xv.x += (double)z.x * f;
        xv.y += (double)z.y * f;
        x[k] = xv;
    }
}

__global__ __launch_bounds__(kThreads) void k_zero64(size_t count, double2* __restrict__ x) {
    for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < count; i += (size_t)gridDim.x * kThreads)
        x[i] = make_double2(0, 0);
}

struct RefineWork {
    double2* r;     // [S, n] FP64 residual
    double2* y;     // [S, n] FP64 matvec output
    float2* r32;    // [S, n]
    float2* d32;    // [S, n]
    double2* part;  // [S, kVecBlocks]
    double* norms;  // [S]
    double* scale;  // [S]
    double* unscale;  // [S]
    void* inner;    // complex64 BiCGSTAB workspace
};

size_t refine_bytes(int n, int S) {
    const size_t v64 = align256(sizeof(double2) * (size_t)n * S), v32 = align256(sizeof(float2) * (size_t)n * S);
    return 2 * v64 + 2 * v32 + align256(sizeof(double2) * (size_t)S * kVecBlocks) + 3 * align256(sizeof(double) * S) +
           work_bytes<float>(n, S);
}

RefineWork carve_refine(void* ws, int n, int S) {
    RefineWork w;
    char* p = (char*)ws;
    const size_t v64 = align256(sizeof(double2) * (size_t)n * S), v32 = align256(sizeof(float2) * (size_t)n * S);
    w.r = (double2*)p; p += v64;
    w.y = (double2*)p; p += v64;
    w.r32 = (float2*)p; p += v32;
    w.d32 = (float2*)p; p += v32;
    w.part = (double2*)p; p += align256(sizeof(double2) * (size_t)S * kVecBlocks);
    w.norms = (double*)p; p += align256(sizeof(double) * S);
    w.scale = (double*)p; p += align256(sizeof(double) * S);
    w.unscale = (double*)p; p += align256(sizeof(double) * S);
    w.inner = p;
    return w;
}

// FP64 residual operator: y = (I + T) x, from a stored c128 matrix.
struct ResidualOp {
    const void* A64;  // [n, n] c128, add_identity applies
    int add_identity;
};

int refine_impl(int n, int S, const void* A32, int add_identity32, ResidualOp op, const double2* b, double2* x,
                double tol, double inner_tol, int maxiter, int max_refine, void* ws, int* outer_out, int* inner_out,
                double* resid_out, int* status_out, hipStream_t st) {
    RefineWork w = carve_refine(ws, n, S);
    const dim3 grid(kVecBlocks, S), block(kThreads);
    std::vector<double> bn2(S), rn2(S), prev(S, INFINITY), sc(S), usc(S);
    std::vector<int> inner_total(S, 0), status(S, MPA_NOT_CONVERGED), outer(S, 0);
    std::vector<int> it(S), ist(S);
    std::vector<double> ires(S);

    hipLaunchKernelGGL(k_zero64, dim3(1024), block, 0, st, (size_t)n * S, x);
    hipLaunchKernelGGL(k_norm2, grid, block, 0, st, n, b, w.part);
    hipLaunchKernelGGL(k_norm_final, dim3(S), block, 0, st, w.part, w.norms);
    MPA_LAUNCH_CHECK();
    MPA_HIP_CHECK(hipMemcpyAsync(bn2.data(), w.norms, sizeof(double) * S, hipMemcpyDeviceToHost, st));
    MPA_HIP_CHECK(hipMemcpyAsync(w.r, b, sizeof(double2) * (size_t)n * S, hipMemcpyDeviceToDevice, st));
    MPA_HIP_CHECK(hipStreamSynchronize(st));
    rn2 = bn2;
    for (int outer_it = 0;; outer_it++) {
        bool all_done = true;
        for (int s = 0; s < S; s++) {
            const bool finite = std::isfinite(rn2[s]);
            if (!finite) status[s] = MPA_NONFINITE;
            else if (rn2[s] <= tol * tol * bn2[s]) status[s] = MPA_OK;
            else if (outer_it > 0 && rn2[s] > 0.25 * prev[s]) status[s] = MPA_NOT_CONVERGED;  // stagnation
            else status[s] = -1;  // keep refining
            const bool active = status[s] == -1 && outer_it < max_refine;
            if (status[s] == -1 && !active) status[s] = MPA_NOT_CONVERGED;
            sc[s] = active ? 1.0 / std::sqrt(rn2[s]) : 0.0;
            usc[s] = active ? std::sqrt(rn2[s]) : 0.0;
            all_done &= !active;
        }
        if (all_done) break;
        MPA_HIP_CHECK(hipMemcpyAsync(w.scale, sc.data(), sizeof(double) * S, hipMemcpyHostToDevice, st));
        MPA_HIP_CHECK(hipMemcpyAsync(w.unscale, usc.data(), sizeof(double) * S, hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_to32, grid, block, 0, st, n, w.r, w.scale, w.r32);
        MPA_LAUNCH_CHECK();
        int rc = bicgstab_impl<float>(n, S, A32, add_identity32, w.r32, w.d32, inner_tol, maxiter, w.inner, it.data(),
                                      ires.data(), ist.data(), 1, st);
        if (rc != MPA_OK && rc != MPA_NOT_CONVERGED && rc != MPA_NONFINITE) return rc;
        for (int s = 0; s < S; s++) {
            if (usc[s] != 0.0) {
                inner_total[s] += it[s];
                outer[s] = outer_it + 1;
            }
        }
        hipLaunchKernelGGL(k_accum, grid, block, 0, st, n, w.d32, w.unscale, x);
        MPA_LAUNCH_CHECK();
        rc = matvec<double>(n, S, op.A64, op.add_identity, x, w.y, st);
        if (rc) return rc;
        hipLaunchKernelGGL(k_residual64, grid, block, 0, st, n, b, w.y, w.r, w.part);
        hipLaunchKernelGGL(k_norm_final, dim3(S), block, 0, st, w.part, w.norms);
        MPA_LAUNCH_CHECK();
        prev = rn2;
        MPA_HIP_CHECK(hipMemcpyAsync(rn2.data(), w.norms, sizeof(double) * S, hipMemcpyDeviceToHost, st));
        MPA_HIP_CHECK(hipStreamSynchronize(st));
    }
    int worst = MPA_OK;
    for (int s = 0; s < S; s++) {
        if (outer_out) outer_out[s] = outer[s];
        if (inner_out) inner_out[s] = inner_total[s];
        if (resid_out) resid_out[s] = bn2[s] > 0 ? std::sqrt(rn2[s] / bn2[s]) : 0.0;
        if (status_out) status_out[s] = status[s];
        if (status[s] != MPA_OK && worst == MPA_OK) worst = status[s];
    }
    return worst;
}

}  // namespace

extern "C" {

int mpa_matvec(int n, int nrhs, int dtype, const void* A, int add_identity, const void* x, void* y, void* stream) {
    if (n < 1 || nrhs < 1 || !A || !x || !y) {
        mpa_set_error("matvec: bad argument");
        return MPA_ERR_ARG;
    }
    hipStream_t st = (hipStream_t)stream;
    if (dtype == MPA_C128) return matvec<double>(n, nrhs, A, add_identity, x, y, st);
    if (dtype == MPA_C64) return matvec<float>(n, nrhs, A, add_identity, x, y, st);
    mpa_set_error("matvec: unknown dtype %d", dtype);
    return MPA_ERR_ARG;
}

size_t mpa_bicgstab_workspace_bytes(int n, int nrhs, int dtype) {
    return dtype == MPA_C64 ? work_bytes<float>(n, nrhs) : work_bytes<double>(n, nrhs);
}

int mpa_bicgstab(int n, int nrhs, int dtype, const void* A, int add_identity, const void* b, void* x, double tol,
                 int maxiter, void* workspace, int* iters_out, double* resid_out, int* status_out, int sync_every,
                 void* stream) {
    if (n < 1 || nrhs < 1 || nrhs > 65535 || !A || !b || !x || !workspace || !(tol >= 0) || maxiter < 0) {
        mpa_set_error("bicgstab: bad argument");
        return MPA_ERR_ARG;
    }
    hipStream_t st = (hipStream_t)stream;
    if (dtype == MPA_C128)
        return bicgstab_impl<double>(n, nrhs, A, add_identity, b, x, tol, maxiter, workspace, iters_out, resid_out,
                                     status_out, sync_every, st);
    if (dtype == MPA_C64)
        return bicgstab_impl<float>(n, nrhs, A, add_identity, b, x, tol, maxiter, workspace, iters_out, resid_out,
                                    status_out, sync_every, st);
    mpa_set_error("bicgstab: unknown dtype %d", dtype);
    return MPA_ERR_ARG;
}

}  // extern "C"

extern "C" {

size_t mpa_refine_workspace_bytes(int n, int nrhs) { return refine_bytes(n, nrhs); }

int mpa_solve_refined(int n, int nrhs, const void* A32, int add_identity32, const void* A64, int add_identity64,
                      const void* b, void* x, double tol, double inner_tol, int maxiter, int max_refine,
                      void* workspace, int* outer_out, int* inner_out, double* resid_out, int* status_out,
                      void* stream) {
    if (n < 1 || nrhs < 1 || nrhs > 65535 || !A32 || !A64 || !b || !x || !workspace || !(tol > 0) ||
        !(inner_tol > 0) || maxiter < 1 || max_refine < 1) {
        mpa_set_error("solve_refined: bad argument");
        return MPA_ERR_ARG;
    }
    ResidualOp op{A64, add_identity64};
    return refine_impl(n, nrhs, A32, add_identity32, op, (const double2*)b, (double2*)x, tol, inner_tol, maxiter,
                       max_refine, workspace, outer_out, inner_out, resid_out, status_out, (hipStream_t)stream);
}

}  // extern "C"
