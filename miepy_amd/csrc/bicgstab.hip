// Batched BiCGSTAB over many independent systems (I + T_c) x_c = b_c.
//
// Replaces the solve of cpp/src/interactions.cpp:347-368 (Eigen::BiCGSTAB,
// identity preconditioner, row-block GEMV :80-91) for any number of systems
// at once.  The iteration is Eigen's (IterativeLinearSolvers/BiCGSTAB.h):
// zero initial guess, stop at ||r||^2 <= tol^2 ||b||^2 or after maxiter
// iterations, restart when |rho| < eps^2 ||r0||^2.  The reference hard-codes
// tol 1e-5 and maxiter 1000 (:362-363); here both are per-system inputs.
// Two per-system options cover the reference's other BiCGSTAB entry points:
// a block-diagonal right preconditioner (y = M^-1 p, z = M^-1 s, x += alpha y
// + w z; solve_linear_system_preconditioned, :291-345, and
// bicgstab_preconditioned_profiled, :230-288), and MPA_SOLVE_REF_LOOP, the
// reference's hand-written loop (bicgstab / bicgstab_profiled, :98-209):
// x0 = b, stop at ||r|| < tol absolute or after maxiter + 1 iterations, no
// restart.
//
// Every system keeps its matrix, Krylov vectors and per-RHS scalars in HBM.
// One launch of each kernel covers all systems of the batch through a work
// list (system, RHS, workgroup), so a batch of 256 small clusters costs the
// launches of one solve, and a system that converged turns every kernel that
// touches it into a no-op: systems of uneven size and iteration count share
// the launches, and new systems can join between iterations (the serving
// loop's continuous batching).  Dot products are two-stage and deterministic.
// A single solve is a batch of one.

#include <cmath>
#include <cstring>
#include <vector>

#include "gemv.hpp"
#include "krylov.hpp"
#include "matfree.hpp"

using namespace mpa;

namespace {

// ---- vector kernels (one workgroup per (system, RHS, group)) -------------------

#define VEC_PROLOGUE                                     \
    const int3 wk = work[blockIdx.x];                    \
    const SysDev& d = sys[wk.x];                         \
    const int s = wk.y, g = wk.z;                        \
    const int n = d.n;                                   \
    const size_t base = (size_t)s * n;                   \
    const int stride = d.G * kThreads;

template <typename T>
__global__ __launch_bounds__(kThreads) void k_init(const SysDev* __restrict__ sys, const int3* __restrict__ work) {
    using C = cplx_t<T>;
    VEC_PROLOGUE
    const C* b = (const C*)d.b;
    C *x = (C*)d.x, *r = (C*)d.r, *r0 = (C*)d.r0, *p = (C*)d.p, *v = (C*)d.v;
    C *y = (C*)d.y, *zz = (C*)d.z;  // the preconditioned operands (alias p, s without a preconditioner)
    const bool x0_b = (d.flags & MPA_SOLVE_REF_LOOP) != 0;  // r = b - A b is formed by the restart kernels
    double2 acc[1] = {make_double2(0, 0)};
    for (int i = g * kThreads + threadIdx.x; i < n; i += stride) {
        const size_t k = base + i;
        const C bv = b[k];
        r[k] = bv;
        r0[k] = bv;
        const C z = cmake<T>(0, 0);
        p[k] = z;
        v[k] = z;
        if (d.minv) {  // defined before any product reads them (a product may run while a restart is pending)
            y[k] = z;
            zz[k] = z;
        }
        x[k] = x0_b ? bv : z;
        const double2 bd = to_d(bv);
        acc[0].x += bd.x * bd.x + bd.y * bd.y;
    }
    block_sum<1>(acc);
    if (threadIdx.x == 0) d.part[((size_t)s * d.G + g) * kSlots + 3] = acc[0];
}

__global__ __launch_bounds__(kThreads) void k_init_scalars(const SysDev* __restrict__ sys, const int2* __restrict__ cw) {
    const int2 c2 = cw[blockIdx.x];
    const SysDev& d = sys[c2.x];
    const int s = c2.y;
    const double2 bb = reduce_partials(d.part, d.G, s, 3);
    if (threadIdx.x == 0) {
        Scal& c = d.scal[s];
        const bool ref = (d.flags & MPA_SOLVE_REF_LOOP) != 0;
        c.rho = c.alpha = c.w = make_double2(1, 0);
        c.alpha_next = c.w_next = make_double2(1, 0);
        c.rhs_sq = bb.x;
        c.r0_sq = bb.x;
        c.rnorm2 = bb.x;
        c.tol2 = ref ? d.tol * d.tol : d.tol * d.tol * bb.x;
        c.it = 0;
        c.restarts = 0;
        c.maxiter = ref ? d.maxiter + 1 : d.maxiter;
        c.status = MPA_OK;
        c.restart_pending = 0;
        c.init_pending = 0;
        c.rho_next = make_double2(bb.x, 0);  // <r0, r> = ||b||^2 with x0 = 0
        if (ref) {
            // x0 = b: the restart kernels form r = b - A b, r0 = r and make the
            // loop's initial test ||r|| < tol (b = 0 gives x = 0, r = 0 at once)
            c.done = bb.x == 0.0 ? 1 : 0;
            c.restart_pending = c.init_pending = c.done ? 0 : 1;
        } else {
            c.done = (bb.x == 0.0 || !(bb.x > c.tol2) || d.maxiter <= 0) ? 1 : 0;
        }
        if (bb.x != bb.x) {
            c.done = 1;
            c.restart_pending = c.init_pending = 0;
            c.status = MPA_NONFINITE;
        }
    }
}

// p = r + beta (p - w v), beta = (rho_new / rho)(alpha / w)
template <typename T>
__global__ __launch_bounds__(kThreads) void k_update_p(const SysDev* __restrict__ sys, const int3* __restrict__ work) {
    using C = cplx_t<T>;
    VEC_PROLOGUE
    const Scal& c = d.scal[s];
    if (c.done || c.restart_pending) return;
    const double2 beta = cmul(cdiv(c.rho_next, c.rho), cdiv(c.alpha, c.w));
    const C bt = from_d<T>(beta), wt = from_d<T>(c.w);
    C *p = (C*)d.p, *r = (C*)d.r, *v = (C*)d.v;
    for (int i = g * kThreads + threadIdx.x; i < n; i += stride) {
        const size_t k = base + i;
        p[k] = cadd(r[k], cmul(bt, csub(p[k], cmul(wt, v[k]))));
    }
}

// Right preconditioner: y = M^-1 p (which = 0) or z = M^-1 s (which = 1),
// M^-1 the block-diagonal [n, bs] stack of the system's block inverses
// (apply_block_preconditioner, cpp/src/interactions.cpp:211-228).  One
// thread per output row; the bs entries of the row's input block are shared
// by the block's rows through the cache.  FP64 accumulation.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_precond(const SysDev* __restrict__ sys, const int3* __restrict__ work,
                                                      int which) {
    using C = cplx_t<T>;
    VEC_PROLOGUE
    if (!d.minv) return;
    const Scal& c = d.scal[s];
    if (c.done || c.restart_pending) return;
    const C* src = (const C*)(which ? d.s : d.p) + base;
    C* dst = (C*)(which ? d.z : d.y) + base;
    const C* M = (const C*)d.minv;
    const int bs = d.bs;
    for (int i = g * kThreads + threadIdx.x; i < n; i += stride) {
        const C* row = M + (size_t)i * bs;
        const C* in = src + (i - i % bs);
        double2 acc = make_double2(0, 0);
        for (int k = 0; k < bs; k++) acc = cadd(acc, cmul(to_d(row[k]), to_d(in[k])));
        dst[i] = from_d<T>(acc);
    }
}

// partial <r0, v>
template <typename T>
__global__ __launch_bounds__(kThreads) void k_dot_r0v(const SysDev* __restrict__ sys, const int3* __restrict__ work) {
    using C = cplx_t<T>;
    VEC_PROLOGUE
    if (d.scal[s].done || d.scal[s].restart_pending) return;
    const C *r0 = (const C*)d.r0, *v = (const C*)d.v;
    double2 acc[1] = {make_double2(0, 0)};
    for (int i = g * kThreads + threadIdx.x; i < n; i += stride) acc[0] = cadd(acc[0], cmulc(to_d(r0[base + i]), to_d(v[base + i])));
    block_sum<1>(acc);
    if (threadIdx.x == 0) d.part[((size_t)s * d.G + g) * kSlots + 0] = acc[0];
}

// alpha = rho / <r0, v>; s = r - alpha v
template <typename T>
__global__ __launch_bounds__(kThreads) void k_update_s(const SysDev* __restrict__ sys, const int3* __restrict__ work) {
    using C = cplx_t<T>;
    VEC_PROLOGUE
    if (d.scal[s].done || d.scal[s].restart_pending) return;
    const double2 dv = reduce_partials(d.part, d.G, s, 0);
    const double2 alpha = cdiv(d.scal[s].rho_next, dv);
    const C at = from_d<T>(alpha);
    C *sv = (C*)d.s, *r = (C*)d.r, *v = (C*)d.v;
    for (int i = g * kThreads + threadIdx.x; i < n; i += stride) sv[base + i] = csub(r[base + i], cmul(at, v[base + i]));
    if (g == 0 && threadIdx.x == 0) d.scal[s].alpha_next = alpha;
}

// partials <t, s>, <t, t>
template <typename T>
__global__ __launch_bounds__(kThreads) void k_dot_ts(const SysDev* __restrict__ sys, const int3* __restrict__ work) {
    using C = cplx_t<T>;
    VEC_PROLOGUE
    if (d.scal[s].done || d.scal[s].restart_pending) return;
    const C *tv = (const C*)d.t, *sv = (const C*)d.s;
    double2 acc[2] = {make_double2(0, 0), make_double2(0, 0)};
    for (int i = g * kThreads + threadIdx.x; i < n; i += stride) {
        const double2 t = to_d(tv[base + i]);
        acc[0] = cadd(acc[0], cmulc(t, to_d(sv[base + i])));
        acc[1].x += t.x * t.x + t.y * t.y;
    }
    block_sum<2>(acc);
    if (threadIdx.x == 0) {
        d.part[((size_t)s * d.G + g) * kSlots + 1] = acc[0];
        d.part[((size_t)s * d.G + g) * kSlots + 2] = acc[1];
    }
}

// w = <t,s>/<t,t>; x += alpha p + w s; r = s - w t; partials ||r||^2, <r0, r>
template <typename T>
__global__ __launch_bounds__(kThreads) void k_update_xr(const SysDev* __restrict__ sys, const int3* __restrict__ work) {
    using C = cplx_t<T>;
    VEC_PROLOGUE
    if (d.scal[s].done || d.scal[s].restart_pending) return;
    const double2 ts = reduce_partials(d.part, d.G, s, 1);
    const double2 tt = reduce_partials(d.part, d.G, s, 2);
    const double2 om = tt.x > 0 ? cscale(ts, 1.0 / tt.x) : make_double2(0, 0);
    const C at = from_d<T>(d.scal[s].alpha_next), wt = from_d<T>(om);
    // y, z alias p, s without a preconditioner
    C *x = (C*)d.x, *y = (C*)d.y, *z = (C*)d.z, *sv = (C*)d.s, *tv = (C*)d.t, *r = (C*)d.r, *r0 = (C*)d.r0;
    const bool pc = d.minv != nullptr;
    double2 acc[2] = {make_double2(0, 0), make_double2(0, 0)};
    for (int i = g * kThreads + threadIdx.x; i < n; i += stride) {
        const size_t k = base + i;
        const C s_ = sv[k];
        x[k] = cadd(x[k], cadd(cmul(at, y[k]), cmul(wt, pc ? z[k] : s_)));
        const C rv = csub(s_, cmul(wt, tv[k]));
        r[k] = rv;
        const double2 rd = to_d(rv);
        acc[0].x += rd.x * rd.x + rd.y * rd.y;
        acc[1] = cadd(acc[1], cmulc(to_d(r0[k]), rd));
    }
    block_sum<2>(acc);
    if (threadIdx.x == 0) {
        d.part[((size_t)s * d.G + g) * kSlots + 3] = acc[0];
        d.part[((size_t)s * d.G + g) * kSlots + 4] = acc[1];
    }
    if (g == 0 && threadIdx.x == 0) d.scal[s].w_next = om;
}

// commit the iteration's scalars; convergence, maxiter and restart checks
__global__ __launch_bounds__(kThreads) void k_commit(const SysDev* __restrict__ sys, const int2* __restrict__ cw) {
    const int2 c2 = cw[blockIdx.x];
    const SysDev& d = sys[c2.x];
    const int s = c2.y;
    Scal& c = d.scal[s];
    if (c.done || c.restart_pending) return;
    const double2 rr = reduce_partials(d.part, d.G, s, 3);
    const double2 r0r = reduce_partials(d.part, d.G, s, 4);
    if (threadIdx.x == 0) {
        c.rho = c.rho_next;
        c.alpha = c.alpha_next;
        c.w = c.w_next;
        c.rnorm2 = rr.x;
        c.it += 1;
        c.rho_next = r0r;
        const double eps = 2.220446049250313e-16;  // NumTraits<double>::epsilon()
        const bool ref = (d.flags & MPA_SOLVE_REF_LOOP) != 0;
        if (!(rr.x == rr.x) || !(r0r.x == r0r.x) || !(r0r.y == r0r.y)) {
            c.done = 1;
            c.status = MPA_NONFINITE;
        } else if (ref ? rr.x < c.tol2 : !(rr.x > c.tol2)) {  // ||r|| < tol (reference loop), Eigen's <=
            c.done = 1;
        } else if (c.it >= c.maxiter) {
            c.done = 1;
            c.status = MPA_NOT_CONVERGED;
        } else if (!ref && sqrt(r0r.x * r0r.x + r0r.y * r0r.y) < eps * eps * c.r0_sq) {
            c.restart_pending = 1;
        }
    }
}

// restart: r = b - t (t = A x computed just before), r0 = r, partial ||r||^2
template <typename T>
__global__ __launch_bounds__(kThreads) void k_restart(const SysDev* __restrict__ sys, const int3* __restrict__ work) {
    using C = cplx_t<T>;
    VEC_PROLOGUE
    if (!d.scal[s].restart_pending) return;
    const C *b = (const C*)d.b, *tv = (const C*)d.t;
    C *r = (C*)d.r, *r0 = (C*)d.r0;
    double2 acc[1] = {make_double2(0, 0)};
    for (int i = g * kThreads + threadIdx.x; i < n; i += stride) {
        const size_t k = base + i;
        const C rv = csub(b[k], tv[k]);
        r[k] = rv;
        r0[k] = rv;
        const double2 rd = to_d(rv);
        acc[0].x += rd.x * rd.x + rd.y * rd.y;
    }
    block_sum<1>(acc);
    if (threadIdx.x == 0) d.part[((size_t)s * d.G + g) * kSlots + 3] = acc[0];
}

__global__ __launch_bounds__(kThreads) void k_restart_commit(const SysDev* __restrict__ sys, const int2* __restrict__ cw) {
    const int2 c2 = cw[blockIdx.x];
    const SysDev& d = sys[c2.x];
    const int s = c2.y;
    if (!d.scal[s].restart_pending) return;
    const double2 rr = reduce_partials(d.part, d.G, s, 3);
    if (threadIdx.x == 0) {
        Scal& c = d.scal[s];
        // Eigen: rho = r0_sqnorm = ||r||^2 replaces this iteration's rho; the
        // committed rho stays as rho_old for beta = (rho/rho_old)(alpha/w).
        c.r0_sq = rr.x;
        c.rho_next = make_double2(rr.x, 0);
        c.restart_pending = 0;
        if (c.init_pending) {  // x0 = b: the reference loop's initial test on ||b - A b||
            c.init_pending = 0;
            c.rnorm2 = rr.x;
            if (!(rr.x == rr.x)) {
                c.done = 1;
                c.status = MPA_NONFINITE;
            } else if (rr.x < c.tol2) {
                c.done = 1;
            }
        } else if (c.restarts++ == 0) {
            c.it = 0;
        }
    }
}

// copy every RHS's scalars into one contiguous array for a single readback
__global__ __launch_bounds__(64) void k_gather_scalars(const SysDev* __restrict__ sys, const int2* __restrict__ cw,
                                                        int count, Scal* __restrict__ out) {
    const int k = blockIdx.x * 64 + threadIdx.x;
    if (k < count) out[k] = sys[cw[k].x].scal[cw[k].y];
}

// ---- host side ---------------------------------------------------------------------

size_t vec_bytes(int n, int S, int esize) { return align256((size_t)esize * 2 * n * S); }

size_t state_bytes(int n, int S, int dtype, bool precond = false) {
    const int es = dtype == MPA_C128 ? 8 : 4;
    return (precond ? 8 : 6) * vec_bytes(n, S, es) + align256(sizeof(Scal) * S) +
           align256(sizeof(double2) * (size_t)S * groups_for(n) * kSlots);
}

// Scratch for descriptors and work lists of a batch (device).
struct Plan {
    std::vector<SysDev> sys;
    std::vector<int3> vec_work;
    std::vector<int2> scal_work;
    std::vector<MvDev> mv_ps, mv_st, mv_xt;
    std::vector<int2> mv_work[kTileClasses];  // (job, row block) per tile class; same for every job set
    std::vector<MfJob> mf_ps, mf_st, mf_xt;   // matrix-free systems (launched by value)
    bool any_precond = false;
    // device copies
    SysDev* d_sys = nullptr;
    int3* d_vec = nullptr;
    int2* d_scal = nullptr;
    MvDev *d_mv_ps = nullptr, *d_mv_st = nullptr, *d_mv_xt = nullptr;
    int2* d_mv_work[kTileClasses] = {};
    int n_mv_work[kTileClasses] = {};
    Scal* d_scal_out = nullptr;  // gathered scalars, [sum(nrhs)]
    std::vector<char> img;       // host image of the descriptor region, uploaded in one copy
};

size_t plan_bytes(int nsys, const mpa_system* s, int dtype) {
    size_t vec = 0, sc = 0, mvw = 0;
    for (int i = 0; i < nsys; i++) {
        vec += (size_t)s[i].nrhs * groups_for(s[i].n);
        sc += s[i].nrhs;
        mvw += gemv_row_blocks(s[i].n, s[i].nrhs, dtype == MPA_C128 ? 8 : 4);
    }
    return align256(sizeof(SysDev) * nsys) + align256(sizeof(int3) * vec) + align256(sizeof(int2) * sc) +
           3 * align256(sizeof(MvDev) * nsys) + kTileClasses * 256 + align256(sizeof(int2) * mvw) +
           align256(sizeof(Scal) * sc);
}

int build_plan(Plan& P, int nsys, const mpa_system* S_, int dtype, void* scratch, hipStream_t st) {
    P.sys.resize(nsys);
    P.vec_work.clear();
    P.scal_work.clear();
    P.mv_ps.resize(nsys);
    P.mv_st.resize(nsys);
    P.mv_xt.resize(nsys);
    for (auto& w : P.mv_work) w.clear();
    P.mf_ps.clear();
    P.mf_st.clear();
    P.mf_xt.clear();
    P.any_precond = false;
    const int es = dtype == MPA_C128 ? 8 : 4;
    for (int i = 0; i < nsys; i++) {
        const mpa_system& q = S_[i];
        SysDev& d = P.sys[i];
        const int n = q.n, S = q.nrhs;
        char* p = (char*)q.state;
        const size_t vb = vec_bytes(n, S, es);
        d.A = q.A;
        d.b = q.b;
        d.x = q.x;
        d.r = p; p += vb;
        d.r0 = p; p += vb;
        d.p = p; p += vb;
        d.v = p; p += vb;
        d.s = p; p += vb;
        d.t = p; p += vb;
        d.scal = (Scal*)p; p += align256(sizeof(Scal) * S);
        d.part = (double2*)p; p += align256(sizeof(double2) * (size_t)S * groups_for(n) * kSlots);
        d.minv = q.minv;
        d.bs = q.minv ? q.block_size : 0;
        if (q.minv) {  // after the unpreconditioned layout: state_bytes(n, S, dtype, true)
            d.y = p; p += vb;
            d.z = p; p += vb;
            P.any_precond = true;
        } else {
            d.y = d.p;
            d.z = d.s;
        }
        d.tol = q.tol;
        d.n = n;
        d.S = S;
        d.G = groups_for(n);
        d.add_identity = q.add_identity;
        d.maxiter = q.maxiter;
        d.flags = q.flags;
        d.pad = 0;
        for (int s = 0; s < S; s++) {
            P.scal_work.push_back(make_int2(i, s));
            for (int g = 0; g < d.G; g++) P.vec_work.push_back(make_int3(i, s, g));
        }
        P.mv_ps[i] = MvDev{q.A, d.y, d.v, n, S, q.add_identity, i};
        P.mv_st[i] = MvDev{q.A, d.z, d.t, n, S, q.add_identity, i};
        P.mv_xt[i] = MvDev{q.A, q.x, d.t, n, S, q.add_identity, -1};
        if (q.op) {
            const mpa_cluster_operator& o = *q.op;
            if (!mpa_get_table(o.lmax)) return MPA_ERR_ARG;  // built here, never inside a graph capture
            const MfJob base{o.N, o.lmax, S, q.add_identity ? 1 : 0, o.positions, (const double2*)o.mie, o.k,
                             nullptr, nullptr, (double2*)o.workspace, i, 0, (const double2*)o.tmatrix, nullptr,
                             dtype == MPA_C64 ? 1 : 0, 0};
            MfJob j = base;
            j.x = (const double2*)d.y;
            j.y = (double2*)d.v;
            P.mf_ps.push_back(j);
            j.x = (const double2*)d.z;
            j.y = (double2*)d.t;
            P.mf_st.push_back(j);
            j.x = (const double2*)q.x;
            j.y = (double2*)d.t;
            j.skip = -1;
            P.mf_xt.push_back(j);
            continue;
        }
        const int rb = gemv_row_blocks(n, S, es);
        std::vector<int2>& mw = P.mv_work[gemv_class(S)];
        for (int b = 0; b < rb; b++) mw.push_back(make_int2(i, b));
    }
    for (auto& w : P.mv_work) order_work_longest_first(w, [&](int i) { return S_[i].n; });
    char* p = (char*)scratch;
    P.d_sys = (SysDev*)p; p += align256(sizeof(SysDev) * nsys);
    P.d_vec = (int3*)p; p += align256(sizeof(int3) * P.vec_work.size());
    P.d_scal = (int2*)p; p += align256(sizeof(int2) * P.scal_work.size());
    P.d_mv_ps = (MvDev*)p; p += align256(sizeof(MvDev) * nsys);
    P.d_mv_st = (MvDev*)p; p += align256(sizeof(MvDev) * nsys);
    P.d_mv_xt = (MvDev*)p; p += align256(sizeof(MvDev) * nsys);
    for (int c = 0; c < kTileClasses; c++) {
        P.d_mv_work[c] = (int2*)p;
        P.n_mv_work[c] = (int)P.mv_work[c].size();
        p += align256(sizeof(int2) * P.mv_work[c].size());
    }
    P.d_scal_out = (Scal*)p;
    // every descriptor array and work list in one host image with the device
    // layout, sent in one copy (one pageable transfer instead of one per array)
    const size_t nbytes = (size_t)(p - (char*)scratch);
    P.img.assign(nbytes, 0);
    auto put = [&](const void* dst, const void* src, size_t bytes) {
        if (bytes) std::memcpy(P.img.data() + ((const char*)dst - (const char*)scratch), src, bytes);
    };
    put(P.d_sys, P.sys.data(), sizeof(SysDev) * nsys);
    put(P.d_vec, P.vec_work.data(), sizeof(int3) * P.vec_work.size());
    put(P.d_scal, P.scal_work.data(), sizeof(int2) * P.scal_work.size());
    put(P.d_mv_ps, P.mv_ps.data(), sizeof(MvDev) * nsys);
    put(P.d_mv_st, P.mv_st.data(), sizeof(MvDev) * nsys);
    put(P.d_mv_xt, P.mv_xt.data(), sizeof(MvDev) * nsys);
    for (int c = 0; c < kTileClasses; c++) put(P.d_mv_work[c], P.mv_work[c].data(), sizeof(int2) * P.n_mv_work[c]);
    // the host image is staged by the runtime before the call returns
    MPA_HIP_CHECK(hipMemcpyAsync(scratch, P.img.data(), nbytes, hipMemcpyHostToDevice, st));
    return MPA_OK;
}

// All matvec jobs of one phase: the batched GEMV over stored matrices and
// the matrix-free systems.  With both kinds present (and MPA_OPT_OVERLAP),
// the GEMV, bound by HBM, runs on the device's side stream beside the
// matrix-free products, bound by the matrix cores: a fork event recorded on
// `st` starts it, a join event recorded after it is waited on by `st`, so
// everything after the phase sees both (in a graph capture the fork and
// join make the side stream's launch a parallel branch of the graph).
template <typename T>
int matvec_phase(const Plan& P, const MvDev* d_jobs, const std::vector<MfJob>& mf, hipStream_t st,
                 bool overlap = true) {
    bool any = false;
    for (int c = 0; c < kTileClasses; c++) any |= P.n_mv_work[c] > 0;
    if (any && !mf.empty() && overlap && mpa_option(MPA_OPT_OVERLAP) != 0 && !mpa_profile_active()) {
        SideLease lease;
        const SideStream* side = lease.get();
        if (side) {
            MPA_HIP_CHECK(hipEventRecord(side->fork, st));
            MPA_HIP_CHECK(hipStreamWaitEvent(side->stream, side->fork, 0));
            const int rc1 = gemv_launch(sizeof(T), d_jobs, P.d_sys, P.d_mv_work, P.n_mv_work, side->stream);
            const int rc2 = matfree_launch_many(mf, P.d_sys, st);
            // join whatever was launched, also after an error, so no branch is left open
            MPA_HIP_CHECK(hipEventRecord(side->join, side->stream));
            MPA_HIP_CHECK(hipStreamWaitEvent(st, side->join, 0));
            return rc1 ? rc1 : rc2;
        }
    }
    if (any) {
        const int rc = gemv_launch(sizeof(T), d_jobs, P.d_sys, P.d_mv_work, P.n_mv_work, st);
        if (rc) return rc;
    }
    return mf.empty() ? MPA_OK : matfree_launch_many(mf, P.d_sys, st);
}

template <typename T>
int launch_start(const Plan& P, hipStream_t st) {
    const int nv = (int)P.vec_work.size(), ns = (int)P.scal_work.size();
    hipLaunchKernelGGL(k_init<T>, dim3(nv), dim3(kThreads), 0, st, P.d_sys, P.d_vec);
    hipLaunchKernelGGL(k_init_scalars, dim3(ns), dim3(kThreads), 0, st, P.d_sys, P.d_scal);
    MPA_LAUNCH_CHECK();
    return MPA_OK;
}

template <typename T>
int launch_iteration(const Plan& P, hipStream_t st) {
    const int nv = (int)P.vec_work.size(), ns = (int)P.scal_work.size();
    hipLaunchKernelGGL(k_update_p<T>, dim3(nv), dim3(kThreads), 0, st, P.d_sys, P.d_vec);
    if (P.any_precond) hipLaunchKernelGGL(k_precond<T>, dim3(nv), dim3(kThreads), 0, st, P.d_sys, P.d_vec, 0);
    int rc = matvec_phase<T>(P, P.d_mv_ps, P.mf_ps, st);
    if (rc) return rc;
    hipLaunchKernelGGL(k_dot_r0v<T>, dim3(nv), dim3(kThreads), 0, st, P.d_sys, P.d_vec);
    hipLaunchKernelGGL(k_update_s<T>, dim3(nv), dim3(kThreads), 0, st, P.d_sys, P.d_vec);
    if (P.any_precond) hipLaunchKernelGGL(k_precond<T>, dim3(nv), dim3(kThreads), 0, st, P.d_sys, P.d_vec, 1);
    rc = matvec_phase<T>(P, P.d_mv_st, P.mf_st, st);
    if (rc) return rc;
    hipLaunchKernelGGL(k_dot_ts<T>, dim3(nv), dim3(kThreads), 0, st, P.d_sys, P.d_vec);
    hipLaunchKernelGGL(k_update_xr<T>, dim3(nv), dim3(kThreads), 0, st, P.d_sys, P.d_vec);
    hipLaunchKernelGGL(k_commit, dim3(ns), dim3(kThreads), 0, st, P.d_sys, P.d_scal);
    MPA_LAUNCH_CHECK();
    return MPA_OK;
}

template <typename T>
int launch_restart(const Plan& P, hipStream_t st) {
    const int nv = (int)P.vec_work.size(), ns = (int)P.scal_work.size();
    int rc = matvec_phase<T>(P, P.d_mv_xt, P.mf_xt, st);
    if (rc) return rc;
    hipLaunchKernelGGL(k_restart<T>, dim3(nv), dim3(kThreads), 0, st, P.d_sys, P.d_vec);
    hipLaunchKernelGGL(k_restart_commit, dim3(ns), dim3(kThreads), 0, st, P.d_sys, P.d_scal);
    MPA_LAUNCH_CHECK();
    return MPA_OK;
}

// Read back the scalars of every RHS (synchronizes the stream): one gather
// kernel and one copy through pinned host memory, however many systems.
// noalloc: straight into the caller's pageable vector (no pinned staging, which may allocate).
int read_scalars(const Plan& P, std::vector<Scal>& out, hipStream_t st, bool noalloc = false) {
    const int total = (int)P.scal_work.size();
    out.resize(total);
    hipLaunchKernelGGL(k_gather_scalars, dim3((total + 63) / 64), dim3(64), 0, st, P.d_sys, P.d_scal, total,
                       P.d_scal_out);
    MPA_LAUNCH_CHECK();
    const size_t bytes = sizeof(Scal) * total;
    if (noalloc) {
        MPA_HIP_CHECK(hipMemcpyAsync(out.data(), P.d_scal_out, bytes, hipMemcpyDeviceToHost, st));
        MPA_HIP_CHECK(hipStreamSynchronize(st));
        return MPA_OK;
    }
    PinnedLease pin(bytes);
    void* dst = pin.ptr() ? pin.ptr() : (void*)out.data();
    MPA_HIP_CHECK(hipMemcpyAsync(dst, P.d_scal_out, bytes, hipMemcpyDeviceToHost, st));
    MPA_HIP_CHECK(hipStreamSynchronize(st));
    if (dst != out.data()) std::memcpy(out.data(), dst, bytes);
    return MPA_OK;
}

// k BiCGSTAB iterations of a plan recorded once as a HIP graph and replayed.
struct IterGraph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    IterGraph() = default;
    IterGraph(const IterGraph&) = delete;
    IterGraph& operator=(const IterGraph&) = delete;
    ~IterGraph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    }
};

template <typename T>
int launch_iteration(const Plan& P, hipStream_t st);

template <typename T>
int capture_iterations(const Plan& P, int k, IterGraph& g) {
    hipStream_t cap = mpa_capture_stream();
    if (!cap) {
        mpa_set_error("bicgstab: no capture stream");
        return MPA_ERR_HIP;
    }
    MPA_HIP_CHECK(hipStreamBeginCapture(cap, hipStreamCaptureModeRelaxed));
    int rc = MPA_OK;
    for (int i = 0; i < k && rc == MPA_OK; i++) rc = launch_iteration<T>(P, cap);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(cap, &graph);
    if (rc != MPA_OK) {
        if (e == hipSuccess && graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    MPA_HIP_CHECK(e);
    g.graph = graph;
    MPA_HIP_CHECK(hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0));
    return MPA_OK;
}

template <typename T>
// noalloc: eager launches only (no graph capture or instantiation) and no pinned staging.
int solve_batch(int nsys, const mpa_system* S_, void* scratch, int* iters, double* resid, int* status, int sync_every,
                hipStream_t st, bool noalloc = false) {
    Plan P;
    const int dtype = sizeof(T) == 8 ? MPA_C128 : MPA_C64;
    int rc = build_plan(P, nsys, S_, dtype, scratch, st);
    if (rc) return rc;
    rc = launch_start<T>(P, st);
    if (rc) return rc;
    int maxit = 0;
    for (int i = 0; i < nsys; i++) maxit = std::max(maxit, S_[i].maxiter);
    const int k_sync = sync_every > 0 ? sync_every : 1;
    std::vector<Scal> host;
    long launched = 0;
    IterGraph graph;
    const bool use_graph = !noalloc && mpa_option(MPA_OPT_GRAPHS) != 0 && !mpa_profile_active();
    if (use_graph) {
        rc = capture_iterations<T>(P, k_sync, graph);
        if (rc) return rc;
    }
    // Rounds of k_sync iterations between host checks.  The first check
    // comes after min(4, maxiter) iterations (every solve runs a few; a
    // system done earlier turns the rest into early-exit launches); from
    // then on, as many rounds as the slowest RHS's observed residual decay
    // says it still needs (70 % of the estimate, 1..8 rounds; the first
    // estimate measures the decay from ||b|| at iteration 0), so a solve of
    // a few dozen iterations reads its scalars back a handful of times.
    std::vector<double> prev_r2;
    std::vector<int> prev_it;
    auto launch_rounds = [&](int nr) -> int {
        for (int round = 0; round < nr; round++) {
            if (use_graph) {
                MPA_HIP_CHECK(hipGraphLaunch(graph.exec, st));
            } else {
                for (int k = 0; k < k_sync; k++) {
                    const int r2 = launch_iteration<T>(P, st);
                    if (r2) return r2;
                }
            }
        }
        launched += (long)nr * k_sync;
        return MPA_OK;
    };
    int rounds = std::max(1, std::min(4, maxit) / k_sync);
    bool from_start = maxit > 0;  // the decay so far is measured from iteration 0
    for (int i = 0; i < nsys; i++) from_start &= (S_[i].flags & MPA_SOLVE_REF_LOOP) == 0;  // x0 = b: restart first
    if (from_start && (rc = launch_rounds(rounds))) return rc;
    for (;;) {
        rc = read_scalars(P, host, st, noalloc);
        if (rc) return rc;
        bool all_done = true, any_restart = false;
        for (const Scal& c : host) {
            all_done &= c.done != 0;
            any_restart |= c.restart_pending != 0;
        }
        if (any_restart) {
            rc = launch_restart<T>(P, st);
            if (rc) return rc;
            prev_r2.clear();
            from_start = false;
            rounds = 1;
            continue;
        }
        if (all_done) break;
        if (launched > 4L * (maxit + 8)) {
            mpa_set_error("bicgstab: iteration budget exhausted");
            return MPA_NOT_CONVERGED;
        }
        if (from_start || prev_r2.size() == host.size()) {
            double need = 1.0;
            for (size_t k = 0; k < host.size(); k++) {
                const Scal& c = host[k];
                if (c.done) continue;
                const double p2 = from_start ? c.rhs_sq : prev_r2[k];
                const int di = c.it - (from_start ? 0 : prev_it[k]);
                if (di <= 0 || !(p2 > 0) || !(c.rnorm2 > 0) || !(c.tol2 > 0)) continue;
                const double rate = std::log(c.rnorm2 / p2) / di;  // per iteration
                if (!(rate < -1e-3)) continue;                    // no measurable decay: one round
                need = std::max(need, 0.7 * std::log(c.tol2 / c.rnorm2) / rate / k_sync);
            }
            rounds = (int)std::min(8.0, std::max(1.0, std::floor(need)));
        }
        from_start = false;
        prev_r2.resize(host.size());
        prev_it.resize(host.size());
        for (size_t k = 0; k < host.size(); k++) {
            prev_r2[k] = host[k].rnorm2;
            prev_it[k] = host[k].it;
        }
        if ((rc = launch_rounds(rounds))) return rc;
    }
    int worst = MPA_OK;
    for (size_t k = 0; k < host.size(); k++) {
        if (iters) iters[k] = host[k].it;
        if (resid) resid[k] = host[k].rhs_sq > 0 ? std::sqrt(host[k].rnorm2 / host[k].rhs_sq) : 0.0;
        if (status) status[k] = host[k].status;
        if (host[k].status != MPA_OK && worst == MPA_OK) worst = host[k].status;
    }
    return worst;
}

bool valid_systems(int nsys, const mpa_system* s, int dtype) {
    if (nsys < 1 || !s || (dtype != MPA_C128 && dtype != MPA_C64)) return false;
    for (int i = 0; i < nsys; i++) {
        if (s[i].n < 2 || (s[i].n & 1) || s[i].nrhs < 1 || !s[i].b || !s[i].x || !s[i].state || !(s[i].tol >= 0) ||
            s[i].maxiter < 0 || (s[i].flags & ~MPA_SOLVE_REF_LOOP) != 0)
            return false;
        if (s[i].minv && (s[i].block_size < 1 || s[i].n % s[i].block_size != 0)) return false;
        if (s[i].op) {
            const mpa_cluster_operator& o = *s[i].op;
            if (s[i].A || o.N < 1 || o.lmax < 1 || o.lmax > MPA_LMAX_MAX || !o.positions ||
                !o.mie == !o.tmatrix || !o.workspace || !(o.k > 0) ||
                (long long)2 * lmax_to_rmax(o.lmax) * o.N != s[i].n)
                return false;
            if (dtype == MPA_C64 && !matfree_c64_supported(o.N, o.lmax, o.tmatrix != nullptr)) return false;
        } else if (!s[i].A) {
            return false;
        }
    }
    return true;
}

}  // namespace

// ---- allocation-free BiCGSTAB ----------------------------------------------------------
//
// mpa_bicgstab_noalloc is mpa_bicgstab with its workspace (Krylov vectors,
// partial sums, per-RHS scalars, descriptors and work lists) in this static
// device array instead of a caller-provided buffer, for callers that cannot
// allocate device memory.  Every GPU the library runs on holds one copy of
// the array until the process ends.  Calls on one GPU share it and it keeps
// the last call's contents: the solve's start kernels initialize every
// vector and scalar it reads, and calls that overlap on one GPU (several host
// threads, or streams not ordered with each other) race on it.  Host code
// never passes `scratch` itself to a kernel or a copy: the array's name is a
// host-side stand-in, and its device address (per device) comes from
// hipGetSymbolAddress.
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

constexpr int kNoAllocMaxDevices = 64;
std::mutex g_noalloc_mu;
void* g_noalloc_ptr[kNoAllocMaxDevices] = {};

// The device address of `scratch` on the current device, looked up on the
// first call on that device (with it current) and kept.
int noalloc_scratch(void** out) {
    int dev = 0;
    MPA_HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kNoAllocMaxDevices) {
        mpa_set_error("bicgstab_noalloc: device %d is past the %d devices this library tracks", dev,
                      kNoAllocMaxDevices);
        return MPA_ERR_ARG;
    }
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    if (!g_noalloc_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            mpa_set_error("bicgstab_noalloc: hipGetSymbolAddress of the static scratch failed on device %d: %s", dev,
                          hipGetErrorString(e));
            return MPA_ERR_HIP;
        }
        g_noalloc_ptr[dev] = p;
    }
    *out = g_noalloc_ptr[dev];
    return MPA_OK;
}

}  // namespace

// ---- batched engine, exposed for the refinement driver and the serving loop ------

int mpa_internal_solve_batch(int nsys, const mpa_system* systems, int dtype, void* scratch, int* iters, double* resid,
                             int* status, int sync_every, hipStream_t st) {
    if (dtype == MPA_C128) return solve_batch<double>(nsys, systems, scratch, iters, resid, status, sync_every, st);
    return solve_batch<float>(nsys, systems, scratch, iters, resid, status, sync_every, st);
}

extern "C" {

size_t mpa_system_state_bytes(int n, int nrhs, int dtype) { return state_bytes(n, nrhs, dtype); }

size_t mpa_system_state_bytes_ex(int n, int nrhs, int dtype, int preconditioned) {
    return state_bytes(n, nrhs, dtype, preconditioned != 0);
}

size_t mpa_batch_scratch_bytes(int nsys, const mpa_system* systems, int dtype) {
    return plan_bytes(nsys, systems, dtype);
}

int mpa_bicgstab_batched(int nsys, const mpa_system* systems, int dtype, void* scratch, int* iters_out,
                         double* resid_out, int* status_out, int sync_every, void* stream) {
    if (!valid_systems(nsys, systems, dtype) || !scratch) {
        mpa_set_error("bicgstab_batched: bad argument");
        return MPA_ERR_ARG;
    }
    return mpa_internal_solve_batch(nsys, systems, dtype, scratch, iters_out, resid_out, status_out, sync_every,
                                    (hipStream_t)stream);
}

int mpa_systems_start(int nsys, const mpa_system* systems, int dtype, void* scratch, void* stream) {
    if (!valid_systems(nsys, systems, dtype) || !scratch) {
        mpa_set_error("systems_start: bad argument");
        return MPA_ERR_ARG;
    }
    hipStream_t st = (hipStream_t)stream;
    Plan P;
    int rc = build_plan(P, nsys, systems, dtype, scratch, st);
    if (rc) return rc;
    return dtype == MPA_C128 ? launch_start<double>(P, st) : launch_start<float>(P, st);
}

int mpa_systems_advance(int nsys, const mpa_system* systems, int dtype, void* scratch, int iterations, int* done_out,
                        int* iters_out, double* resid_out, int* status_out, void* stream) {
    if (!valid_systems(nsys, systems, dtype) || !scratch || iterations < 0) {
        mpa_set_error("systems_advance: bad argument");
        return MPA_ERR_ARG;
    }
    hipStream_t st = (hipStream_t)stream;
    Plan P;
    int rc = build_plan(P, nsys, systems, dtype, scratch, st);
    if (rc) return rc;
    std::vector<Scal> host;
    rc = read_scalars(P, host, st);
    if (rc) return rc;
    bool restart = false;
    for (const Scal& c : host) restart |= c.restart_pending != 0;
    if (restart) {
        rc = dtype == MPA_C128 ? launch_restart<double>(P, st) : launch_restart<float>(P, st);
        if (rc) return rc;
    }
    for (int k = 0; k < iterations; k++) {
        rc = dtype == MPA_C128 ? launch_iteration<double>(P, st) : launch_iteration<float>(P, st);
        if (rc) return rc;
    }
    rc = read_scalars(P, host, st);
    if (rc) return rc;
    for (size_t k = 0; k < host.size(); k++) {
        if (done_out) done_out[k] = host[k].done;
        if (iters_out) iters_out[k] = host[k].it;
        if (resid_out) resid_out[k] = host[k].rhs_sq > 0 ? std::sqrt(host[k].rnorm2 / host[k].rhs_sq) : 0.0;
        if (status_out) status_out[k] = host[k].status;
    }
    return MPA_OK;
}

size_t mpa_bicgstab_workspace_bytes(int n, int nrhs, int dtype) {
    mpa_system s{};
    s.n = n;
    s.nrhs = nrhs;
    return state_bytes(n, nrhs, dtype) + plan_bytes(1, &s, dtype);
}

int mpa_bicgstab(int n, int nrhs, int dtype, const void* A, int add_identity, const void* b, void* x, double tol,
                 int maxiter, void* workspace, int* iters_out, double* resid_out, int* status_out, int sync_every,
                 void* stream) {
    if (n < 2 || (n & 1) || nrhs < 1 || !A || !b || !x || !workspace || !(tol >= 0) || maxiter < 0 ||
        (dtype != MPA_C128 && dtype != MPA_C64)) {
        mpa_set_error("bicgstab: bad argument (n must be even and positive)");
        return MPA_ERR_ARG;
    }
    mpa_system s{};
    s.n = n;
    s.nrhs = nrhs;
    s.add_identity = add_identity;
    s.A = A;
    s.b = b;
    s.x = x;
    s.state = workspace;
    s.tol = tol;
    s.maxiter = maxiter;
    void* scratch = (char*)workspace + state_bytes(n, nrhs, dtype);
    return mpa_internal_solve_batch(1, &s, dtype, scratch, iters_out, resid_out, status_out, sync_every,
                                    (hipStream_t)stream);
}

size_t mpa_bicgstab_ex_workspace_bytes(int n, int nrhs, int dtype, int preconditioned) {
    mpa_system s{};
    s.n = n;
    s.nrhs = nrhs;
    return state_bytes(n, nrhs, dtype, preconditioned != 0) + plan_bytes(1, &s, dtype);
}

int mpa_bicgstab_ex(int n, int nrhs, int dtype, const void* A, int add_identity, const void* b, void* x, double tol,
                    int maxiter, int flags, const void* minv, int block_size, void* workspace, int* iters_out,
                    double* resid_out, int* status_out, int sync_every, void* stream) {
    mpa_system s{};
    s.n = n;
    s.nrhs = nrhs;
    s.add_identity = add_identity;
    s.flags = flags;
    s.A = A;
    s.b = b;
    s.x = x;
    s.state = workspace;
    s.tol = tol;
    s.maxiter = maxiter;
    s.minv = minv;
    s.block_size = block_size;
    if (!A || !workspace || !valid_systems(1, &s, dtype)) {
        mpa_set_error("bicgstab_ex: bad argument (n even and positive, block_size dividing n, known flags)");
        return MPA_ERR_ARG;
    }
    void* scratch = (char*)workspace + state_bytes(n, nrhs, dtype, minv != nullptr);
    return mpa_internal_solve_batch(1, &s, dtype, scratch, iters_out, resid_out, status_out, sync_every,
                                    (hipStream_t)stream);
}

int mpa_bicgstab_noalloc(int n, int nrhs, int dtype, const void* A, int add_identity, const void* b, void* x,
                         double tol, int maxiter, int* iters_out, double* resid_out, int* status_out, void* stream) {
    if (n < 2 || (n & 1) || nrhs < 1 || !A || !b || !x || !(tol >= 0) || maxiter < 0 ||
        (dtype != MPA_C128 && dtype != MPA_C64)) {
        mpa_set_error("bicgstab_noalloc: bad argument (n must be even and positive)");
        return MPA_ERR_ARG;
    }
    const size_t need = mpa_bicgstab_workspace_bytes(n, nrhs, dtype);
    if (need > kNoAllocScratchBytes) {
        mpa_set_error("bicgstab_noalloc: n = %d with %d right-hand sides needs %zu bytes of workspace, more than the "
                      "%zu-byte static scratch; call mpa_bicgstab with a workspace instead",
                      n, nrhs, need, kNoAllocScratchBytes);
        return MPA_ERR_ARG;
    }
    void* ws = nullptr;
    const int rc = noalloc_scratch(&ws);
    if (rc) return rc;
    mpa_system s{};
    s.n = n;
    s.nrhs = nrhs;
    s.add_identity = add_identity;
    s.A = A;
    s.b = b;
    s.x = x;
    s.state = ws;
    s.tol = tol;
    s.maxiter = maxiter;
    void* plan = (char*)ws + state_bytes(n, nrhs, dtype);
    hipStream_t st = (hipStream_t)stream;
    if (dtype == MPA_C128) return solve_batch<double>(1, &s, plan, iters_out, resid_out, status_out, 1, st, true);
    return solve_batch<float>(1, &s, plan, iters_out, resid_out, status_out, 1, st, true);
}

int mpa_bicgstab_matfree(const mpa_cluster_operator* op, int nrhs, int add_identity, const void* b, void* x,
                         double tol, int maxiter, void* workspace, int* iters_out, double* resid_out, int* status_out,
                         int sync_every, void* stream) {
    if (!op || op->N < 1 || op->lmax < 1 || op->lmax > MPA_LMAX_MAX) {
        mpa_set_error("bicgstab_matfree: bad operator");
        return MPA_ERR_ARG;
    }
    mpa_system s{};
    s.n = 2 * lmax_to_rmax(op->lmax) * op->N;
    s.nrhs = nrhs;
    s.add_identity = add_identity;
    s.b = b;
    s.x = x;
    s.state = workspace;
    s.tol = tol;
    s.maxiter = maxiter;
    s.op = op;
    if (!valid_systems(1, &s, MPA_C128) || !workspace) {
        mpa_set_error("bicgstab_matfree: bad argument");
        return MPA_ERR_ARG;
    }
    void* scratch = (char*)workspace + state_bytes(s.n, nrhs, MPA_C128);
    return mpa_internal_solve_batch(1, &s, MPA_C128, scratch, iters_out, resid_out, status_out, sync_every,
                                    (hipStream_t)stream);
}

}  // extern "C"
