// Mixed precision for batches of cluster systems: complex64 BiCGSTAB
// corrections, FP64 iterative refinement.
//
// Solves the reference's system (I + T) x = b (cpp/src/interactions.cpp:347-368)
// to an FP64 tolerance.  Per outer step and per RHS still above tolerance:
//   r = b - A64 x (FP64), d = BiCGSTAB_c64(A32, r / ||r||) to inner_tol,
//   x += ||r|| d (FP64).
// Outer steps stop per RHS at ||r|| <= tol ||b||, after max_refine
// corrections, or when a correction reduces ||r|| by less than 4x.  The
// inner solves of all systems run as one batch of the complex64 engine;
// the FP64 residual matvecs of all systems run as one batched c128 matvec.

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "gemv.hpp"
#include "krylov.hpp"
#include "matfree.hpp"

using namespace mpa;

int mpa_internal_solve_batch(int nsys, const mpa_system* systems, int dtype, void* scratch, int* iters, double* resid,
                             int* status, int sync_every, hipStream_t st);

namespace {

struct RefDev {
    const double2* b;
    double2* x;
    double2* r;
    double2* y;
    float2* r32;
    float2* d32;
    double2* part;  // [S][G]
    double* norms;  // [S]
    double* scale;  // [S]
    double* unscale;  // [S]
    int n, S, G, pad;
};

#define REF_PROLOGUE                       \
    const int3 wk = work[blockIdx.x];      \
    const RefDev& d = sys[wk.x];           \
    const int s = wk.y, g = wk.z;          \
    const int n = d.n;                     \
    const size_t base = (size_t)s * n;     \
    const int stride = d.G * kThreads;

// x = 0, r = b, partial ||b||^2
__global__ __launch_bounds__(kThreads) void k_ref_start(const RefDev* __restrict__ sys, const int3* __restrict__ work) {
    REF_PROLOGUE
    double2 acc[1] = {make_double2(0, 0)};
    for (int i = g * kThreads + threadIdx.x; i < n; i += stride) {
        const double2 bv = d.b[base + i];
        d.r[base + i] = bv;
        d.x[base + i] = make_double2(0, 0);
        acc[0].x += bv.x * bv.x + bv.y * bv.y;
    }
    block_sum<1>(acc);
    if (threadIdx.x == 0) d.part[(size_t)s * d.G + g] = acc[0];
}

// norms[s] = sum of the partials (fixed order), one block per (system, RHS)
__global__ __launch_bounds__(kThreads) void k_ref_norms(const RefDev* __restrict__ sys, const int2* __restrict__ cw) {
    const int2 c2 = cw[blockIdx.x];
    const RefDev& d = sys[c2.x];
    const int s = c2.y;
    __shared__ double red[kThreads];
    double a = 0;
    for (int g = threadIdx.x; g < d.G; g += kThreads) a += d.part[(size_t)s * d.G + g].x;
    red[threadIdx.x] = a;
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) d.norms[s] = red[0];
}

// Per-RHS values between the systems' records and one flat array in
// (system, RHS) order (the order of the cw list): mode 0 gathers norms into
// flat[0, total); mode 1 scatters scale = flat[total + k], unscale =
// flat[2 total + k]; mode 2 scatters unscale only.  One copy of the flat
// array then replaces one copy per system.
__global__ __launch_bounds__(kThreads) void k_ref_flat(const RefDev* __restrict__ sys, const int2* __restrict__ cw,
                                                       int total, double* __restrict__ flat, int mode) {
    const int k = blockIdx.x * kThreads + threadIdx.x;
    if (k >= total) return;
    const int2 c2 = cw[k];
    const RefDev& d = sys[c2.x];
    if (mode == 0) {
        flat[k] = d.norms[c2.y];
    } else {
        if (mode == 1) d.scale[c2.y] = flat[total + k];
        d.unscale[c2.y] = flat[2 * total + k];
    }
}

// r32 = r * scale, rounded to complex64
__global__ __launch_bounds__(kThreads) void k_ref_to32(const RefDev* __restrict__ sys, const int3* __restrict__ work) {
    REF_PROLOGUE
    const double f = d.scale[s];
    for (int i = g * kThreads + threadIdx.x; i < n; i += stride) {
        const double2 z = d.r[base + i];
        d.r32[base + i] = make_float2((float)(z.x * f), (float)(z.y * f));
    }
}

// x += d32 * unscale (FP64)
__global__ __launch_bounds__(kThreads) void k_ref_accum(const RefDev* __restrict__ sys, const int3* __restrict__ work) {
    REF_PROLOGUE
    const double f = d.unscale[s];
    if (f == 0.0) return;
    for (int i = g * kThreads + threadIdx.x; i < n; i += stride) {
        const float2 z = d.d32[base + i];
        double2 xv = d.x[base + i];
        xv.x += (double)z.x * f;
        xv.y += (double)z.y * f;
        d.x[base + i] = xv;
    }
}

// r = b - y, partial ||r||^2
__global__ __launch_bounds__(kThreads) void k_ref_residual(const RefDev* __restrict__ sys, const int3* __restrict__ work) {
    REF_PROLOGUE
    double2 acc[1] = {make_double2(0, 0)};
    for (int i = g * kThreads + threadIdx.x; i < n; i += stride) {
        const double2 rv = csub(d.b[base + i], d.y[base + i]);
        d.r[base + i] = rv;
        acc[0].x += rv.x * rv.x + rv.y * rv.y;
    }
    block_sum<1>(acc);
    if (threadIdx.x == 0) d.part[(size_t)s * d.G + g] = acc[0];
}

size_t v64(int n, int S) { return align256(sizeof(double2) * (size_t)n * S); }
size_t v32(int n, int S) { return align256(sizeof(float2) * (size_t)n * S); }

// Per-RHS control record of the iteration-level refinement
// (mpa_refine_systems_start / _advance), kept in the system's state between
// calls.  The system's phase is held by its RHS 0.
struct RefCtl {
    double bn2, rn2, prev, inner_tol;
    int outer, inner_it, status, phase;  // phase: 0 decide, 1 correction running, 2 finished
};

// Control records of every (system, RHS) between the systems' states and one
// flat array (cw order): to_flat 1 gathers, 0 scatters; ctlp[i] is system
// i's record array.  One copy of the flat array replaces one per system.
__global__ __launch_bounds__(kThreads) void k_ref_ctl_flat(RefCtl* const* __restrict__ ctlp,
                                                           const int2* __restrict__ cw, int total,
                                                           RefCtl* __restrict__ flat, int to_flat) {
    const int k = blockIdx.x * kThreads + threadIdx.x;
    if (k >= total) return;
    const int2 c2 = cw[k];
    if (to_flat) flat[k] = ctlp[c2.x][c2.y];
    else ctlp[c2.x][c2.y] = flat[k];
}

size_t own_bytes(int n, int S) {
    return 2 * v64(n, S) + 2 * v32(n, S) + align256(sizeof(double2) * (size_t)S * groups_for(n)) +
           3 * align256(sizeof(double) * S) + align256(sizeof(RefCtl) * S);
}

bool valid(int nsys, const mpa_refine_system* s) {
    if (nsys < 1 || !s) return false;
    for (int i = 0; i < nsys; i++) {
        if (s[i].n < 2 || (s[i].n & 1) || s[i].nrhs < 1 || !s[i].b || !s[i].x || !s[i].state ||
            !(s[i].tol > 0) || !(s[i].inner_tol > 0) || s[i].maxiter < 1 || s[i].max_refine < 1)
            return false;
        if (s[i].op) {
            const mpa_cluster_operator& o = *s[i].op;
            if (s[i].A32 || s[i].A64 || o.N < 1 || o.lmax < 1 || o.lmax > MPA_LMAX_MAX || !o.positions || !o.mie ||
                o.tmatrix || !o.workspace || !(o.k > 0) || (long long)2 * lmax_to_rmax(o.lmax) * o.N != s[i].n ||
                !matfree_c64_supported(o.N, o.lmax, false))
                return false;
        } else if (!s[i].A32 || !s[i].A64) {
            return false;
        }
    }
    return true;
}

// The FP64 residual products of one refinement step: the batched GEMV over
// the stored complex128 matrices and the FP64 matrix-free products; with
// both present (and MPA_OPT_OVERLAP) the GEMV runs on the device's side
// stream beside the matrix-free products, joined before the caller's next
// launch, as in the solver's product phases.
int residual_products(const MvDev* d_mv, int2* const* d_mw, const int* n_mw, const std::vector<MfJob>& mf,
                      hipStream_t st) {
    bool any_mv = false;
    for (int c = 0; c < kTileClasses; c++) any_mv |= n_mw[c] > 0;
    if (any_mv && !mf.empty() && mpa_option(MPA_OPT_OVERLAP) != 0 && !mpa_profile_active()) {
        SideLease lease;
        const SideStream* side = lease.get();
        if (side) {
            MPA_HIP_CHECK(hipEventRecord(side->fork, st));
            MPA_HIP_CHECK(hipStreamWaitEvent(side->stream, side->fork, 0));
            const int rc1 = gemv_launch(8, d_mv, nullptr, d_mw, n_mw, side->stream);
            const int rc2 = matfree_launch_many(mf, nullptr, st);
            MPA_HIP_CHECK(hipEventRecord(side->join, side->stream));
            MPA_HIP_CHECK(hipStreamWaitEvent(st, side->join, 0));
            return rc1 ? rc1 : rc2;
        }
    }
    int rc = MPA_OK;
    if (any_mv && (rc = gemv_launch(8, d_mv, nullptr, d_mw, n_mw, st))) return rc;
    if (!mf.empty() && (rc = matfree_launch_many(mf, nullptr, st))) return rc;
    return MPA_OK;
}

int refine_batch(int nsys, const mpa_refine_system* R, void* scratch, int* outer_out, int* inner_out,
                 double* resid_out, int* status_out, hipStream_t st) {
    // descriptors
    std::vector<RefDev> ref(nsys);
    std::vector<mpa_system> inner(nsys);
    std::vector<MvDev> mv(nsys);
    std::vector<int3> vw;
    std::vector<int2> cw;
    std::vector<int2> mw[kTileClasses];
    std::vector<MfJob> mf;
    int total = 0;
    for (int i = 0; i < nsys; i++) {
        const mpa_refine_system& q = R[i];
        const int n = q.n, S = q.nrhs;
        RefDev& d = ref[i];
        char* p = (char*)q.state;
        d.b = (const double2*)q.b;
        d.x = (double2*)q.x;
        d.r = (double2*)p; p += v64(n, S);
        d.y = (double2*)p; p += v64(n, S);
        d.r32 = (float2*)p; p += v32(n, S);
        d.d32 = (float2*)p; p += v32(n, S);
        d.part = (double2*)p; p += align256(sizeof(double2) * (size_t)S * groups_for(n));
        d.norms = (double*)p; p += align256(sizeof(double) * S);
        d.scale = (double*)p; p += align256(sizeof(double) * S);
        d.unscale = (double*)p; p += align256(sizeof(double) * S);
        p += align256(sizeof(RefCtl) * S);  // control records of the iteration-level driver
        d.n = n;
        d.S = S;
        d.G = groups_for(n);
        d.pad = 0;
        mpa_system& in = inner[i];
        in.n = n;
        in.nrhs = S;
        in.add_identity = q.add_identity;
        in.A = q.op ? nullptr : q.A32;
        in.op = q.op;
        in.b = d.r32;
        in.x = d.d32;
        in.state = p;
        in.tol = q.inner_tol;
        in.maxiter = q.maxiter;
        mv[i] = MvDev{q.A64, q.x, d.y, n, S, q.add_identity, -1};
        for (int s = 0; s < S; s++) {
            cw.push_back(make_int2(i, s));
            for (int g = 0; g < d.G; g++) vw.push_back(make_int3(i, s, g));
        }
        if (q.op) {  // FP64 residual with the matrix-free operator
            const mpa_cluster_operator& o = *q.op;
            mf.push_back(MfJob{o.N, o.lmax, S, q.add_identity ? 1 : 0, o.positions, (const double2*)o.mie, o.k,
                               (const double2*)q.x, d.y, (double2*)o.workspace, -1, 0, nullptr, nullptr, 0, 0});
        } else {
            const int rb = gemv_row_blocks(n, S, 8);  // the FP64 residual matvec
            for (int b = 0; b < rb; b++) mw[gemv_class(S)].push_back(make_int2(i, b));
        }
        total += S;
    }
    for (auto& w : mw) order_work_longest_first(w, [&](int i) { return R[i].n; });
    char* p = (char*)scratch;
    RefDev* d_ref = (RefDev*)p; p += align256(sizeof(RefDev) * nsys);
    MvDev* d_mv = (MvDev*)p; p += align256(sizeof(MvDev) * nsys);
    int3* d_vw = (int3*)p; p += align256(sizeof(int3) * vw.size());
    int2* d_cw = (int2*)p; p += align256(sizeof(int2) * cw.size());
    int2* d_mw[kTileClasses];
    int n_mw[kTileClasses];
    for (int c = 0; c < kTileClasses; c++) {
        d_mw[c] = (int2*)p;
        n_mw[c] = (int)mw[c].size();
        p += align256(sizeof(int2) * mw[c].size());
    }
    double* d_flat = (double*)p;  // [3 total]: gathered norms, scale, unscale
    p += align256(sizeof(double) * 3 * total);
    void* inner_scratch = p;
    {  // descriptors and work lists in one host image with the device layout, one copy
        const size_t nbytes = (size_t)(p - (char*)scratch);
        std::vector<char> img(nbytes, 0);
        auto put = [&](const void* dst, const void* src, size_t bytes) {
            if (bytes) std::memcpy(img.data() + ((const char*)dst - (const char*)scratch), src, bytes);
        };
        put(d_ref, ref.data(), sizeof(RefDev) * nsys);
        put(d_mv, mv.data(), sizeof(MvDev) * nsys);
        put(d_vw, vw.data(), sizeof(int3) * vw.size());
        put(d_cw, cw.data(), sizeof(int2) * cw.size());
        for (int c = 0; c < kTileClasses; c++) put(d_mw[c], mw[c].data(), sizeof(int2) * n_mw[c]);
        // pageable source: staged by the runtime before the call returns
        MPA_HIP_CHECK(hipMemcpyAsync(scratch, img.data(), nbytes, hipMemcpyHostToDevice, st));
    }

    const int nv = (int)vw.size(), nc = (int)cw.size();
    hipLaunchKernelGGL(k_ref_start, dim3(nv), dim3(kThreads), 0, st, d_ref, d_vw);
    hipLaunchKernelGGL(k_ref_norms, dim3(nc), dim3(kThreads), 0, st, d_ref, d_cw);
    MPA_LAUNCH_CHECK();
    // host copies of per-RHS state, flattened in (system, RHS) order
    std::vector<double> bn2(total), rn2(total), prev(total, INFINITY), sc(total), usc(total);
    std::vector<int> outer(total, 0), inner_it(total, 0), status(total, MPA_NOT_CONVERGED);
    std::vector<int> it(total), ist(total);
    std::vector<double> ires(total);
    const unsigned flat_grid = (unsigned)((total + kThreads - 1) / kThreads);
    // one system: its own arrays directly; several: one flat array through the gather/scatter kernel
    const bool flat = nsys > 1;
    auto read_norms = [&](std::vector<double>& out) -> int {
        if (flat) {
            hipLaunchKernelGGL(k_ref_flat, dim3(flat_grid), dim3(kThreads), 0, st, d_ref, d_cw, total, d_flat, 0);
            MPA_LAUNCH_CHECK();
        }
        MPA_HIP_CHECK(hipMemcpyAsync(out.data(), flat ? d_flat : ref[0].norms, sizeof(double) * total,
                                     hipMemcpyDeviceToHost, st));
        MPA_HIP_CHECK(hipStreamSynchronize(st));
        return MPA_OK;
    };
    std::vector<double> flat_up(2 * (size_t)total);
    int rc = read_norms(bn2);
    if (rc) return rc;
    rn2 = bn2;
    for (int step = 0;; step++) {
        bool any = false;
        int k = 0;
        for (int i = 0; i < nsys; i++) {
            for (int s = 0; s < ref[i].S; s++, k++) {
                const double tol = R[i].tol;
                int stt;
                if (!std::isfinite(rn2[k])) stt = MPA_NONFINITE;
                else if (rn2[k] <= tol * tol * bn2[k]) stt = MPA_OK;
                else if (step > 0 && rn2[k] > 0.0625 * prev[k]) stt = MPA_NOT_CONVERGED;  // < 4x per step
                else if (step >= R[i].max_refine) stt = MPA_NOT_CONVERGED;
                else stt = -1;
                status[k] = stt < 0 ? MPA_NOT_CONVERGED : stt;
                const bool active = stt < 0;
                sc[k] = active ? 1.0 / std::sqrt(rn2[k]) : 0.0;
                usc[k] = active ? std::sqrt(rn2[k]) : 0.0;
                any |= active;
            }
        }
        if (!any) break;
        // each correction only needs to reach the remaining factor to tol
        // (with a 4x margin), never tighter than the configured inner_tol
        k = 0;
        for (int i = 0; i < nsys; i++) {
            double need = INFINITY;
            for (int s = 0; s < ref[i].S; s++)
                if (usc[k + s] != 0.0) need = std::min(need, 0.25 * R[i].tol * std::sqrt(bn2[k + s] / rn2[k + s]));
            inner[i].tol = std::isfinite(need) ? std::max(R[i].inner_tol, need) : R[i].inner_tol;
            k += ref[i].S;
        }
        if (flat) {
            std::copy(sc.begin(), sc.end(), flat_up.begin());
            std::copy(usc.begin(), usc.end(), flat_up.begin() + total);
            // pageable source: staged by the runtime before the call returns
            MPA_HIP_CHECK(hipMemcpyAsync(d_flat + total, flat_up.data(), sizeof(double) * 2 * total,
                                         hipMemcpyHostToDevice, st));
            hipLaunchKernelGGL(k_ref_flat, dim3(flat_grid), dim3(kThreads), 0, st, d_ref, d_cw, total, d_flat, 1);
        } else {
            MPA_HIP_CHECK(hipMemcpyAsync(ref[0].scale, sc.data(), sizeof(double) * total, hipMemcpyHostToDevice, st));
            MPA_HIP_CHECK(hipMemcpyAsync(ref[0].unscale, usc.data(), sizeof(double) * total, hipMemcpyHostToDevice, st));
        }
        hipLaunchKernelGGL(k_ref_to32, dim3(nv), dim3(kThreads), 0, st, d_ref, d_vw);
        MPA_LAUNCH_CHECK();
        rc = mpa_internal_solve_batch(nsys, inner.data(), MPA_C64, inner_scratch, it.data(), ires.data(), ist.data(), 1,
                                      st);
        if (rc != MPA_OK && rc != MPA_NOT_CONVERGED && rc != MPA_NONFINITE) return rc;
        for (int j = 0; j < total; j++)
            if (usc[j] != 0.0) {
                inner_it[j] += it[j];
                outer[j] = step + 1;
            }
        hipLaunchKernelGGL(k_ref_accum, dim3(nv), dim3(kThreads), 0, st, d_ref, d_vw);
        MPA_LAUNCH_CHECK();
        if ((rc = residual_products(d_mv, d_mw, n_mw, mf, st))) return rc;
        hipLaunchKernelGGL(k_ref_residual, dim3(nv), dim3(kThreads), 0, st, d_ref, d_vw);
        hipLaunchKernelGGL(k_ref_norms, dim3(nc), dim3(kThreads), 0, st, d_ref, d_cw);
        MPA_LAUNCH_CHECK();
        prev = rn2;
        rc = read_norms(rn2);
        if (rc) return rc;
        // a correction that left the residual larger (a failed inner solve) is taken back:
        // x -= ||r|| d, and the RHS keeps its previous residual (it stops as not converged)
        bool any_revert = false;
        for (int j = 0; j < total; j++) {
            const bool revert = usc[j] != 0.0 && !(rn2[j] <= prev[j]);
            usc[j] = revert ? -usc[j] : 0.0;
            if (revert) {
                rn2[j] = prev[j];
                any_revert = true;
            }
        }
        if (any_revert) {
            MPA_HIP_CHECK(hipMemcpyAsync(flat ? d_flat + 2 * total : ref[0].unscale, usc.data(), sizeof(double) * total,
                                         hipMemcpyHostToDevice, st));
            if (flat)
                hipLaunchKernelGGL(k_ref_flat, dim3(flat_grid), dim3(kThreads), 0, st, d_ref, d_cw, total, d_flat, 2);
            hipLaunchKernelGGL(k_ref_accum, dim3(nv), dim3(kThreads), 0, st, d_ref, d_vw);
            MPA_LAUNCH_CHECK();
            MPA_HIP_CHECK(hipStreamSynchronize(st));  // usc is read by the copies above
        }
    }
    int worst = MPA_OK;
    for (int j = 0; j < total; j++) {
        if (outer_out) outer_out[j] = outer[j];
        if (inner_out) inner_out[j] = inner_it[j];
        if (resid_out) resid_out[j] = bn2[j] > 0 ? std::sqrt(rn2[j] / bn2[j]) : 0.0;
        if (status_out) status_out[j] = status[j];
        if (status[j] != MPA_OK && worst == MPA_OK) worst = status[j];
    }
    return worst;
}

// ---- iteration-level refinement (serving: systems join and leave between calls) ----
//
// The same algorithm as refine_batch, split into calls: each call advances
// the running corrections of all given systems by `iterations` complex64
// BiCGSTAB iterations; a system whose correction finished gets its FP64
// update, residual and per-RHS decision in the same call, and its next
// correction is started (or the system is finished).  Per-RHS control
// records live in each system's state, so any subset of systems can be
// advanced together in any call.

struct RefView {
    RefDev d;
    RefCtl* ctl;      // device, [S]
    void* inner_state;
};

RefView ref_view(const mpa_refine_system& q) {
    RefView v{};
    const int n = q.n, S = q.nrhs;
    char* p = (char*)q.state;
    v.d.b = (const double2*)q.b;
    v.d.x = (double2*)q.x;
    v.d.r = (double2*)p; p += v64(n, S);
    v.d.y = (double2*)p; p += v64(n, S);
    v.d.r32 = (float2*)p; p += v32(n, S);
    v.d.d32 = (float2*)p; p += v32(n, S);
    v.d.part = (double2*)p; p += align256(sizeof(double2) * (size_t)S * groups_for(n));
    v.d.norms = (double*)p; p += align256(sizeof(double) * S);
    v.d.scale = (double*)p; p += align256(sizeof(double) * S);
    v.d.unscale = (double*)p; p += align256(sizeof(double) * S);
    v.ctl = (RefCtl*)p; p += align256(sizeof(RefCtl) * S);
    v.inner_state = p;
    v.d.n = n;
    v.d.S = S;
    v.d.G = groups_for(n);
    return v;
}

mpa_system inner_system(const mpa_refine_system& q, const RefView& v, double tol) {
    mpa_system in{};
    in.n = q.n;
    in.nrhs = q.nrhs;
    in.add_identity = q.add_identity;
    in.A = q.op ? nullptr : q.A32;
    in.op = q.op;
    in.b = v.d.r32;
    in.x = v.d.d32;
    in.state = v.inner_state;
    in.tol = tol;
    in.maxiter = q.maxiter;
    return in;
}

// Work lists over a subset of systems (indices into the full descriptor array).
struct SubLists {
    std::vector<int3> vw;
    std::vector<int2> cw;
    std::vector<int2> mw[kTileClasses];
    std::vector<MfJob> mf;
};

SubLists sub_lists(const std::vector<int>& idx, const mpa_refine_system* R, const std::vector<RefView>& V) {
    SubLists L;
    for (int i : idx) {
        const mpa_refine_system& q = R[i];
        for (int s = 0; s < q.nrhs; s++) {
            L.cw.push_back(make_int2(i, s));
            for (int g = 0; g < V[i].d.G; g++) L.vw.push_back(make_int3(i, s, g));
        }
        if (q.op) {
            const mpa_cluster_operator& o = *q.op;
            L.mf.push_back(MfJob{o.N, o.lmax, q.nrhs, q.add_identity ? 1 : 0, o.positions, (const double2*)o.mie, o.k,
                                 (const double2*)q.x, V[i].d.y, (double2*)o.workspace, -1, 0, nullptr, nullptr, 0,
                                 0});
        } else {
            const int rb = gemv_row_blocks(q.n, q.nrhs, 8);
            for (int b = 0; b < rb; b++) L.mw[gemv_class(q.nrhs)].push_back(make_int2(i, b));
        }
    }
    for (auto& w : L.mw) order_work_longest_first(w, [&](int i) { return R[i].n; });
    return L;
}

int refine_iterate(int nsys, const mpa_refine_system* R, void* scratch, bool start, int iterations, int* done_out,
                   int* outer_out, int* inner_out, double* resid_out, int* status_out, hipStream_t st) {
    std::vector<RefView> V(nsys);
    std::vector<RefDev> ref(nsys);
    std::vector<MvDev> mv(nsys);
    std::vector<int> first(nsys + 1, 0);
    for (int i = 0; i < nsys; i++) {
        V[i] = ref_view(R[i]);
        ref[i] = V[i].d;
        mv[i] = MvDev{R[i].A64, R[i].x, V[i].d.y, R[i].n, R[i].nrhs, R[i].add_identity, -1};
        first[i + 1] = first[i] + R[i].nrhs;
    }
    const int total = first[nsys];
    // scratch: descriptors, one region each for the (subset) work lists, then the inner engine's
    char* p = (char*)scratch;
    RefDev* d_ref = (RefDev*)p; p += align256(sizeof(RefDev) * nsys);
    MvDev* d_mv = (MvDev*)p; p += align256(sizeof(MvDev) * nsys);
    size_t nvw = 0, nmw = 0;
    for (int i = 0; i < nsys; i++) {
        nvw += (size_t)R[i].nrhs * V[i].d.G;
        nmw += gemv_row_blocks(R[i].n, R[i].nrhs, 8);
    }
    int3* d_vw = (int3*)p; p += align256(sizeof(int3) * nvw);
    int2* d_cw = (int2*)p; p += align256(sizeof(int2) * total);
    int2* d_mwbase = (int2*)p; p += kTileClasses * 256 + align256(sizeof(int2) * nmw);
    // control records: per-system pointers and the full (system, RHS) list, then the flat records
    char* ctl_region = p;
    RefCtl** d_ctlp = (RefCtl**)p; p += align256(sizeof(void*) * nsys);
    int2* d_cwall = (int2*)p; p += align256(sizeof(int2) * total);
    const size_t ctl_head = (size_t)(p - ctl_region);
    RefCtl* d_ctlflat = (RefCtl*)p; p += align256(sizeof(RefCtl) * total);
    double* d_flat = (double*)p; p += align256(sizeof(double) * 3 * total);  // per-RHS values of a subset
    void* inner_scratch = p;
    MPA_HIP_CHECK(hipMemcpyAsync(d_ref, ref.data(), sizeof(RefDev) * nsys, hipMemcpyHostToDevice, st));
    MPA_HIP_CHECK(hipMemcpyAsync(d_mv, mv.data(), sizeof(MvDev) * nsys, hipMemcpyHostToDevice, st));
    std::vector<char> ctl_img(ctl_head, 0);
    for (int i = 0; i < nsys; i++) {
        std::memcpy(ctl_img.data() + sizeof(void*) * i, &V[i].ctl, sizeof(void*));
        for (int s = 0; s < R[i].nrhs; s++) {
            const int2 e = make_int2(i, s);
            std::memcpy(ctl_img.data() + ((char*)d_cwall - ctl_region) + sizeof(int2) * (first[i] + s), &e, sizeof(int2));
        }
    }
    // pageable source: staged by the runtime before the call returns
    MPA_HIP_CHECK(hipMemcpyAsync(ctl_region, ctl_img.data(), ctl_head, hipMemcpyHostToDevice, st));
    const unsigned ctl_grid = (unsigned)((total + kThreads - 1) / kThreads);
    auto upload = [&](const SubLists& L, int2** d_mw, int* n_mw) -> int {
        if (!L.vw.empty())
            MPA_HIP_CHECK(hipMemcpyAsync(d_vw, L.vw.data(), sizeof(int3) * L.vw.size(), hipMemcpyHostToDevice, st));
        if (!L.cw.empty())
            MPA_HIP_CHECK(hipMemcpyAsync(d_cw, L.cw.data(), sizeof(int2) * L.cw.size(), hipMemcpyHostToDevice, st));
        char* q = (char*)d_mwbase;
        for (int c = 0; c < kTileClasses; c++) {
            d_mw[c] = (int2*)q;
            n_mw[c] = (int)L.mw[c].size();
            if (n_mw[c])
                MPA_HIP_CHECK(hipMemcpyAsync(d_mw[c], L.mw[c].data(), sizeof(int2) * n_mw[c], hipMemcpyHostToDevice, st));
            q += align256(sizeof(int2) * L.mw[c].size());
        }
        return MPA_OK;
    };
    int2* d_mw[kTileClasses];
    int n_mw[kTileClasses];
    std::vector<RefCtl> ctl(total);
    std::vector<double> nrm(total);
    auto read_ctl = [&]() -> int {
        hipLaunchKernelGGL(k_ref_ctl_flat, dim3(ctl_grid), dim3(kThreads), 0, st, d_ctlp, d_cwall, total, d_ctlflat, 1);
        MPA_LAUNCH_CHECK();
        MPA_HIP_CHECK(hipMemcpyAsync(ctl.data(), d_ctlflat, sizeof(RefCtl) * total, hipMemcpyDeviceToHost, st));
        MPA_HIP_CHECK(hipStreamSynchronize(st));
        return MPA_OK;
    };
    // Per-RHS values of a subset of systems move through d_flat in the order of
    // the subset's uploaded (system, RHS) list d_cw: one gather kernel and one
    // copy, or one copy and one scatter kernel, however many systems take part.
    // two host images (revert values, decisions): the final synchronization of this call ends every copy
    // from them before they go out of scope
    std::vector<double> flat_host(3 * (size_t)total), flat_dec(3 * (size_t)total);
    auto read_norms = [&](const std::vector<int>& idx) -> int {
        int cnt = 0;
        for (int i : idx) cnt += R[i].nrhs;
        hipLaunchKernelGGL(k_ref_flat, dim3((unsigned)((cnt + kThreads - 1) / kThreads)), dim3(kThreads), 0, st, d_ref,
                           d_cw, cnt, d_flat, 0);
        MPA_LAUNCH_CHECK();
        MPA_HIP_CHECK(hipMemcpyAsync(flat_host.data(), d_flat, sizeof(double) * cnt, hipMemcpyDeviceToHost, st));
        MPA_HIP_CHECK(hipStreamSynchronize(st));
        int o = 0;
        for (int i : idx)
            for (int s = 0; s < R[i].nrhs; s++) nrm[first[i] + s] = flat_host[o++];
        return MPA_OK;
    };
    // scale (mode 1: img[cnt + k]) and unscale (img[2 cnt + k]) of the subset in d_cw
    auto write_scales = [&](const std::vector<double>& img, int cnt, int mode) -> int {
        const size_t off = mode == 1 ? (size_t)cnt : 2 * (size_t)cnt;
        const size_t len = mode == 1 ? 2 * (size_t)cnt : (size_t)cnt;
        // pageable source: staged by the runtime before the call returns
        MPA_HIP_CHECK(hipMemcpyAsync(d_flat + off, img.data() + off, sizeof(double) * len,
                                     hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_ref_flat, dim3((unsigned)((cnt + kThreads - 1) / kThreads)), dim3(kThreads), 0, st, d_ref,
                           d_cw, cnt, d_flat, mode);
        MPA_LAUNCH_CHECK();
        return MPA_OK;
    };
    int rc;
    std::vector<int> all(nsys);
    for (int i = 0; i < nsys; i++) all[i] = i;
    if (start) {
        // x = 0, r = b, ||b||^2; every system starts in the decision phase
        SubLists L = sub_lists(all, R, V);
        if ((rc = upload(L, d_mw, n_mw))) return rc;
        hipLaunchKernelGGL(k_ref_start, dim3((unsigned)L.vw.size()), dim3(kThreads), 0, st, d_ref, d_vw);
        hipLaunchKernelGGL(k_ref_norms, dim3((unsigned)L.cw.size()), dim3(kThreads), 0, st, d_ref, d_cw);
        MPA_LAUNCH_CHECK();
        if ((rc = read_norms(all))) return rc;
        for (int k = 0; k < total; k++)
            ctl[k] = RefCtl{nrm[k], nrm[k], INFINITY, 0.0, 0, 0, MPA_NOT_CONVERGED, 0};
    } else {
        if ((rc = read_ctl())) return rc;
    }
    // A. advance the running corrections
    std::vector<int> running, finished_inner;
    for (int i = 0; i < nsys; i++)
        if (ctl[first[i]].phase == 1) running.push_back(i);
    if (!running.empty() && iterations > 0) {
        std::vector<mpa_system> in;
        int cnt = 0;
        for (int i : running) {
            in.push_back(inner_system(R[i], V[i], ctl[first[i]].inner_tol));
            cnt += R[i].nrhs;
        }
        std::vector<int> dn(cnt), its(cnt), stt(cnt);
        std::vector<double> res(cnt);
        rc = mpa_systems_advance((int)in.size(), in.data(), MPA_C64, inner_scratch, iterations, dn.data(), its.data(),
                                 res.data(), stt.data(), st);
        if (rc) return rc;
        int o = 0;
        for (int i : running) {
            bool all_done = true;
            for (int s = 0; s < R[i].nrhs; s++) all_done &= dn[o + s] != 0;
            if (all_done) {
                for (int s = 0; s < R[i].nrhs; s++) {
                    RefCtl& c = ctl[first[i] + s];
                    if (c.status == -1) {  // this RHS had a correction running
                        c.inner_it += its[o + s];
                        c.outer += 1;
                    }
                }
                finished_inner.push_back(i);
            }
            o += R[i].nrhs;
        }
    }
    // B. finished corrections: x += ||r|| d, FP64 residual and norms
    if (!finished_inner.empty()) {
        SubLists L = sub_lists(finished_inner, R, V);
        if ((rc = upload(L, d_mw, n_mw))) return rc;
        hipLaunchKernelGGL(k_ref_accum, dim3((unsigned)L.vw.size()), dim3(kThreads), 0, st, d_ref, d_vw);
        MPA_LAUNCH_CHECK();
        if ((rc = residual_products(d_mv, d_mw, n_mw, L.mf, st))) return rc;
        hipLaunchKernelGGL(k_ref_residual, dim3((unsigned)L.vw.size()), dim3(kThreads), 0, st, d_ref, d_vw);
        hipLaunchKernelGGL(k_ref_norms, dim3((unsigned)L.cw.size()), dim3(kThreads), 0, st, d_ref, d_cw);
        MPA_LAUNCH_CHECK();
        if ((rc = read_norms(finished_inner))) return rc;
        std::vector<int> revert_sys;
        const int cnt = (int)L.cw.size();
        int o = 0;
        for (int i : finished_inner) {
            bool any = false;
            for (int s = 0; s < R[i].nrhs; s++, o++) {
                RefCtl& c = ctl[first[i] + s];
                const double now = nrm[first[i] + s];
                flat_host[2 * (size_t)cnt + o] = 0.0;
                if (c.status == -1 && !(now <= c.rn2)) {
                    // the correction left the residual larger (a failed inner solve): take it back
                    flat_host[2 * (size_t)cnt + o] = -std::sqrt(c.rn2);
                    c.prev = c.rn2;  // residual unchanged: the decision below stops this RHS
                    any = true;
                } else {
                    c.prev = c.rn2;
                    c.rn2 = now;
                }
            }
            if (any) revert_sys.push_back(i);
            ctl[first[i]].phase = 0;
        }
        if (!revert_sys.empty()) {
            // unscale = -||r|| for the reverting RHS, 0 for the others of the finished set (every one of
            // them is decided afresh below, which sets its scales again before any further use)
            if ((rc = write_scales(flat_host, cnt, 2))) return rc;
            SubLists Lr = sub_lists(revert_sys, R, V);
            if ((rc = upload(Lr, d_mw, n_mw))) return rc;
            hipLaunchKernelGGL(k_ref_accum, dim3((unsigned)Lr.vw.size()), dim3(kThreads), 0, st, d_ref, d_vw);
            MPA_LAUNCH_CHECK();
        }
    }
    // C. decisions: per RHS converged / stalled / out of corrections, or the next correction
    std::vector<int> deciding;
    for (int i = 0; i < nsys; i++)
        if (ctl[first[i]].phase == 0) deciding.push_back(i);
    if (!deciding.empty()) {
        std::vector<int> starting;
        for (int i : deciding) {
            bool any = false;
            double need = INFINITY;
            for (int s = 0; s < R[i].nrhs; s++) {
                // every RHS is decided afresh, as in refine_batch: one that took no part in the last
                // correction keeps its residual, so a converged RHS stays converged and a stalled one
                // stays stalled
                RefCtl& c = ctl[first[i] + s];
                const double tol = R[i].tol;
                int d;
                if (!std::isfinite(c.rn2)) d = MPA_NONFINITE;
                else if (c.rn2 <= tol * tol * c.bn2) d = MPA_OK;
                else if (c.outer > 0 && c.rn2 > 0.0625 * c.prev) d = MPA_NOT_CONVERGED;  // < 4x per step
                else if (c.outer >= R[i].max_refine) d = MPA_NOT_CONVERGED;
                else d = -1;
                c.status = d;
                if (d == -1) {
                    any = true;
                    need = std::min(need, 0.25 * tol * std::sqrt(c.bn2 / c.rn2));
                }
            }
            if (any) {
                const double itol = std::isfinite(need) ? std::max(R[i].inner_tol, need) : R[i].inner_tol;
                for (int s = 0; s < R[i].nrhs; s++) ctl[first[i] + s].inner_tol = itol;
                ctl[first[i]].phase = 1;
                starting.push_back(i);
            } else {
                ctl[first[i]].phase = 2;
            }
        }
        if (!starting.empty()) {
            SubLists L = sub_lists(starting, R, V);
            if ((rc = upload(L, d_mw, n_mw))) return rc;
            // scale = 1 / ||r||, unscale = ||r|| of every RHS that starts a correction (0 for the others),
            // in the order of the starting set's (system, RHS) list
            const int cnt = (int)L.cw.size();
            int o = 0;
            for (int i : starting)
                for (int s = 0; s < R[i].nrhs; s++, o++) {
                    const RefCtl& c = ctl[first[i] + s];
                    const bool active = c.status == -1;
                    flat_dec[(size_t)cnt + o] = active ? 1.0 / std::sqrt(c.rn2) : 0.0;
                    flat_dec[2 * (size_t)cnt + o] = active ? std::sqrt(c.rn2) : 0.0;
                }
            if ((rc = write_scales(flat_dec, cnt, 1))) return rc;
            hipLaunchKernelGGL(k_ref_to32, dim3((unsigned)L.vw.size()), dim3(kThreads), 0, st, d_ref, d_vw);
            MPA_LAUNCH_CHECK();
            std::vector<mpa_system> in;
            for (int i : starting) in.push_back(inner_system(R[i], V[i], ctl[first[i]].inner_tol));
            if ((rc = mpa_systems_start((int)in.size(), in.data(), MPA_C64, inner_scratch, st))) return rc;
        }
    }
    MPA_HIP_CHECK(hipMemcpyAsync(d_ctlflat, ctl.data(), sizeof(RefCtl) * total, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_ref_ctl_flat, dim3(ctl_grid), dim3(kThreads), 0, st, d_ctlp, d_cwall, total, d_ctlflat, 0);
    MPA_LAUNCH_CHECK();
    MPA_HIP_CHECK(hipStreamSynchronize(st));
    for (int i = 0; i < nsys; i++)
        for (int s = 0; s < R[i].nrhs; s++) {
            const int k = first[i] + s;
            const RefCtl& c = ctl[k];
            const bool fin = ctl[first[i]].phase == 2;
            if (done_out) done_out[k] = fin ? 1 : 0;
            if (outer_out) outer_out[k] = c.outer;
            if (inner_out) inner_out[k] = c.inner_it;
            if (resid_out) resid_out[k] = c.bn2 > 0 ? std::sqrt(c.rn2 / c.bn2) : 0.0;
            if (status_out) status_out[k] = c.status == -1 ? MPA_NOT_CONVERGED : c.status;
        }
    return MPA_OK;
}

size_t refine_scratch_bytes(int nsys, const mpa_refine_system* R) {
    size_t vw = 0, cw = 0, mw = 0;
    std::vector<mpa_system> inner(nsys);
    for (int i = 0; i < nsys; i++) {
        vw += (size_t)R[i].nrhs * groups_for(R[i].n);
        cw += R[i].nrhs;
        mw += gemv_row_blocks(R[i].n, R[i].nrhs, 8);
        inner[i].n = R[i].n;
        inner[i].nrhs = R[i].nrhs;
    }
    return align256(sizeof(RefDev) * nsys) + align256(sizeof(MvDev) * nsys) + align256(sizeof(int3) * vw) +
           align256(sizeof(int2) * cw) + kTileClasses * 256 + align256(sizeof(int2) * mw) +
           align256(sizeof(double) * 3 * cw) + align256(sizeof(void*) * nsys) + align256(sizeof(int2) * cw) +
           align256(sizeof(RefCtl) * cw) + mpa_batch_scratch_bytes(nsys, inner.data(), MPA_C64);
}

}  // namespace

extern "C" {

size_t mpa_refine_state_bytes(int n, int nrhs) { return own_bytes(n, nrhs) + mpa_system_state_bytes(n, nrhs, MPA_C64); }

size_t mpa_refine_scratch_bytes(int nsys, const mpa_refine_system* systems) {
    return refine_scratch_bytes(nsys, systems);
}

int mpa_solve_refined_batched(int nsys, const mpa_refine_system* systems, void* scratch, int* outer_out,
                              int* inner_out, double* resid_out, int* status_out, void* stream) {
    if (!valid(nsys, systems) || !scratch) {
        mpa_set_error("solve_refined_batched: bad argument");
        return MPA_ERR_ARG;
    }
    return refine_batch(nsys, systems, scratch, outer_out, inner_out, resid_out, status_out, (hipStream_t)stream);
}

int mpa_refine_systems_start(int nsys, const mpa_refine_system* systems, void* scratch, void* stream) {
    if (!valid(nsys, systems) || !scratch) {
        mpa_set_error("refine_systems_start: bad argument");
        return MPA_ERR_ARG;
    }
    return refine_iterate(nsys, systems, scratch, true, 0, nullptr, nullptr, nullptr, nullptr, nullptr,
                          (hipStream_t)stream);
}

int mpa_refine_systems_advance(int nsys, const mpa_refine_system* systems, void* scratch, int iterations,
                               int* done_out, int* outer_out, int* inner_out, double* resid_out, int* status_out,
                               void* stream) {
    if (!valid(nsys, systems) || !scratch || iterations < 0) {
        mpa_set_error("refine_systems_advance: bad argument");
        return MPA_ERR_ARG;
    }
    return refine_iterate(nsys, systems, scratch, false, iterations, done_out, outer_out, inner_out, resid_out,
                          status_out, (hipStream_t)stream);
}

size_t mpa_refine_workspace_bytes(int n, int nrhs) {
    mpa_refine_system s{};
    s.n = n;
    s.nrhs = nrhs;
    return mpa_refine_state_bytes(n, nrhs) + refine_scratch_bytes(1, &s);
}

int mpa_solve_refined(int n, int nrhs, const void* A32, int add_identity32, const void* A64, int add_identity64,
                      const void* b, void* x, double tol, double inner_tol, int maxiter, int max_refine,
                      void* workspace, int* outer_out, int* inner_out, double* resid_out, int* status_out,
                      void* stream) {
    if (add_identity32 != add_identity64) {
        mpa_set_error("solve_refined: A32 and A64 must describe the same system (add_identity differs)");
        return MPA_ERR_ARG;
    }
    mpa_refine_system s{};
    s.n = n;
    s.nrhs = nrhs;
    s.add_identity = add_identity32;
    s.A32 = A32;
    s.A64 = A64;
    s.b = b;
    s.x = x;
    s.state = workspace;
    s.tol = tol;
    s.inner_tol = inner_tol;
    s.maxiter = maxiter;
    s.max_refine = max_refine;
    if (!valid(1, &s)) {
        mpa_set_error("solve_refined: bad argument");
        return MPA_ERR_ARG;
    }
    void* scratch = (char*)workspace + mpa_refine_state_bytes(n, nrhs);
    return refine_batch(1, &s, scratch, outer_out, inner_out, resid_out, status_out, (hipStream_t)stream);
}

}  // extern "C"
