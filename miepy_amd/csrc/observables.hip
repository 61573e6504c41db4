// Observables of a solved cluster: scattered/interior coefficients,
// per-particle cross-sections, force and torque, cluster totals.
//
// Replaces, for all particles and sources in one launch:
//   p_scat = p_inc * mie_scat[:, :, n_idx], p_int = p_inc * mie_int[:, ::-1, n_idx]
//       (src/miepy/sphere_cluster.py:715-717)
//   particle_cross_sections           (cpp/src/flux.cpp:5-30)
//   force / force_all                 (cpp/src/forces.cpp:7-86, 139-154)
//   torque / torque_all               (cpp/src/forces.cpp:88-137, 156-171)
//   cross_sections() cluster sums     (src/miepy/sphere_cluster.py:442-454)
// All in FP64.  One thread per (particle, source); totals are reduced in a
// fixed order by one workgroup per source.

#include <algorithm>

#include "common.hpp"

namespace {

constexpr int kThreads = 256;
// GSL_CONST_MKSA_VACUUM_PERMITTIVITY, the constant compiled into the
// reference's force/torque (cpp/src/forces.cpp:12)
constexpr double kEps0 = 8.854187817e-12;
constexpr double kPi = 3.14159265358979323846;

__device__ inline double re_mulc(double2 a, double2 b) {  // Re(a * conj(b))
    return a.x * b.x + a.y * b.y;
}
// a * conj(b)
__device__ inline double2 mulc2(double2 a, double2 b) { return make_double2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y); }

// cross-sections per multipole of one particle (flux.cpp:5-30); Cp [3][2][lmax]
__device__ void particle_xs(int lmax, const double2* ps_, const double2* pi_, const double2* pr_, double k,
                            double* Cp) {
    const int rmax = lmax_to_rmax(lmax);
    const double factor = 4 * kPi / (k * k);
    for (int a = 0; a < 2; a++)
        for (int n = 1; n <= lmax; n++) {
            double cabs = 0, cext = 0;
            for (int m = -n; m <= n; m++) {
                const int r = a * rmax + mode_r(n, m);
                const double2 sc = ps_[r];
                cabs += factor * (re_mulc(sc, pi_[r]) - (sc.x * sc.x + sc.y * sc.y));
                cext += factor * re_mulc(sc, pr_[r]);
            }
            Cp[0 * 2 * lmax + a * lmax + n - 1] = cext - cabs;
            Cp[1 * 2 * lmax + a * lmax + n - 1] = cabs;
            Cp[2 * 2 * lmax + a * lmax + n - 1] = cext;
        }
}

// 2 P[r] conj(Q[r1]) - P[r] conj(Qi[r1]) - Pi[r] conj(Q[r1])
__device__ inline double2 fterm(int r, int r1, const double2* P, const double2* Pi, const double2* Q,
                                const double2* Qi) {
    double2 t = cscale(mulc2(P[r], Q[r1]), 2.0);
    t = csub(t, mulc2(P[r], Qi[r1]));
    return csub(t, mulc2(Pi[r], Q[r1]));
}

// force on one particle (forces.cpp:7-86); ps_/pi_ are [2 rmax]
__device__ void particle_force(int lmax, const double2* ps_, const double2* pi_, double k, double eps_b, double* Fo) {
    const int rmax = lmax_to_rmax(lmax);
    const double2* p = ps_;
    const double2* q = ps_ + rmax;
    const double2* pi = pi_;
    const double2* qi = pi_ + rmax;
    double2 Fxy = make_double2(0, 0), Fz = make_double2(0, 0);
    const double Axy = kPi / (k * k) * kEps0 * eps_b;
    const double Az = -2 * Axy;
    for (int n = 1; n <= lmax; n++) {
        for (int m = -n; m <= n; m++) {
            const int r = n * n + n + m - 1;
            if (m != n) {
                const double f = Axy * sqrt((double)((n + m + 1) * (n - m))) / (double)(n * (n + 1));
                const int r1 = r + 1;
                Fxy = cadd(Fxy, cscale(cadd(fterm(r, r1, p, pi, q, qi), fterm(r, r1, q, qi, p, pi)), f));
            }
            Fz = cadd(Fz, cscale(fterm(r, r, p, pi, q, qi), Az * m / (double)(n * (n + 1))));
            if (n < lmax) {
                const double den = (double)((2 * n + 3) * (2 * n + 1));
                double f = -Axy * sqrt((double)((n + m + 2) * (n + m + 1) * n * (n + 2)) / den) / (n + 1);
                int r1 = (n + 1) * (n + 1) + (n + 1) + m;
                Fxy = cadd(Fxy, cscale(cadd(fterm(r, r1, p, pi, p, pi), fterm(r, r1, q, qi, q, qi)), f));
                f = Axy * sqrt((double)((n - m + 1) * (n - m + 2) * n * (n + 2)) / den) / (n + 1);
                r1 = (n + 1) * (n + 1) + (n + 1) + m - 2;
                Fxy = cadd(Fxy, cscale(cadd(fterm(r1, r, p, pi, p, pi), fterm(r1, r, q, qi, q, qi)), f));
                f = Az * sqrt((double)((n - m + 1) * (n + m + 1) * n * (n + 2)) / den) / (n + 1);
                r1 = (n + 1) * (n + 1) + (n + 1) + m - 1;
                Fz = cadd(Fz, cscale(cadd(fterm(r1, r, p, pi, p, pi), fterm(r1, r, q, qi, q, qi)), f));
            }
        }
    }
    Fo[0] = Fxy.x;
    Fo[1] = Fxy.y;
    Fo[2] = Fz.x;
}

// torque on one particle (forces.cpp:88-137)
__device__ void particle_torque(int lmax, const double2* ps_, const double2* pi_, double k, double eps_b, double* To) {
    const int rmax = lmax_to_rmax(lmax);
    const double2* p = ps_;
    const double2* q = ps_ + rmax;
    const double2* pi = pi_;
    const double2* qi = pi_ + rmax;
    const double A = -2 * kPi / (k * k * k) * kEps0 * eps_b;
    double T0 = 0, T1 = 0, T2 = 0;
    for (int n = 1; n <= lmax; n++) {
        for (int m = -n; m <= n; m++) {
            const int r = n * n + n + m - 1;
            if (m != n) {
                const double f = -A * sqrt((double)((n - m) * (n + m + 1)));
                const int r1 = r + 1;
                const double2 pp = cadd(mulc2(p[r], p[r1]), mulc2(q[r], q[r1]));
                const double2 x1 = cadd(cadd(mulc2(p[r1], pi[r]), mulc2(p[r], pi[r1])),
                                        cadd(mulc2(q[r1], qi[r]), mulc2(q[r], qi[r1])));
                const double2 y1 = cadd(csub(mulc2(p[r1], pi[r]), mulc2(p[r], pi[r1])),
                                        csub(mulc2(q[r1], qi[r]), mulc2(q[r], qi[r1])));
                T0 += f * (pp.x - 0.5 * x1.x);
                T1 += f * (pp.y + 0.5 * y1.y);
            }
            T2 += A * m * ((p[r].x * p[r].x + p[r].y * p[r].y) + (q[r].x * q[r].x + q[r].y * q[r].y) -
                           (re_mulc(p[r], pi[r]) + re_mulc(q[r], qi[r])));
        }
    }
    To[0] = T0;
    To[1] = T1;
    To[2] = T2;
}

// fused, for particle i and source s: p_scat / p_int from p_inc, then cross-sections, force, torque
__device__ void observables_one(int N, int lmax, int s, int i, const double2* __restrict__ p_inc,
                                const double2* __restrict__ p_src, const double2* __restrict__ mie_scat,
                                const double2* __restrict__ mie_int, double k, double eps_b,
                                double2* __restrict__ p_scat, double2* __restrict__ p_int, double* __restrict__ C_part,
                                double* __restrict__ F, double* __restrict__ Tq) {
    const int rmax = lmax_to_rmax(lmax);
    const size_t base = ((size_t)s * N + i) * 2 * rmax;
    for (int n = 1; n <= lmax; n++)
        for (int m = -n; m <= n; m++) {
            const int r = mode_r(n, m);
            for (int a = 0; a < 2; a++) {
                const double2 pv = p_inc[base + a * rmax + r];
                p_scat[base + a * rmax + r] = cmul(pv, mie_scat[((size_t)i * 2 + a) * lmax + n - 1]);
                if (p_int) p_int[base + a * rmax + r] = cmul(pv, mie_int[((size_t)i * 2 + (1 - a)) * lmax + n - 1]);
            }
        }
    particle_xs(lmax, p_scat + base, p_inc + base, p_src + base, k, C_part + ((size_t)s * N + i) * 6 * lmax);
    particle_force(lmax, p_scat + base, p_inc + base, k, eps_b, F + ((size_t)s * N + i) * 3);
    if (Tq) particle_torque(lmax, p_scat + base, p_inc + base, k, eps_b, Tq + ((size_t)s * N + i) * 3);
}

__global__ __launch_bounds__(kThreads) void k_observables(int N, int lmax, int S, const double2* __restrict__ p_inc,
                                                          const double2* __restrict__ p_src,
                                                          const double2* __restrict__ mie_scat,
                                                          const double2* __restrict__ mie_int, double k, double eps_b,
                                                          double2* __restrict__ p_scat, double2* __restrict__ p_int,
                                                          double* __restrict__ C_part, double* __restrict__ F,
                                                          double* __restrict__ Tq) {
    const int idx = blockIdx.x * kThreads + threadIdx.x;
    if (idx >= N * S) return;
    const int s = idx / N, i = idx - s * N;
    observables_one(N, lmax, s, i, p_inc, p_src, mie_scat, mie_int, k, eps_b, p_scat, p_int, C_part, F, Tq);
}

// ---- many clusters per launch: descriptors travel by value in the kernel
// arguments (up to kObsBatchMax clusters per launch), work ids (source,
// particle) dealt to the clusters in order; each cluster's arithmetic is
// that of its own launch.
constexpr int kObsBatchMax = 32;

struct ObsBatch {
    int count;
    int first[kObsBatchMax + 1];  // prefix of each cluster's workgroups
    mpa_obs_cluster c[kObsBatchMax];
};

__device__ __forceinline__ int obs_cluster_of(const ObsBatch& B, int v) {
    int q = 0;
    while (q + 1 < B.count && v >= B.first[q + 1]) q++;
    return q;
}

__global__ __launch_bounds__(kThreads) void k_observables_b(ObsBatch B) {
    const int q = obs_cluster_of(B, blockIdx.x);
    const mpa_obs_cluster& C = B.c[q];
    const int idx = (blockIdx.x - B.first[q]) * kThreads + threadIdx.x;
    if (idx >= C.N * C.nsrc) return;
    const int s = idx / C.N, i = idx - s * C.N;
    observables_one(C.N, C.lmax, s, i, (const double2*)C.p_inc, (const double2*)C.p_src,
                    (const double2*)C.mie_scat, (const double2*)C.mie_int, C.k, C.eps_b, (double2*)C.p_scat,
                    (double2*)C.p_int, C.C_part, C.F, C.T);
}

// force_all / torque_all with explicit p_scat and p_inc ([M, 2 rmax] rows)
__global__ __launch_bounds__(kThreads) void k_force_torque(int M, int lmax, const double2* __restrict__ p_scat,
                                                           const double2* __restrict__ p_inc, double k, double eps_b,
                                                           double* __restrict__ F, double* __restrict__ Tq) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= M) return;
    const size_t base = (size_t)i * 2 * lmax_to_rmax(lmax);
    if (F) particle_force(lmax, p_scat + base, p_inc + base, k, eps_b, F + (size_t)i * 3);
    if (Tq) particle_torque(lmax, p_scat + base, p_inc + base, k, eps_b, Tq + (size_t)i * 3);
}

// particle_cross_sections for M particles ([M, 2 rmax] rows) -> [M, 3, 2, lmax]
__global__ __launch_bounds__(kThreads) void k_cross_sections(int M, int lmax, const double2* __restrict__ p_scat,
                                                             const double2* __restrict__ p_inc,
                                                             const double2* __restrict__ p_src, double k,
                                                             double* __restrict__ C_part) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= M) return;
    const size_t base = (size_t)i * 2 * lmax_to_rmax(lmax);
    particle_xs(lmax, p_scat + base, p_inc + base, p_src + base, k, C_part + (size_t)i * 6 * lmax);
}

// cluster totals per source: sum over particles (fixed order) of the
// per-particle sums over multipoles; out [S][3] (scat, abs, ext)
__device__ void totals_one(int N, int lmax, int s, const double* __restrict__ C_part, double* __restrict__ C_total) {
    const int L = 2 * lmax;
    double acc[3] = {0, 0, 0};
    for (int i = threadIdx.x; i < N; i += kThreads) {
        const double* Cp = C_part + ((size_t)s * N + i) * 3 * L;
        for (int c = 0; c < 3; c++) {
            double t = 0;
            for (int e = 0; e < L; e++) t += Cp[c * L + e];
            acc[c] += t;
        }
    }
    __shared__ double red[3][kThreads];
    for (int c = 0; c < 3; c++) red[c][threadIdx.x] = acc[c];
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w)
            for (int c = 0; c < 3; c++) red[c][threadIdx.x] += red[c][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int c = 0; c < 3; c++) C_total[s * 3 + c] = red[c][0];
}

__global__ __launch_bounds__(kThreads) void k_totals(int N, int lmax, const double* __restrict__ C_part,
                                                     double* __restrict__ C_total) {
    totals_one(N, lmax, blockIdx.x, C_part, C_total);
}

// one workgroup per (cluster, source) of the batch with C_total set; first[] counts sources
__global__ __launch_bounds__(kThreads) void k_totals_b(ObsBatch B) {
    const int q = obs_cluster_of(B, blockIdx.x);
    const mpa_obs_cluster& C = B.c[q];
    totals_one(C.N, C.lmax, blockIdx.x - B.first[q], C.C_part, C.C_total);
}

// y_j = Tm_j x_j for every particle and right-hand side (block-diagonal
// product): p_scat = T p_inc of cpp's `cluster` path (src/miepy/cluster.py:657,
// einsum "naibj,nbj->nai").  One thread per output entry; Tm_j rows are read
// through the caches.
__global__ __launch_bounds__(256) void k_tmatrix_apply(int N, int bs, int S, const double2* __restrict__ tm,
                                                       const double2* __restrict__ x, double2* __restrict__ y) {
    const size_t n = (size_t)bs * N;
    const size_t total = (size_t)S * n;
    for (size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x; idx < total; idx += (size_t)gridDim.x * 256) {
        const size_t s = idx / n, rest = idx - s * n;
        const size_t j = rest / bs, e = rest - j * bs;
        const double2* trow = tm + (j * bs + e) * bs;
        const double2* xj = x + s * n + j * bs;
        double2 acc = make_double2(0.0, 0.0);
        for (int c = 0; c < bs; c++) cfma(acc, trow[c], xj[c]);
        y[idx] = acc;
    }
}

}  // namespace

int mpa_internal_tmatrix_apply(int N, int lmax, int nrhs, const double* tmatrix, const void* x, void* y,
                               hipStream_t st) {
    const int bs = 2 * lmax_to_rmax(lmax);
    const size_t total = (size_t)nrhs * bs * N;
    const unsigned blocks = (unsigned)std::min<size_t>(4096, (total + 255) / 256);
    hipLaunchKernelGGL(k_tmatrix_apply, dim3(blocks), dim3(256), 0, st, N, bs, nrhs, (const double2*)tmatrix,
                       (const double2*)x, (double2*)y);
    MPA_LAUNCH_CHECK();
    return MPA_OK;
}

extern "C" int mpa_tmatrix_apply(int N, int lmax, int nrhs, const double* tmatrix, const void* x, void* y,
                                 void* stream) {
    if (N < 1 || nrhs < 1 || lmax < 1 || lmax > MPA_LMAX_MAX || !tmatrix || !x || !y || x == y) {
        mpa_set_error("tmatrix_apply: bad argument");
        return MPA_ERR_ARG;
    }
    return mpa_internal_tmatrix_apply(N, lmax, nrhs, tmatrix, x, y, (hipStream_t)stream);
}

extern "C" int mpa_cluster_observables(int N, int lmax, int nsrc, const double* p_inc, const double* p_src,
                                       const double* mie_scat, const double* mie_int, double k, double eps_b,
                                       double* p_scat, double* p_int, double* C_part, double* C_total, double* F,
                                       double* T, void* stream) {
    if (N < 1 || nsrc < 1 || lmax < 1 || lmax > MPA_LMAX_MAX || !p_inc || !p_src || !mie_scat || !p_scat || !C_part ||
        !F || !(k > 0) || (p_int && !mie_int)) {
        mpa_set_error("cluster_observables: bad argument");
        return MPA_ERR_ARG;
    }
    hipStream_t st = (hipStream_t)stream;
    const int total = N * nsrc;
    hipLaunchKernelGGL(k_observables, dim3((total + kThreads - 1) / kThreads), dim3(kThreads), 0, st, N, lmax, nsrc,
                       (const double2*)p_inc, (const double2*)p_src, (const double2*)mie_scat,
                       (const double2*)mie_int, k, eps_b, (double2*)p_scat, (double2*)p_int, C_part, F, T);
    if (C_total) hipLaunchKernelGGL(k_totals, dim3(nsrc), dim3(kThreads), 0, st, N, lmax, C_part, C_total);
    MPA_LAUNCH_CHECK();
    return MPA_OK;
}

extern "C" int mpa_cluster_observables_batched(int count, const mpa_obs_cluster* clusters, void* stream) {
    if (count < 0 || (count > 0 && !clusters)) {
        mpa_set_error("cluster_observables_batched: bad argument");
        return MPA_ERR_ARG;
    }
    for (int q = 0; q < count; q++) {
        const mpa_obs_cluster& C = clusters[q];
        if (C.N < 1 || C.nsrc < 1 || C.lmax < 1 || C.lmax > MPA_LMAX_MAX || !C.p_inc || !C.p_src || !C.mie_scat ||
            !C.p_scat || !C.C_part || !C.F || !(C.k > 0) || (C.p_int && !C.mie_int) ||
            (long long)C.N * C.nsrc > 0x7fffffffLL - kThreads) {
            mpa_set_error("cluster_observables_batched: bad cluster %d", q);
            return MPA_ERR_ARG;
        }
    }
    hipStream_t st = (hipStream_t)stream;
    for (int b = 0; b < count; b += kObsBatchMax) {
        const int cnt = std::min(kObsBatchMax, count - b);
        ObsBatch Bo{}, Bt{};
        Bo.count = cnt;
        long long wg = 0;
        for (int q = 0; q < cnt; q++) {
            const mpa_obs_cluster& C = clusters[b + q];
            Bo.c[q] = C;
            Bo.first[q + 1] = Bo.first[q] + (C.N * C.nsrc + kThreads - 1) / kThreads;
            wg += (C.N * C.nsrc + kThreads - 1) / kThreads;
            if (C.C_total) {
                Bt.c[Bt.count] = C;
                Bt.first[Bt.count + 1] = Bt.first[Bt.count] + C.nsrc;
                Bt.count++;
            }
        }
        if (wg > 0x7fffffffLL) {
            mpa_set_error("cluster_observables_batched: batch too large for one launch");
            return MPA_ERR_ARG;
        }
        hipLaunchKernelGGL(k_observables_b, dim3((unsigned)Bo.first[cnt]), dim3(kThreads), 0, st, Bo);
        if (Bt.count) hipLaunchKernelGGL(k_totals_b, dim3((unsigned)Bt.first[Bt.count]), dim3(kThreads), 0, st, Bt);
        MPA_LAUNCH_CHECK();
    }
    return MPA_OK;
}

extern "C" int mpa_force_torque(int M, int lmax, const double* p_scat, const double* p_inc, double k, double eps_b,
                                double* F, double* T, void* stream) {
    if (M < 1 || lmax < 1 || lmax > MPA_LMAX_MAX || !p_scat || !p_inc || (!F && !T) || !(k > 0)) {
        mpa_set_error("force_torque: bad argument");
        return MPA_ERR_ARG;
    }
    hipLaunchKernelGGL(k_force_torque, dim3((M + kThreads - 1) / kThreads), dim3(kThreads), 0, (hipStream_t)stream, M,
                       lmax, (const double2*)p_scat, (const double2*)p_inc, k, eps_b, F, T);
    MPA_LAUNCH_CHECK();
    return MPA_OK;
}

extern "C" int mpa_particle_cross_sections(int M, int lmax, const double* p_scat, const double* p_inc,
                                           const double* p_src, double k, double* C_part, void* stream) {
    if (M < 1 || lmax < 1 || lmax > MPA_LMAX_MAX || !p_scat || !p_inc || !p_src || !C_part || !(k > 0)) {
        mpa_set_error("particle_cross_sections: bad argument");
        return MPA_ERR_ARG;
    }
    hipLaunchKernelGGL(k_cross_sections, dim3((M + kThreads - 1) / kThreads), dim3(kThreads), 0, (hipStream_t)stream,
                       M, lmax, (const double2*)p_scat, (const double2*)p_inc, (const double2*)p_src, k, C_part);
    MPA_LAUNCH_CHECK();
    return MPA_OK;
}
