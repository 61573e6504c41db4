// Cluster coefficients and the cluster's per-multipole cross-sections.
//
// p_cluster[s] = sum_i of particle i's scattered-field coefficients
// p_scat[s][i] re-expanded about one origin: for every output mode (n, m)
// (n <= lmax_out) and input mode (v, u) (v <= lmax_in),
//   p_cluster[0][r] += a A + b B,  p_cluster[1][r] += a B + b A,
// with (a, b) = p_scat[s][i][:, r'] and (A, B) the incident-mode translation
// coefficients of rij = origin - r_i (j_p in place of h_p), and a particle
// exactly at the origin adding its coefficients directly
// (src/miepy/vsh/cluster_coefficients.py:14-53; scalar vsh_translation,
// cpp/src/vsh_translation.cpp:428-470).  The translation coefficients come
// from the compact table of L = max(lmax_out, lmax_in) (every (n, m; v, u)
// with n, v <= L); the per-particle Bessel-Legendre products are formed once
// per particle in LDS.  Particles are split over workgroups in a fixed way
// and the partial sums are added in workgroup order: deterministic.
//
// Cross-sections per multipole (cpp/src/flux.cpp:32-57): with
// f = 4 pi / k^2, Cscat[a][n-1] = f sum_m |p_cluster[a][r]|^2,
// Cext[a][n-1] = f sum_m Re(conj(p_src[a][r]) p_cluster[a][r]),
// Cabs = Cext - Cscat.

#include <cmath>

#include "translation.hpp"

using namespace mpa;

namespace {

constexpr int kCT = 256;
constexpr int kMaxChunks = 128;  // workgroups over particles, one partial sum each

int chunks_for(int N) { return N < kMaxChunks ? N : kMaxChunks; }

// mie (optional, [N][2][lmax_in]) scales the input first: p_scat = mie (.) x per
// (particle, polarization, n).  A particle is at the origin when the
// translation vector is exactly zero or shorter than origin_tol.
__global__ __launch_bounds__(kCT) void k_cluster_coef(int N, int L, int lmax_in, int lmax_out, int S,
                                                      const double* __restrict__ pos, double ox, double oy, double oz,
                                                      const double2* __restrict__ p_scat,
                                                      const double2* __restrict__ mie, double origin_tol, double k,
                                                      TabDev tab, double2* __restrict__ part) {
    __shared__ double2 HL[(2 * MPA_LMAX_MAX + 2) * (2 * MPA_LMAX_MAX + 3) / 2];
    __shared__ double2 eph[2 * MPA_LMAX_MAX + 1];
    const int rmax_in = lmax_to_rmax(lmax_in), rmax_out = lmax_to_rmax(lmax_out), rmaxL = lmax_to_rmax(L);
    const int pmax = 2 * L + 1;
    const int total = S * rmax_out;
    for (int u0 = 0; u0 < total; u0 += kCT) {
        const int u = u0 + threadIdx.x;
        const bool mine = u < total;
        const int s = mine ? u / rmax_out : 0, r = mine ? u - s * rmax_out : 0;
        double2 acc0 = make_double2(0, 0), acc1 = make_double2(0, 0);
        for (int i = blockIdx.x; i < N; i += gridDim.x) {
            const double dx = ox - pos[3 * i], dy = oy - pos[3 * i + 1], dz = oz - pos[3 * i + 2];
            const bool at_origin = (dx == 0.0 && dy == 0.0 && dz == 0.0) ||
                                   sqrt(dx * dx + dy * dy + dz * dz) < origin_tol;
            __syncthreads();  // the previous particle's HL is no longer read
            if (!at_origin && (int)threadIdx.x <= pmax) pair_hl(HL, eph, threadIdx.x, L, dx, dy, dz, k, 2);
            __syncthreads();
            if (!mine) continue;
            const double2* ps = p_scat + ((size_t)s * N + i) * 2 * rmax_in;
            const double2* mi = mie ? mie + (size_t)i * 2 * lmax_in : nullptr;
            if (at_origin) {
                if (r < rmax_in) {
                    const int v = mode_n(r);
                    acc0 = cadd(acc0, mi ? cmul(mi[v - 1], ps[r]) : ps[r]);
                    acc1 = cadd(acc1, mi ? cmul(mi[lmax_in + v - 1], ps[rmax_in + r]) : ps[rmax_in + r]);
                }
                continue;
            }
            for (int rp = 0; rp < rmax_in; rp++) {
                double2 xa, xb;
                tuple_xab(tab, r * rmaxL + rp, HL, eph, xa, xb);
                double2 a = ps[rp], b = ps[rmax_in + rp];
                if (mi) {
                    const int v = mode_n(rp);
                    a = cmul(mi[v - 1], a);
                    b = cmul(mi[lmax_in + v - 1], b);
                }
                acc0 = cadd(acc0, cadd(cmul(a, xa), cmul(b, xb)));
                acc1 = cadd(acc1, cadd(cmul(a, xb), cmul(b, xa)));
            }
        }
        if (mine) {
            double2* out = part + ((size_t)blockIdx.x * S + s) * 2 * rmax_out;
            out[r] = acc0;
            out[rmax_out + r] = acc1;
        }
    }
}

// (A, B) = the translation coefficients of tuples (m, n, u, v) for the
// translation vectors (rad, theta, phi)[g] and one VSH mode: the scalar
// vsh_translation of cpp/src/vsh_translation.cpp:428-470 (and its numpy
// form over arrays of vectors, :366-426; the row order m enters negated
// there, as in the table).  One workgroup per vector: threads 0..pmax form
// the products of the radial function and P_p^mu in LDS, then one thread per
// tuple sums its table terms.  mnuv [count][4]; out [g][count][2].
__global__ __launch_bounds__(kCT) void k_vsh_translation(int count, const int* __restrict__ mnuv, int L,
                                                         const double* __restrict__ rad,
                                                         const double* __restrict__ theta,
                                                         const double* __restrict__ phi, double k, int zkind,
                                                         TabDev tab, double2* __restrict__ out) {
    __shared__ double2 HL[(2 * MPA_LMAX_MAX + 2) * (2 * MPA_LMAX_MAX + 3) / 2];
    __shared__ double2 eph[2 * MPA_LMAX_MAX + 1];
    const int g = blockIdx.x;
    const int rmaxL = lmax_to_rmax(L);
    const double st = sin(theta[g]);
    const double dx = rad[g] * st * cos(phi[g]), dy = rad[g] * st * sin(phi[g]), dz = rad[g] * cos(theta[g]);
    if ((int)threadIdx.x <= 2 * L + 1) pair_hl(HL, eph, threadIdx.x, L, dx, dy, dz, k, zkind);
    __syncthreads();
    for (int q = threadIdx.x; q < count; q += kCT) {
        const int m = mnuv[4 * q], n = mnuv[4 * q + 1], u = mnuv[4 * q + 2], v = mnuv[4 * q + 3];
        double2 xa, xb;
        tuple_xab(tab, mode_r(n, m) * rmaxL + mode_r(v, u), HL, eph, xa, xb);
        out[2 * ((size_t)g * count + q)] = xa;
        out[2 * ((size_t)g * count + q) + 1] = xb;
    }
}

// out[e] = sum over chunks c (in order) of part[c][e], e < count; with
// rows > 0 the [count / rows][rows] result is stored transposed, [rows][count / rows]
__global__ __launch_bounds__(kCT) void k_cluster_sum(const double2* __restrict__ part, int chunks, int count,
                                                     int rows, double2* __restrict__ out) {
    const int e = blockIdx.x * kCT + threadIdx.x;
    if (e >= count) return;
    double2 acc = part[e];
    for (int c = 1; c < chunks; c++) acc = cadd(acc, part[(size_t)c * count + e]);
    if (rows > 0) {
        const int col = e / rows, row = e - col * rows;
        out[(size_t)row * (count / rows) + col] = acc;
    } else {
        out[e] = acc;
    }
}

// The source matrix of a sphere cluster's T-matrix: origin mode (b, n, m)
// (n <= lmax_out) expanded about particle i in modes (a, v, u) (v <= lmax_p)
// with the incident-mode translation coefficients (A, B) of the tuple
// (u, v, m, n) and d = r_i - origin: A where a == b, B otherwise; a particle
// within 1e-15 of the origin takes the identity on the modes both sides hold
// (src/miepy/tmatrix/common.py:199-247).  Stored one origin mode per row,
// out[b * rmax_out + r][i][a][s] (each row a right-hand side of the cluster
// system).  One workgroup per particle; threads over (r, s).
__global__ __launch_bounds__(kCT) void k_cluster_source(int N, int L, int lmax_p, int lmax_out,
                                                        const double* __restrict__ pos, double ox, double oy,
                                                        double oz, double k, TabDev tab, double2* __restrict__ out) {
    __shared__ double2 HL[(2 * MPA_LMAX_MAX + 2) * (2 * MPA_LMAX_MAX + 3) / 2];
    __shared__ double2 eph[2 * MPA_LMAX_MAX + 1];
    const int i = blockIdx.x;
    const int rmax_p = lmax_to_rmax(lmax_p), rmax_out = lmax_to_rmax(lmax_out), rmaxL = lmax_to_rmax(L);
    const size_t n = (size_t)N * 2 * rmax_p;
    const double dx = pos[3 * i] - ox, dy = pos[3 * i + 1] - oy, dz = pos[3 * i + 2] - oz;
    const bool at_origin = sqrt(dx * dx + dy * dy + dz * dz) < 1e-15;
    if (!at_origin && (int)threadIdx.x <= 2 * L + 1) pair_hl(HL, eph, threadIdx.x, L, dx, dy, dz, k, 2);
    __syncthreads();
    for (int q = threadIdx.x; q < rmax_out * rmax_p; q += kCT) {
        const int r = q / rmax_p, s = q - r * rmax_p;
        double2 xa = make_double2(0, 0), xb = make_double2(0, 0);
        if (at_origin) {
            if (r == s) xa.x = 1.0;
        } else {
            tuple_xab(tab, s * rmaxL + r, HL, eph, xa, xb);
        }
        double2* o0 = out + (size_t)r * n + (size_t)i * 2 * rmax_p;               // origin polarization 0
        double2* o1 = out + (size_t)(rmax_out + r) * n + (size_t)i * 2 * rmax_p;  // origin polarization 1
        o0[s] = xa;
        o0[rmax_p + s] = xb;
        o1[s] = xb;
        o1[rmax_p + s] = xa;
    }
}

// C[s] = (Cscat, Cabs, Cext) [3][2][lmax]; one thread per (s, a, n)
__global__ __launch_bounds__(kCT) void k_cluster_xs(int lmax, int S, const double2* __restrict__ pc,
                                                    const double2* __restrict__ p0, double k, double* __restrict__ C) {
    const int u = blockIdx.x * kCT + threadIdx.x;
    if (u >= S * 2 * lmax) return;
    const int s = u / (2 * lmax), an = u - s * 2 * lmax, a = an / lmax, n = an - a * lmax + 1;
    const int rmax = lmax_to_rmax(lmax);
    const double factor = 4.0 * M_PI / (k * k);
    double cs = 0.0, ce = 0.0;
    for (int m = -n; m <= n; m++) {
        const size_t idx = (size_t)s * 2 * rmax + a * rmax + mode_r(n, m);
        const double2 c = pc[idx], q = p0[idx];
        cs += factor * (c.x * c.x + c.y * c.y);
        ce += factor * (q.x * c.x + q.y * c.y);
    }
    double* out = C + (size_t)s * 3 * 2 * lmax;
    out[an] = cs;
    out[2 * lmax + an] = ce - cs;
    out[4 * lmax + an] = ce;
}

}  // namespace

extern "C" {

size_t mpa_cluster_coefficients_workspace_bytes(int N, int lmax_out, int nrhs) {
    if (N < 1 || lmax_out < 1 || nrhs < 1) return 0;
    return sizeof(double2) * (size_t)chunks_for(N) * nrhs * 2 * lmax_to_rmax(lmax_out);
}

int mpa_cluster_coefficients(int N, int lmax_in, int lmax_out, int nrhs, const double* positions, double ox,
                             double oy, double oz, const void* p_scat, double k, void* p_cluster, void* workspace,
                             void* stream) {
    if (N < 1 || nrhs < 1 || lmax_in < 1 || lmax_out < 1 || lmax_in > MPA_LMAX_MAX || lmax_out > MPA_LMAX_MAX ||
        !positions || !p_scat || !p_cluster || !workspace || !(k > 0)) {
        mpa_set_error("cluster_coefficients: bad argument (1 <= lmax_in, lmax_out <= %d)", MPA_LMAX_MAX);
        return MPA_ERR_ARG;
    }
    const int L = lmax_in > lmax_out ? lmax_in : lmax_out;
    const TransTable* T = mpa_get_table(L);
    if (!T) return MPA_ERR_HIP;
    hipStream_t st = (hipStream_t)stream;
    const int chunks = chunks_for(N);
    const TabDev tab{T->cmeta, T->ccoef, T->clidx};
    hipLaunchKernelGGL(k_cluster_coef, dim3(chunks), dim3(kCT), 0, st, N, L, lmax_in, lmax_out, nrhs, positions, ox,
                       oy, oz, (const double2*)p_scat, (const double2*)nullptr, 0.0, k, tab, (double2*)workspace);
    const int count = nrhs * 2 * lmax_to_rmax(lmax_out);
    hipLaunchKernelGGL(k_cluster_sum, dim3((count + kCT - 1) / kCT), dim3(kCT), 0, st, (const double2*)workspace,
                       chunks, count, 0, (double2*)p_cluster);
    MPA_LAUNCH_CHECK();
    return MPA_OK;
}

int mpa_cluster_source(int N, int lmax_p, int lmax_out, const double* positions, double ox, double oy, double oz,
                       double k, void* out, void* stream) {
    if (N < 1 || N > 65535 || lmax_p < 1 || lmax_out < 1 || lmax_p > MPA_LMAX_MAX || lmax_out > MPA_LMAX_MAX ||
        !positions || !out || !(k > 0)) {
        mpa_set_error("cluster_source: bad argument (1 <= N <= 65535, 1 <= lmax_p, lmax_out <= %d)", MPA_LMAX_MAX);
        return MPA_ERR_ARG;
    }
    const int L = lmax_p > lmax_out ? lmax_p : lmax_out;
    const TransTable* T = mpa_get_table(L);
    if (!T) return MPA_ERR_HIP;
    const TabDev tab{T->cmeta, T->ccoef, T->clidx};
    hipLaunchKernelGGL(k_cluster_source, dim3(N), dim3(kCT), 0, (hipStream_t)stream, N, L, lmax_p, lmax_out,
                       positions, ox, oy, oz, k, tab, (double2*)out);
    MPA_LAUNCH_CHECK();
    return MPA_OK;
}

int mpa_cluster_tmatrix(int N, int lmax_p, int lmax_out, const double* positions, double ox, double oy, double oz,
                        const void* mie, const void* X, double k, void* tmatrix, void* workspace, void* stream) {
    if (N < 1 || lmax_p < 1 || lmax_out < 1 || lmax_p > MPA_LMAX_MAX || lmax_out > MPA_LMAX_MAX || !positions ||
        !mie || !X || !tmatrix || !workspace || !(k > 0)) {
        mpa_set_error("cluster_tmatrix: bad argument (1 <= lmax_p, lmax_out <= %d)", MPA_LMAX_MAX);
        return MPA_ERR_ARG;
    }
    const int L = lmax_p > lmax_out ? lmax_p : lmax_out;
    const TransTable* T = mpa_get_table(L);
    if (!T) return MPA_ERR_HIP;
    hipStream_t st = (hipStream_t)stream;
    const int chunks = chunks_for(N);
    const int rows = 2 * lmax_to_rmax(lmax_out);  // one right-hand side per origin mode
    const TabDev tab{T->cmeta, T->ccoef, T->clidx};
    hipLaunchKernelGGL(k_cluster_coef, dim3(chunks), dim3(kCT), 0, st, N, L, lmax_p, lmax_out, rows, positions, ox,
                       oy, oz, (const double2*)X, (const double2*)mie, 1e-15, k, tab, (double2*)workspace);
    const int count = rows * rows;
    hipLaunchKernelGGL(k_cluster_sum, dim3((count + kCT - 1) / kCT), dim3(kCT), 0, st, (const double2*)workspace,
                       chunks, count, rows, (double2*)tmatrix);
    MPA_LAUNCH_CHECK();
    return MPA_OK;
}

int mpa_vsh_translation(int count, const int* mnuv, int lmax, int nvec, const double* rad, const double* theta,
                        const double* phi, double k, int mode, void* out, void* stream) {
    if (count < 1 || nvec < 1 || nvec > 65535 || !mnuv || !rad || !theta || !phi || !out || lmax < 1 ||
        lmax > MPA_LMAX_MAX || !(k > 0) || mode < 0 || mode > 3) {
        mpa_set_error("vsh_translation: bad argument (1 <= lmax <= %d, mode 0-3)", MPA_LMAX_MAX);
        return MPA_ERR_ARG;
    }
    const TransTable* T = mpa_get_table(lmax);
    if (!T) return MPA_ERR_HIP;
    const TabDev tab{T->cmeta, T->ccoef, T->clidx};
    const int zkind = mode == 0 ? 0 : (mode == 1 ? 1 : 2);  // outgoing, ingoing, incident / interior
    hipLaunchKernelGGL(k_vsh_translation, dim3(nvec), dim3(kCT), 0, (hipStream_t)stream, count, mnuv, lmax, rad,
                       theta, phi, k, zkind, tab, (double2*)out);
    MPA_LAUNCH_CHECK();
    return MPA_OK;
}

int mpa_cluster_cross_sections(int lmax, int nrhs, const void* p_cluster, const void* p_src, double k, double* C,
                               void* stream) {
    if (lmax < 1 || nrhs < 1 || !p_cluster || !p_src || !C || !(k > 0)) {
        mpa_set_error("cluster_cross_sections: bad argument");
        return MPA_ERR_ARG;
    }
    const int count = nrhs * 2 * lmax;
    hipLaunchKernelGGL(k_cluster_xs, dim3((count + kCT - 1) / kCT), dim3(kCT), 0, (hipStream_t)stream, lmax, nrhs,
                       (const double2*)p_cluster, (const double2*)p_src, k, C);
    MPA_LAUNCH_CHECK();
    return MPA_OK;
}

}  // extern "C"
