// Per-wavelength inputs of the cluster solve on the device: Mie coefficients
// of every sphere and the plane-wave source coefficients p_src of every
// particle.
//
// Mie coefficients (src/miepy/mie_single/mie_sphere.py:16-76): with
// x = k a, m = sqrt(eps / eps_b), mt = m mu_b / mu, z = m x and the
// Riccati-Bessel functions psi_n = x j_n, xi_n = x h_n^(1)
// (src/miepy/special_functions.py:127-153),
//   a = (mt psi(z) psi'(x) - psi(x) psi'(z)) / (mt psi(z) xi'(x) - xi(x) psi'(z))
//   b = (psi(z) psi'(x) - mt psi(x) psi'(z)) / (psi(z) xi'(x) - mt xi(x) psi'(z))
//   c = m (psi(x) xi'(x) - xi(x) psi'(x)) / (psi(z) xi'(x) - mt xi(x) psi'(z))
//   d = m (psi(x) xi'(x) - xi(x) psi'(x)) / (mt psi(z) xi'(x) - xi(x) psi'(z))
// and a = psi'(x)/xi'(x), b = psi(x)/xi(x), c = d = 0 for a perfect conductor.
// The reference evaluates j_n of the complex argument with scipy; here psi_n
// of a real or complex argument comes from the upward recurrence while
// n <= |z| (stable there) and above |z| from the logarithmic derivative
// D_n = psi_n'/psi_n, by downward recurrence D_{n-1} = n/z - 1/(D_n + n/z)
// started well above max(lmax, |z|), as psi_n = psi_{n-1} / (D_n + n/z);
// x y_n by its upward recurrence (stable: y_n grows); derivatives
// f_n' = f_{n-1} - (n/x) f_n (D_n psi_n above |z|).  Checked against
// 40-digit values: within 2e-11 relative over x = 0.01-40 (the scipy-based
// reference formula is within 1.4e-10 there).
// One thread per particle, FP64.
//
// Plane waves (src/miepy/sources/plane_waves.py:126-156): p_src[s][i][a][r] =
// sum over the waves w of source s of e^{i (k r_i . khat_w + phase_w)}
// T_w[a][r], where the host folds amplitude, |E_mn|, pi/tau at the wave's
// theta, the polarization and e^{-i m phi} into the per-wave table T_w
// (rmax values per wave).

#include <cmath>

#include "common.hpp"

namespace {

constexpr int kIT = 256;
constexpr int kMieMaxL = 20;      // orders per particle
constexpr int kMieMaxStart = 400;  // downward-recurrence start bound (|m x| up to ~380)

__device__ inline double2 csqrt_principal(double2 w) {
    const double r = hypot(w.x, w.y);
    if (r == 0.0) return make_double2(0.0, w.y);
    if (w.x >= 0.0) {
        const double t = sqrt(0.5 * (r + w.x));
        return make_double2(t, w.y / (2.0 * t));
    }
    const double t = sqrt(0.5 * (r - w.x));
    return make_double2(fabs(w.y) / (2.0 * t), copysign(t, w.y));
}

__device__ inline double2 cdivs(double2 a, double2 b) {  // Smith's division (no overflow of |b|^2)
    if (fabs(b.x) >= fabs(b.y)) {
        const double r = b.y / b.x, d = b.x + b.y * r;
        return make_double2((a.x + a.y * r) / d, (a.y - a.x * r) / d);
    }
    const double r = b.x / b.y, d = b.y + b.x * r;
    return make_double2((a.x * r + a.y) / d, (a.y * r - a.x) / d);
}

__device__ inline double2 csin(double2 z) {
    double s, c;
    sincos(z.x, &s, &c);
    return make_double2(s * cosh(z.y), c * sinh(z.y));
}

// psi_n(z), psi_n'(z) for n = 0..L (complex z, z != 0).  The recurrence
// starts from max(kMieMaxL, |z|) + 16 whatever L is, so a particle's values
// do not depend on the lmax of the call (many clusters share one call).
__device__ inline void riccati_psi(double2 z, int L, double2* psi, double2* dpsi) {
    const double az = hypot(z.x, z.y);
    int nst = (int)(fmax((double)kMieMaxL, az) + 16.0);
    if (nst > kMieMaxStart) nst = kMieMaxStart;
    double2 D = make_double2(0.0, 0.0);
    double2 Dn[kMieMaxL + 1];
    for (int n = nst; n >= 1; n--) {
        const double2 nz = cdivs(make_double2((double)n, 0.0), z);
        if (n <= L) Dn[n] = D;
        D = csub(nz, cdivs(make_double2(1.0, 0.0), cadd(D, nz)));
    }
    // n <= |z|: the upward recurrence psi_n = (2n-1)/z psi_{n-1} - psi_{n-2}
    // is stable there (psi_n oscillates, and the ratio psi_{n-1}/psi_n would
    // pass through its zeros); above |z| the ratios of D_n
    int n0 = (int)floor(az);
    if (n0 > L) n0 = L;
    psi[0] = csin(z);
    if (n0 >= 1) {
        double sn, cs;
        sincos(z.x, &sn, &cs);
        const double2 cz = make_double2(cs * cosh(z.y), -sn * sinh(z.y));
        psi[1] = csub(cdivs(psi[0], z), cz);
    }
    for (int n = 2; n <= n0; n++)
        psi[n] = csub(cmul(cdivs(make_double2((double)(2 * n - 1), 0.0), z), psi[n - 1]), psi[n - 2]);
    for (int n = 1; n <= L; n++) {
        const double2 nz = cdivs(make_double2((double)n, 0.0), z);
        if (n <= n0) {
            dpsi[n] = csub(psi[n - 1], cmul(nz, psi[n]));
        } else {
            psi[n] = cdivs(psi[n - 1], cadd(Dn[n], nz));
            dpsi[n] = cmul(Dn[n], psi[n]);
        }
    }
}

// Mie coefficients of particle i: mie_scat[i][0][n-1] = a_n, [1] = b_n;
// mie_int[i][0] = c_n, [1] = d_n.  Complex inputs interleaved (re, im).
__global__ __launch_bounds__(kIT) void k_mie(int N, int lmax, const double* __restrict__ radius,
                                             const double2* __restrict__ eps, const double2* __restrict__ mu,
                                             const double* __restrict__ k, const double2* __restrict__ eps_b,
                                             const double2* __restrict__ mu_b, const int* __restrict__ conducting,
                                             double2* __restrict__ mie_scat, double2* __restrict__ mie_int) {
    const int i = blockIdx.x * kIT + threadIdx.x;
    if (i >= N) return;
    const double x = k[i] * radius[i];
    const double2 m = csqrt_principal(cdivs(eps[i], eps_b[i]));
    const double2 mt = cdivs(cmul(m, mu_b[i]), mu[i]);
    const bool cond = conducting && conducting[i] != 0;
    double2 px[kMieMaxL + 1], dpx[kMieMaxL + 1], pz[kMieMaxL + 1], dpz[kMieMaxL + 1];
    riccati_psi(make_double2(x, 0.0), lmax, px, dpx);
    if (!cond) riccati_psi(cmul(m, make_double2(x, 0.0)), lmax, pz, dpz);
    double s, c;
    sincos(x, &s, &c);
    double chi2 = -c, chi1 = -c / x - s;  // chi_n = x y_n for n = 0, 1
    double2* ms = mie_scat + (size_t)i * 2 * lmax;
    double2* mi = mie_int + (size_t)i * 2 * lmax;
    for (int n = 1; n <= lmax; n++) {
        double chi, chim1;
        if (n == 1) {
            chi = chi1;
            chim1 = chi2;
        } else {
            chi = (2 * n - 1) / x * chi1 - chi2;
            chim1 = chi1;
            chi2 = chi1;
            chi1 = chi;
        }
        const double2 xi = make_double2(px[n].x, px[n].y + chi);
        const double dchi = chim1 - n / x * chi;
        const double2 dxi = make_double2(dpx[n].x, dpx[n].y + dchi);
        double2 a, b, cc, d;
        if (cond) {
            a = cdivs(dpx[n], dxi);
            b = cdivs(px[n], xi);
            cc = d = make_double2(0.0, 0.0);
        } else {
            const double2 mpz = cmul(mt, pz[n]);
            const double2 mxz = cmul(mt, xi);
            const double2 den_a = csub(cmul(mpz, dxi), cmul(xi, dpz[n]));
            const double2 den_b = csub(cmul(pz[n], dxi), cmul(mxz, dpz[n]));
            a = cdivs(csub(cmul(mpz, dpx[n]), cmul(px[n], dpz[n])), den_a);
            b = cdivs(csub(cmul(pz[n], dpx[n]), cmul(cmul(mt, px[n]), dpz[n])), den_b);
            const double2 w = csub(cmul(cmul(m, px[n]), dxi), cmul(cmul(m, xi), dpx[n]));
            cc = cdivs(w, den_b);
            d = cdivs(w, den_a);
        }
        ms[n - 1] = a;
        ms[lmax + n - 1] = b;
        mi[n - 1] = cc;
        mi[lmax + n - 1] = d;
    }
}

// p_src[s][i][a][r] = sum_{w in [wbeg[s], wbeg[s+1])} e^{i (k r_i . khat_w + phase_w)} T_w[a][r];
// one thread per (s, i, a r).
__global__ __launch_bounds__(kIT) void k_plane_wave(int N, int rmax, int nsrc, const double* __restrict__ pos,
                                                    double k, const int* __restrict__ wbeg,
                                                    const double* __restrict__ khat, const double* __restrict__ phase,
                                                    const double2* __restrict__ table, double2* __restrict__ p_src) {
    const size_t u = (size_t)blockIdx.x * kIT + threadIdx.x;
    const size_t per = (size_t)2 * rmax;
    if (u >= (size_t)nsrc * N * per) return;
    const int ar = (int)(u % per);
    const size_t si = u / per;
    const int i = (int)(si % N), s = (int)(si / N);
    const double x = pos[3 * i], y = pos[3 * i + 1], z = pos[3 * i + 2];
    double2 acc = make_double2(0.0, 0.0);
    for (int w = wbeg[s]; w < wbeg[s + 1]; w++) {
        const double ph = k * (x * khat[3 * w] + y * khat[3 * w + 1] + z * khat[3 * w + 2]) + phase[w];
        double sn, cs;
        sincos(ph, &sn, &cs);
        acc = cadd(acc, cmul(make_double2(cs, sn), table[(size_t)w * per + ar]));
    }
    p_src[u] = acc;
}

}  // namespace

extern "C" {

int mpa_mie_coefficients(int N, int lmax, const double* radius, const void* eps, const void* mu, const double* k,
                         const void* eps_b, const void* mu_b, const int* conducting, void* mie_scat, void* mie_int,
                         void* stream) {
    if (N < 1 || lmax < 1 || lmax > kMieMaxL || !radius || !eps || !mu || !k || !eps_b || !mu_b || !mie_scat ||
        !mie_int) {
        mpa_set_error("mie_coefficients: bad argument (1 <= lmax <= %d)", kMieMaxL);
        return MPA_ERR_ARG;
    }
    hipLaunchKernelGGL(k_mie, dim3((N + kIT - 1) / kIT), dim3(kIT), 0, (hipStream_t)stream, N, lmax, radius,
                       (const double2*)eps, (const double2*)mu, k, (const double2*)eps_b, (const double2*)mu_b,
                       conducting, (double2*)mie_scat, (double2*)mie_int);
    MPA_LAUNCH_CHECK();
    return MPA_OK;
}

int mpa_plane_wave_structure(int N, int lmax, int nsrc, const double* positions, double k, const int* wave_begin,
                             const double* khat, const double* phase, const void* table, void* p_src, void* stream) {
    if (N < 1 || lmax < 1 || nsrc < 1 || !positions || !wave_begin || !khat || !phase || !table || !p_src) {
        mpa_set_error("plane_wave_structure: bad argument");
        return MPA_ERR_ARG;
    }
    const int rmax = lmax_to_rmax(lmax);
    const size_t total = (size_t)nsrc * N * 2 * rmax;
    hipLaunchKernelGGL(k_plane_wave, dim3((unsigned)((total + kIT - 1) / kIT)), dim3(kIT), 0, (hipStream_t)stream, N,
                       rmax, nsrc, positions, k, wave_begin, khat, phase, (const double2*)table, (double2*)p_src);
    MPA_LAUNCH_CHECK();
    return MPA_OK;
}

}  // extern "C"
