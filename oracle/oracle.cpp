// CPU FP64 oracle for the MiePy cluster solve.
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the shipped package (miepy_amd/) may
// load or call this library; only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py use it, and only as the checker.
//
// This is a restatement, in plain C++ with OpenMP and without GSL or Eigen,
// of the reference's C++ path for the cluster solve.  Every function cites the
// reference file:line it follows (paths relative to the reference root).
// Where the reference calls into GSL or Eigen (absent here), the published
// algorithm of that routine is restated and named.
//
// Complex arrays are interleaved (re, im) doubles, i.e. std::complex<double>.
// Matrices are row-major, like the reference's ComplexMatrix
// (cpp/src/vec.hpp:14).

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <vector>
#include <omp.h>

typedef std::complex<double> cd;

namespace {

const double PI = 3.14159265358979323846;  // cpp/src/vec.hpp:25

// cpp/src/special.cpp:15-17 (tgamma-based factorial)
double factorial(double n) { return std::tgamma(n + 1); }

double neg1pow(int k) { return (k % 2 == 0) ? 1.0 : -1.0; }

// i^p computed exactly (the reference evaluates pow(1i, p) through
// std::pow, which differs from this in the last bit only).
cd ipow(int p) {
    switch (((p % 4) + 4) % 4) {
        case 0: return cd(1, 0);
        case 1: return cd(0, 1);
        case 2: return cd(-1, 0);
        default: return cd(0, -1);
    }
}

int lmax_to_rmax(int lmax) { return lmax * (lmax + 2); }            // cpp/src/indices.cpp:8-10
int rmax_to_lmax(int rmax) { return -1 + int(std::sqrt(1.0 + rmax)); }  // cpp/src/indices.cpp:3-6

// ---------------------------------------------------------------------------
// Spherical Bessel / Hankel by upward recurrence
// ---------------------------------------------------------------------------

constexpr int kMaxDegH = 128;

// cpp/src/special.cpp:90-112 (spherical_hn_recursion): j_0 = sin z / z,
// y_0 = -cos z / z, j_1 = j_0/z + y_0, y_1 = y_0/z - j_0, then the upward
// three-term recurrence f_i = (2i-1)/z f_{i-1} - f_{i-2} for both.
void hn_recursion(int nmax, double z, cd* out) {
    double s = std::sin(z), c = std::cos(z);
    double jn[kMaxDegH], yn[kMaxDegH];
    jn[0] = s / z;
    yn[0] = -c / z;
    if (nmax >= 1) {
        jn[1] = jn[0] / z + yn[0];
        yn[1] = yn[0] / z - jn[0];
    }
    for (int i = 2; i <= nmax; i++) {
        jn[i] = (2 * i - 1) / z * jn[i - 1] - jn[i - 2];
        yn[i] = (2 * i - 1) / z * yn[i - 1] - yn[i - 2];
    }
    for (int i = 0; i <= nmax; i++) out[i] = cd(jn[i], yn[i]);
}

// cpp/src/special.cpp:54-88 (spherical_hn, single order, same recurrence)
cd spherical_hn(int n, double z) {
    cd v[kMaxDegH];
    hn_recursion(n, z, v);
    return v[n];
}

// cpp/src/special.cpp:20-44 (spherical_jn, complex argument, upward recurrence)
cd spherical_jn(int n, cd z) {
    cd s = std::sin(z), c = std::cos(z);
    cd jn_2p = s / z;
    cd jn = jn_2p;
    if (n > 0) {
        cd jn_p = jn_2p / z - c / z;
        jn = jn_p;
        for (int i = 2; i <= n; i++) {
            jn = double(2 * i - 1) / z * jn_p - jn_2p;
            jn_2p = jn_p;
            jn_p = jn;
        }
    }
    return jn;
}

// ---------------------------------------------------------------------------
// Associated Legendre functions, unnormalized, no Condon-Shortley phase
// ---------------------------------------------------------------------------

// Restates gsl_sf_legendre_array(GSL_SF_LEGENDRE_NONE, ...) as used at
// cpp/src/special.cpp:180-204: for m >= 0,
//   P_m^m = (2m-1)!! (1-x^2)^{m/2},  P_{m+1}^m = x (2m+1) P_m^m,
//   (l-m) P_l^m = (2l-1) x P_{l-1}^m - (l+m-1) P_{l-2}^m
// (the textbook recurrences; GSL's NONE normalization carries no (-1)^m).
// Output P[l * (lmax+1) + m] for 0 <= m <= l <= lmax (caller-owned buffer).
void legendre_nonneg(int lmax, double x, double* P) {
    const int w = lmax + 1;
    double sx = std::sqrt(std::max(0.0, (1 - x) * (1 + x)));
    double pmm = 1.0;
    for (int m = 0; m <= lmax; m++) {
        if (m > 0) pmm *= (2 * m - 1) * sx;
        P[m * w + m] = pmm;
        if (m + 1 <= lmax) P[(m + 1) * w + m] = x * (2 * m + 1) * pmm;
        for (int l = m + 2; l <= lmax; l++)
            P[l * w + m] = ((2 * l - 1) * x * P[(l - 1) * w + m] - (l + m - 1) * P[(l - 2) * w + m]) / (l - m);
    }
}

constexpr int kMaxDeg = 64;

// cpp/src/special.cpp:180-204: index n(n+2)-n+m, negative m through
// P_n^{-|m|} = (-1)^{|m|} (n-|m|)!/(n+|m|)! P_n^{|m|} (special.cpp:193).
void legendre_recursion(int nmax, double z, double* out) {
    std::vector<double> P((nmax + 1) * (nmax + 1));
    legendre_nonneg(nmax, z, P.data());
    for (int n = 0; n <= nmax; n++) {
        for (int m = -n; m <= n; m++) {
            double leg = P[n * (nmax + 1) + std::abs(m)];
            if (m < 0) leg *= neg1pow(m) * factorial(n + m) / factorial(n - m);
            out[n * (n + 2) - n + m] = leg;
        }
    }
}

// cpp/src/special.cpp:148-178 (associated_legendre, single value; n = 1
// closed forms at :150-156 are the same numbers as the recurrence).
double associated_legendre(int n, int m, double z) {
    if (std::abs(m) > n) return 0.0;
    if (n == 1) {
        if (m == 0) return z;
        if (m == 1) return std::sqrt(1 - z * z);
        if (m == -1) return -std::sqrt(1 - z * z) / 2;
    }
    std::vector<double> P((n + 1) * (n + 1));
    legendre_nonneg(n, z, P.data());
    double leg = P[n * (n + 1) + std::abs(m)];
    if (m < 0) leg *= neg1pow(m) * factorial(n + m) / factorial(n - m);
    return leg;
}

// d/dx P_n^{|m|}(x) for |x| < 1 from (x^2-1) P' = n x P_n - (n+m) P_{n-1},
// the relation gsl_sf_legendre_deriv_array evaluates (linear in P at fixed m,
// so it holds with or without the Condon-Shortley phase).
double legendre_deriv_nonneg(int n, int m, double x) {
    std::vector<double> P((n + 1) * (n + 1));
    legendre_nonneg(n, x, P.data());
    double pn = P[n * (n + 1) + m];
    double pn1 = (n - 1 >= m) ? P[(n - 1) * (n + 1) + m] : 0.0;
    return (n * x * pn - (n + m) * pn1) / (x * x - 1);
}

// cpp/src/special.cpp:209-231
double pi_func(int n, int m, double theta) {
    if (theta == 0) {
        if (m == 1) return n * (n + 1) / 2.0;
        else if (m == -1) return 1 / 2.0;
        else return 0;
    } else if (theta == PI) {
        if (m == 1) return neg1pow(n + 1) * n * (n + 1) / 2.0;
        else if (m == -1) return neg1pow(n + 1) / 2.0;
        else return 0;
    }
    double z = std::cos(theta);
    return m / std::sin(theta) * associated_legendre(n, m, z);
}

// cpp/src/special.cpp:233-270
double tau_func(int n, int m, double theta) {
    if (theta == 0) {
        if (m == 1) return n * (n + 1) / 2.0;
        else if (m == -1) return -1 / 2.0;
        else return 0;
    } else if (theta == PI) {
        if (m == 1) return neg1pow(n) * n * (n + 1) / 2.0;
        else if (m == -1) return -neg1pow(n) / 2.0;
        else return 0;
    }
    double z = std::cos(theta);
    double leg_p = -std::sin(theta) * legendre_deriv_nonneg(n, std::abs(m), z);
    if (m < 0) return neg1pow(m) * factorial(n + m) / factorial(n - m) * leg_p;
    return leg_p;
}

// ---------------------------------------------------------------------------
// Wigner 3j, Gaunt coefficients
// ---------------------------------------------------------------------------

// Restates gsl_sf_coupling_3j (called at cpp/src/special.cpp:272-274) by the
// Racah formula (Edmonds, "Angular Momentum in Quantum Mechanics", eq.
// 3.7.3), summed in long double.  Returns 0 when a selection rule fails.
double wigner_3j(int j1, int j2, int j3, int m1, int m2, int m3) {
    if (m1 + m2 + m3 != 0) return 0.0;
    if (j1 < 0 || j2 < 0 || j3 < 0) return 0.0;
    if (std::abs(m1) > j1 || std::abs(m2) > j2 || std::abs(m3) > j3) return 0.0;
    if (j3 < std::abs(j1 - j2) || j3 > j1 + j2) return 0.0;
    auto lf = [](int k) -> long double { return std::tgamma((long double)k + 1); };
    long double tri = lf(j1 + j2 - j3) * lf(j1 - j2 + j3) * lf(-j1 + j2 + j3) / lf(j1 + j2 + j3 + 1);
    long double pre = std::sqrt(tri * lf(j1 + m1) * lf(j1 - m1) * lf(j2 + m2) * lf(j2 - m2) *
                                lf(j3 + m3) * lf(j3 - m3));
    int kmin = std::max({0, j2 - j3 - m1, j1 - j3 + m2});
    int kmax = std::min({j1 + j2 - j3, j1 - m1, j2 + m2});
    long double sum = 0;
    for (int k = kmin; k <= kmax; k++) {
        long double den = lf(k) * lf(j3 - j2 + k + m1) * lf(j3 - j1 + k - m2) * lf(j1 + j2 - j3 - k) *
                          lf(j1 - k - m1) * lf(j2 - k + m2);
        sum += ((k % 2) ? -1.0L : 1.0L) / den;
    }
    int ph = j1 - j2 - m3;
    long double sign = (((ph % 2) + 2) % 2) ? -1.0L : 1.0L;
    return (double)(sign * pre * sum);
}

// cpp/src/special.cpp:276-285
double a_func(int m, int n, int u, int v, int p) {
    double numerator = factorial(n + m) * factorial(v + u) * factorial(p - m - u);
    double denominator = factorial(n - m) * factorial(v - u) * factorial(p + m + u);
    double factor = neg1pow(m + u) * (2 * p + 1) * std::sqrt(numerator / denominator);
    return factor * wigner_3j(n, v, p, 0, 0, 0) * wigner_3j(n, v, p, m, u, -m - u);
}

// cpp/src/special.cpp:287-296
double b_func(int m, int n, int u, int v, int p) {
    double numerator = factorial(n + m) * factorial(v + u) * factorial(p - m - u + 1);
    double denominator = factorial(n - m) * factorial(v - u) * factorial(p + m + u + 1);
    double factor = neg1pow(m + u) * (2 * p + 3) * std::sqrt(numerator / denominator);
    return factor * wigner_3j(n, v, p, 0, 0, 0) * wigner_3j(n, v, p + 1, m, u, -m - u);
}

// cpp/src/special.cpp:298-363 (wigner_3j_batch): values of
// W(j, j2, j3; m1, m2, m3) for j = jmin..jmax, seeded by two direct values and
// continued by the Schulten-Gordon three-term recurrence, with a direct value
// wherever the leading coefficient vanishes (|C| < 1e-30).
struct batch3j {
    int jmin = 0, jmax = -1;
    std::vector<double> values;
};

batch3j wigner_3j_batch(int j2, int j3, int m1, int m2, int m3) {
    batch3j r;
    if (m1 + m2 + m3 != 0) { r.jmin = 0; r.jmax = -1; return r; }
    r.jmin = std::max(std::abs(j2 - j3), std::abs(m1));
    r.jmax = j2 + j3;
    int size = r.jmax - r.jmin + 1;
    if (size <= 0) return r;
    r.values.resize(size);
    if (size <= 2) {
        for (int i = 0; i < size; i++) r.values[i] = wigner_3j(r.jmin + i, j2, j3, m1, m2, m3);
        return r;
    }
    r.values[0] = wigner_3j(r.jmin, j2, j3, m1, m2, m3);
    r.values[1] = wigner_3j(r.jmin + 1, j2, j3, m1, m2, m3);
    int d23 = j2 - j3;
    int s23p1 = j2 + j3 + 1;
    double j2j3 = double(j2) * (j2 + 1) - double(j3) * (j3 + 1);
    double mdiff = m3 - m2;
    for (int i = 1; i < size - 1; i++) {
        int j = r.jmin + i;
        double jd = j, jp1 = j + 1.0;
        double sj = std::sqrt(std::max(0.0, (jd * jd - double(d23) * d23) *
                                                (double(s23p1) * s23p1 - jd * jd) * (jd * jd - double(m1) * m1)));
        double sjp1 = std::sqrt(std::max(0.0, (jp1 * jp1 - double(d23) * d23) *
                                                  (double(s23p1) * s23p1 - jp1 * jp1) *
                                                  (jp1 * jp1 - double(m1) * m1)));
        double A = jp1 * sj;
        double B = (2.0 * j + 1) * (-double(m1) * j2j3 + mdiff * jd * jp1);
        double C = jd * sjp1;
        if (std::abs(C) < 1e-30)
            r.values[i + 1] = wigner_3j(j + 1, j2, j3, m1, m2, m3);
        else
            r.values[i + 1] = (-A * r.values[i - 1] - B * r.values[i]) / C;
    }
    return r;
}

// cpp/src/special.cpp:365-440 (gaunt_batch)
void gaunt_batch(int m, int n, int u, int v, std::vector<double>& a, std::vector<double>& b) {
    int qmax_A = std::min({n, v, (n + v - std::abs(m + u)) / 2});
    int qmax_B = std::min({n, v, (n + v + 1 - std::abs(m + u)) / 2});
    a.assign(qmax_A + 1, 0.0);
    b.assign(qmax_B + 1, 0.0);
    batch3j wz = wigner_3j_batch(n, v, 0, 0, 0);
    batch3j wm = wigner_3j_batch(n, v, -(m + u), m, u);
    double fnm = factorial(n + m), fn_m = factorial(n - m);
    double fvu = factorial(v + u), fv_u = factorial(v - u);
    double sign = neg1pow(m + u);
    for (int q = 0; q <= qmax_A; q++) {
        int p = n + v - 2 * q;
        double ratio = fnm * fvu * factorial(p - m - u) / (fn_m * fv_u * factorial(p + m + u));
        a[q] = sign * (2 * p + 1) * std::sqrt(ratio) * wz.values[p - wz.jmin] * wm.values[p - wm.jmin];
    }
    for (int q = 1; q <= qmax_B; q++) {
        int p = n + v - 2 * q;
        double ratio = fnm * fvu * factorial(p - m - u + 1) / (fn_m * fv_u * factorial(p + m + u + 1));
        b[q] = sign * (2 * p + 3) * std::sqrt(ratio) * wz.values[p - wz.jmin] * wm.values[(p + 1) - wm.jmin];
    }
}

// cpp/src/vsh_translation.cpp:12-34 (vsh_cache constructor; m is the
// already-negated order m' of the caller)
struct vsh_cache {
    double factor;
    int qmax_A, qmax_B;
    std::vector<cd> A, B;
};

vsh_cache make_cache(int n, int m, int v, int u) {
    vsh_cache c;
    c.factor = 0.5 * neg1pow(m) *
               std::sqrt((2 * v + 1) * (2 * n + 1) * factorial(v - u) * factorial(n - m) /
                         (v * (v + 1) * n * (n + 1) * factorial(v + u) * factorial(n + m)));
    c.qmax_A = std::min({n, v, (n + v - std::abs(m + u)) / 2});
    c.qmax_B = std::min({n, v, (n + v + 1 - std::abs(m + u)) / 2});
    std::vector<double> a, b;
    gaunt_batch(m, n, u, v, a, b);
    c.A.resize(c.qmax_A + 1);
    c.B.assign(c.qmax_B + 1, cd(0, 0));
    for (int q = 0; q <= c.qmax_A; q++) {
        int p = n + v - 2 * q;
        c.A[q] = a[q] * ipow(p) * double(n * (n + 1) + v * (v + 1) - p * (p + 1));
    }
    for (int q = 1; q <= c.qmax_B; q++) {
        int p = n + v - 2 * q;
        c.B[q] = b[q] * ipow(p + 1) *
                 std::sqrt((double((p + 1) * (p + 1)) - double((n - v) * (n - v))) *
                           (double((n + v + 1) * (n + v + 1)) - double((p + 1) * (p + 1))));
    }
    return c;
}

// Flat cache over the tuple space (n, m', v, u), m' in [-n, n], v <= lmax.
// The reference keys a std::map by the same tuple (vsh_translation.cpp:36-51).
struct cache_table {
    int lmax;
    std::vector<vsh_cache> entries;
    int key(int n, int m, int v, int u) const {
        int rn = n * (n + 2) - n + m - 1;
        int rv = v * (v + 2) - v + u - 1;
        return rn * lmax_to_rmax(lmax) + rv;
    }
    const vsh_cache& at(int n, int m, int v, int u) const { return entries[key(n, m, v, u)]; }
};

cache_table make_cache_table(int lmax) {
    cache_table t;
    t.lmax = lmax;
    int rmax = lmax_to_rmax(lmax);
    t.entries.resize(rmax * rmax);
    for (int n = 1; n <= lmax; n++)
        for (int m = -n; m <= n; m++)
            for (int v = 1; v <= lmax; v++)
                for (int u = -v; u <= v; u++) t.entries[t.key(n, m, v, u)] = make_cache(n, m, v, u);
    return t;
}

// cpp/src/vsh_translation.cpp:161-237 (vsh_translation_insert_pair, Mie
// overload): A/B translation for the reduced tuple set, four mirrored writes
// scaled by the Mie coefficient of the column particle.
void insert_pair_mie(cd* T, int size, const cd* mie, int lmax, int i, int j, double rad, double theta,
                     double phi, double k, const cache_table& cache) {
    int rmax = lmax_to_rmax(lmax);
    int p_max = 2 * lmax + 1;
    cd zn[kMaxDeg];
    hn_recursion(p_max, k * rad, zn);
    double Pnm[kMaxDeg * kMaxDeg];
    legendre_recursion(p_max, std::cos(theta), Pnm);
    auto MIE = [&](int part, int idx) { return mie[part * 2 * lmax + idx]; };
    auto AT = [&](int r, int c) -> cd& { return T[(size_t)r * size + c]; };

    for (int n = 1; n <= lmax; n++) {
        for (int m = -n; m <= n; m++) {
            for (int v = 1; v <= n; v++) {
                for (int u = -v; u <= v; u++) {
                    if (n == v && u < -m) continue;
                    int mp = -m;
                    const vsh_cache& c = cache.at(n, mp, v, u);
                    cd exp_phi = std::exp(cd(0, double(u + mp) * phi));
                    cd sum = 0;
                    for (int q = 0; q <= c.qmax_A; q++) {
                        int p = n + v - 2 * q;
                        sum += c.A[q] * Pnm[p * (p + 2) - p + (u + mp)] * zn[p];
                    }
                    cd At = c.factor * exp_phi * sum;
                    sum = 0;
                    for (int q = 1; q <= c.qmax_B; q++) {
                        int p = n + v - 2 * q;
                        sum += c.B[q] * Pnm[(p + 1) * (p + 3) - (p + 1) + (u + mp)] * zn[p + 1];
                    }
                    cd Bt = -c.factor * exp_phi * sum;
                    cd transfer[2] = {At, Bt};
                    for (int a = 0; a < 2; a++) {
                        for (int b = 0; b < 2; b++) {
                            cd val = transfer[(a + b) % 2];
                            int idx = i * (2 * rmax) + a * rmax + n * (n + 2) - n + m - 1;
                            int idy = j * (2 * rmax) + b * rmax + v * (v + 2) - v + u - 1;
                            AT(idx, idy) = val * MIE(j, b * lmax + v - 1);

                            idx = j * (2 * rmax) + a * rmax + n * (n + 2) - n + m - 1;
                            idy = i * (2 * rmax) + b * rmax + v * (v + 2) - v + u - 1;
                            AT(idx, idy) = neg1pow(n + v + a + b) * val * MIE(i, b * lmax + v - 1);

                            if ((n == v && u != -m) || (n != v)) {
                                idx = i * (2 * rmax) + b * rmax + v * (v + 2) - v - u - 1;
                                idy = j * (2 * rmax) + a * rmax + n * (n + 2) - n - m - 1;
                                AT(idx, idy) = neg1pow(m + u - a - b) * val * MIE(j, a * lmax + n - 1);

                                idx = j * (2 * rmax) + b * rmax + v * (v + 2) - v - u - 1;
                                idy = i * (2 * rmax) + a * rmax + n * (n + 2) - n - m - 1;
                                AT(idx, idy) = neg1pow(m + u + n + v) * val * MIE(i, a * lmax + n - 1);
                            }
                        }
                    }
                }
            }
        }
    }
}

// cpp/src/vsh_translation.cpp:53-159 (vsh_translation_insert_pair, T-matrix
// overload): full translation block A_matrix for the pair, then
// agg(i,j) += A(i, j') T_j(j', j) and agg(j,i) += A(j, i') T_i(i', i), where
// the reference contracts over the mirrored index (-u') of its translation.
void insert_pair_tmatrix(cd* T, int size, const cd* tmat, int rmax, int i, int j, double rad, double theta,
                         double phi, double k, const cache_table& cache) {
    int lmax = rmax_to_lmax(rmax);
    int p_max = 2 * lmax + 1;
    std::vector<cd> zn(p_max + 1);
    hn_recursion(p_max, k * rad, zn.data());
    std::vector<double> Pnm((p_max + 1) * (p_max + 1));
    legendre_recursion(p_max, std::cos(theta), Pnm.data());
    int bs = 2 * rmax;
    // Local translation blocks: Aij[ida][idc] (row block i, col block j) and
    // Aji[ida][idc] (row block j, col block i).
    std::vector<cd> Aij(bs * bs, cd(0, 0)), Aji(bs * bs, cd(0, 0));
    for (int n = 1; n <= lmax; n++) {
        for (int m = -n; m <= n; m++) {
            for (int v = 1; v <= n; v++) {
                for (int u = -v; u <= v; u++) {
                    if (n == v && u < -m) continue;
                    int mp = -m;
                    const vsh_cache& c = cache.at(n, mp, v, u);
                    cd exp_phi = std::exp(cd(0, double(u + mp) * phi));
                    cd sum = 0;
                    for (int q = 0; q <= c.qmax_A; q++) {
                        int p = n + v - 2 * q;
                        sum += c.A[q] * Pnm[p * (p + 2) - p + (u + mp)] * zn[p];
                    }
                    cd At = c.factor * exp_phi * sum;
                    sum = 0;
                    for (int q = 1; q <= c.qmax_B; q++) {
                        int p = n + v - 2 * q;
                        sum += c.B[q] * Pnm[(p + 1) * (p + 3) - (p + 1) + (u + mp)] * zn[p + 1];
                    }
                    cd Bt = -c.factor * exp_phi * sum;
                    cd transfer[2] = {At, Bt};
                    for (int a = 0; a < 2; a++) {
                        for (int b = 0; b < 2; b++) {
                            cd val = transfer[(a + b) % 2];
                            int ra = a * rmax + n * (n + 2) - n + m - 1;
                            int cb = b * rmax + v * (v + 2) - v + u - 1;
                            Aij[ra * bs + cb] = val;
                            Aji[ra * bs + cb] = neg1pow(n + v + a + b) * val;
                            if ((n == v && u != -m) || (n != v)) {
                                int rb = b * rmax + v * (v + 2) - v - u - 1;
                                int ca = a * rmax + n * (n + 2) - n - m - 1;
                                Aij[rb * bs + ca] = neg1pow(m + u - a - b) * val;
                                Aji[rb * bs + ca] = neg1pow(m + u + n + v) * val;
                            }
                        }
                    }
                }
            }
        }
    }
    // cpp/src/vsh_translation.cpp:132-158: contraction with the mirrored
    // column index idc = bp*rmax + r(vp, -up).
    auto TM = [&](int part, int r, int c) { return tmat[((size_t)part * bs + r) * bs + c]; };
    for (int ida = 0; ida < bs; ida++) {
        for (int idb = 0; idb < bs; idb++) {
            cd sij = 0, sji = 0;
            for (int bp = 0; bp < 2; bp++)
                for (int vp = 1; vp <= lmax; vp++)
                    for (int up = -vp; up <= vp; up++) {
                        int idc = bp * rmax + vp * (vp + 2) - vp - up - 1;
                        sij += Aij[ida * bs + idc] * TM(j, idc, idb);
                        sji += Aji[ida * bs + idc] * TM(i, idc, idb);
                    }
            T[(size_t)(i * bs + ida) * size + j * bs + idb] += sij;
            T[(size_t)(j * bs + ida) * size + i * bs + idb] += sji;
        }
    }
}

// ---------------------------------------------------------------------------
// Dense complex linear algebra (Eigen replaced)
// ---------------------------------------------------------------------------

void gemv(const cd* A, const cd* x, cd* y, int n) {
#pragma omp parallel for schedule(static)
    for (int r = 0; r < n; r++) {
        const cd* row = A + (size_t)r * n;
        cd s = 0;
        for (int c = 0; c < n; c++) s += row[c] * x[c];
        y[r] = s;
    }
}

// Eigen's dot convention: a.dot(b) = sum conj(a_i) b_i.
cd dotc(const std::vector<cd>& a, const std::vector<cd>& b) {
    cd s = 0;
    for (size_t i = 0; i < a.size(); i++) s += std::conj(a[i]) * b[i];
    return s;
}
double sqnorm(const std::vector<cd>& a) {
    double s = 0;
    for (auto& z : a) s += std::norm(z);
    return s;
}

}  // namespace

extern "C" {

// ---- special functions ----------------------------------------------------

void orc_spherical_hn_recursion(int nmax, double z, double* out) {
    hn_recursion(nmax, z, reinterpret_cast<cd*>(out));
}

void orc_spherical_jn_complex(int n, double zr, double zi, double* out) {
    cd r = spherical_jn(n, cd(zr, zi));
    out[0] = r.real();
    out[1] = r.imag();
}

void orc_associated_legendre_recursion(int nmax, double z, double* out) { legendre_recursion(nmax, z, out); }

double orc_associated_legendre(int n, int m, double z) { return associated_legendre(n, m, z); }

double orc_pi_func(int n, int m, double theta) { return pi_func(n, m, theta); }

double orc_tau_func(int n, int m, double theta) { return tau_func(n, m, theta); }

// cpp/src/vsh_functions.cpp:56-58
void orc_Emn(int m, int n, double* out) {
    cd r = ipow(n) * std::sqrt((2 * n + 1) * factorial(n - m) / (n * (n + 1) * factorial(n + m)));
    out[0] = r.real();
    out[1] = r.imag();
}

double orc_wigner_3j(int j1, int j2, int j3, int m1, int m2, int m3) { return wigner_3j(j1, j2, j3, m1, m2, m3); }

// Returns the number of values written (jmax-jmin+1, or 0); *jmin set.
int orc_wigner_3j_batch(int j2, int j3, int m1, int m2, int m3, int* jmin, double* out, int cap) {
    batch3j r = wigner_3j_batch(j2, j3, m1, m2, m3);
    *jmin = r.jmin;
    int size = (int)r.values.size();
    for (int i = 0; i < size && i < cap; i++) out[i] = r.values[i];
    return size;
}

double orc_a_func(int m, int n, int u, int v, int p) { return a_func(m, n, u, v, p); }
double orc_b_func(int m, int n, int u, int v, int p) { return b_func(m, n, u, v, p); }

// a_out has qmax_A+1 entries, b_out qmax_B+1 (b_out[0] = 0).
void orc_gaunt_batch(int m, int n, int u, int v, double* a_out, double* b_out) {
    std::vector<double> a, b;
    gaunt_batch(m, n, u, v, a, b);
    std::copy(a.begin(), a.end(), a_out);
    std::copy(b.begin(), b.end(), b_out);
}

// cpp/src/vsh_translation.cpp:428-470 (vsh_translation, scalar).
// mode: 0 outgoing (h_n), 1 ingoing (conj h_n), 2 incident / 3 interior (j_n),
// as get_zn (cpp/src/vsh_functions.cpp:42-54).  out = [A.re, A.im, B.re, B.im].
void orc_vsh_translation(int m, int n, int u, int v, double rad, double theta, double phi, double k, int mode,
                         double* out) {
    m *= -1;
    auto zn = [&](int p, double x) -> cd {
        if (mode == 2 || mode == 3) return spherical_jn(p, cd(x, 0));
        cd h = spherical_hn(p, x);
        return (mode == 0) ? h : std::conj(h);
    };
    cd factor = 0.5 * neg1pow(m) *
                std::sqrt((2 * v + 1) * (2 * n + 1) * factorial(v - u) * factorial(n - m) /
                          (v * (v + 1) * n * (n + 1) * factorial(v + u) * factorial(n + m))) *
                std::exp(cd(0, double(u + m) * phi));
    double ct = std::cos(theta);
    std::vector<double> a, b;
    gaunt_batch(m, n, u, v, a, b);
    int qmax = std::min({n, v, (n + v - std::abs(m + u)) / 2});
    cd sum = 0;
    for (int q = 0; q <= qmax; q++) {
        int p = n + v - 2 * q;
        cd A = a[q] * ipow(p) * double(n * (n + 1) + v * (v + 1) - p * (p + 1));
        sum += A * associated_legendre(p, u + m, ct) * zn(p, k * rad);
    }
    cd At = factor * sum;
    qmax = std::min({n, v, (n + v + 1 - std::abs(m + u)) / 2});
    sum = 0;
    for (int q = 1; q <= qmax; q++) {
        int p = n + v - 2 * q;
        cd B = b[q] * ipow(p + 1) *
               std::sqrt((double((p + 1) * (p + 1)) - double((n - v) * (n - v))) *
                         (double((n + v + 1) * (n + v + 1)) - double((p + 1) * (p + 1))));
        sum += B * associated_legendre(p + 1, u + m, ct) * zn(p + 1, k * rad);
    }
    cd Bt = -factor * sum;
    out[0] = At.real();
    out[1] = At.imag();
    out[2] = Bt.real();
    out[3] = Bt.imag();
}

// ---- aggregate T-matrix ----------------------------------------------------

// cpp/src/interactions.cpp:370-415 (sphere_aggregate_tmatrix).
// positions [N,3] f64, mie [N, 2*lmax] c128, out [n, n] c128 (n = 2 rmax N).
void orc_sphere_aggregate_tmatrix(const double* positions, const double* mie, int N, int lmax, double k,
                                  double* out) {
    int rmax = lmax_to_rmax(lmax);
    int size = 2 * rmax * N;
    cd* T = reinterpret_cast<cd*>(out);
    std::memset(out, 0, sizeof(cd) * (size_t)size * size);
    if (N == 1) return;
    std::vector<int> iv, jv;
    for (int i = 0; i < N; i++)
        for (int j = i + 1; j < N; j++) { iv.push_back(i); jv.push_back(j); }
    cache_table cache = make_cache_table(lmax);
    const cd* M = reinterpret_cast<const cd*>(mie);
    int P = (int)iv.size();
#pragma omp parallel for schedule(dynamic, 16)
    for (int ij = 0; ij < P; ij++) {
        int i = iv[ij], j = jv[ij];
        double dx = positions[3 * i] - positions[3 * j];
        double dy = positions[3 * i + 1] - positions[3 * j + 1];
        double dz = positions[3 * i + 2] - positions[3 * j + 2];
        double rad = std::sqrt(dx * dx + dy * dy + dz * dz);
        double theta = std::acos(dz / rad);
        double phi = std::atan2(dy, dx);
        insert_pair_mie(T, size, M, lmax, i, j, rad, theta, phi, k, cache);
    }
}

// cpp/src/interactions.cpp:417-462 (particle_aggregate_tmatrix).
// tmatrix [N, 2rmax, 2rmax] c128.
void orc_particle_aggregate_tmatrix(const double* positions, const double* tmatrix, int N, int rmax, double k,
                                    double* out) {
    int size = 2 * rmax * N;
    cd* T = reinterpret_cast<cd*>(out);
    std::memset(out, 0, sizeof(cd) * (size_t)size * size);
    if (N == 1) return;
    cache_table cache = make_cache_table(rmax_to_lmax(rmax));
    const cd* TM = reinterpret_cast<const cd*>(tmatrix);
    // Pairs sharing a particle touch the same rows of T, so the pair loop
    // runs serially here (the oracle is for test sizes).
    for (int i = 0; i < N; i++)
        for (int j = i + 1; j < N; j++) {
            double dx = positions[3 * i] - positions[3 * j];
            double dy = positions[3 * i + 1] - positions[3 * j + 1];
            double dz = positions[3 * i + 2] - positions[3 * j + 2];
            double rad = std::sqrt(dx * dx + dy * dy + dz * dz);
            insert_pair_tmatrix(T, size, TM, rmax, i, j, rad, std::acos(dz / rad), std::atan2(dy, dx), k, cache);
        }
}

// ---- solvers ---------------------------------------------------------------

// Block-diagonal preconditioner apply (cpp/src/interactions.cpp:211-228):
// y[blk] = Minv[blk rows, 0:bs] x[blk] for each bs-long block; Minv is the
// [n, bs] row-major stack of the N block inverses.  Minv == null is the
// identity (a copy).
static void apply_block_precond(const cd* Minv, int bs, const std::vector<cd>& x, std::vector<cd>& y) {
    const int n = (int)x.size();
    if (!Minv) {
        y = x;
        return;
    }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
        const int b0 = (i / bs) * bs;
        cd acc = 0;
        for (int k = 0; k < bs; k++) acc += Minv[(size_t)i * bs + k] * x[b0 + k];
        y[i] = acc;
    }
}

// Restates Eigen::BiCGSTAB (Eigen 3.x, IterativeLinearSolvers/BiCGSTAB.h) as
// driven at cpp/src/interactions.cpp:359-365 (identity preconditioner) and
// :324-333 (BlockDiagonalPreconditioner, Minv != null): zero initial guess,
// stop when ||r||^2 <= tol^2 ||b||^2 or at maxiter, restart on rho
// breakdown; right preconditioning y = M^-1 p, z = M^-1 s, x += alpha y + w z.
// A is the full system matrix (I + T).  Returns iterations; *err = ||r||/||b||.
static int bicgstab_eigen(const cd* A, const cd* bp, int n, const cd* Minv, int bs, double tol, int maxiter, cd* xout,
                   double* err) {
    std::vector<cd> b(bp, bp + n), x(n, cd(0, 0));
    std::vector<cd> r(n), r0, v(n, 0.0), p(n, 0.0), s(n), t(n), tmp(n), y(n), z(n);
    gemv(A, x.data(), tmp.data(), n);
    for (int i = 0; i < n; i++) r[i] = b[i] - tmp[i];
    r0 = r;
    double r0_sq = sqnorm(r0), rhs_sq = sqnorm(b);
    if (rhs_sq == 0) {
        std::fill(xout, xout + n, cd(0, 0));
        *err = 0;
        return 0;
    }
    cd rho = 1, alpha = 1, w = 1;
    double tol2 = tol * tol * rhs_sq;
    double eps = 2.220446049250313e-16;
    double eps2 = eps * eps;
    int it = 0, restarts = 0;
    while (sqnorm(r) > tol2 && it < maxiter) {
        cd rho_old = rho;
        rho = dotc(r0, r);
        if (std::abs(rho) < eps2 * r0_sq) {
            gemv(A, x.data(), tmp.data(), n);
            for (int i = 0; i < n; i++) r[i] = b[i] - tmp[i];
            r0 = r;
            r0_sq = sqnorm(r);
            rho = r0_sq;
            if (restarts++ == 0) it = 0;
        }
        cd beta = (rho / rho_old) * (alpha / w);
        for (int i = 0; i < n; i++) p[i] = r[i] + beta * (p[i] - w * v[i]);
        apply_block_precond(Minv, bs, p, y);
        gemv(A, y.data(), v.data(), n);
        alpha = rho / dotc(r0, v);
        for (int i = 0; i < n; i++) s[i] = r[i] - alpha * v[i];
        apply_block_precond(Minv, bs, s, z);
        gemv(A, z.data(), t.data(), n);
        double tt = sqnorm(t);
        w = (tt > 0) ? dotc(t, s) / tt : cd(0, 0);
        for (int i = 0; i < n; i++) {
            x[i] += alpha * y[i] + w * z[i];
            r[i] = s[i] - w * t[i];
        }
        ++it;
    }
    *err = std::sqrt(sqnorm(r) / rhs_sq);
    std::memcpy(xout, x.data(), sizeof(cd) * n);
    return it;
}

int orc_bicgstab_eigen(const double* A_, const double* b_, int n, double tol, int maxiter, double* x_,
                       double* err) {
    return bicgstab_eigen(reinterpret_cast<const cd*>(A_), reinterpret_cast<const cd*>(b_), n, nullptr, 1, tol,
                          maxiter, reinterpret_cast<cd*>(x_), err);
}

// The same with the block-diagonal preconditioner Minv [n, bs]
// (cpp/src/interactions.cpp:305-345, solve_linear_system_preconditioned).
int orc_bicgstab_eigen_pc(const double* A_, const double* b_, int n, const double* Minv, int bs, double tol,
                          int maxiter, double* x_, double* err) {
    return bicgstab_eigen(reinterpret_cast<const cd*>(A_), reinterpret_cast<const cd*>(b_), n,
                          reinterpret_cast<const cd*>(Minv), bs, tol, maxiter, reinterpret_cast<cd*>(x_), err);
}

// cpp/src/interactions.cpp:152-209 (bicgstab_profiled; bicgstab :98-150 is
// the same loop) and, with Minv != null, :230-289
// (bicgstab_preconditioned_profiled: p_hat = M^-1 p, s_hat = M^-1 s,
// x = x + alpha p_hat + w s_hat): x0 = b, absolute tolerance on ||r||, no
// restart.  Returns iterations.
static int bicgstab_custom(const cd* A, const cd* bp, int n, const cd* Minv, int bs, int maxiter, double tol, cd* xout,
                    double* residual) {
    std::vector<cd> b(bp, bp + n), x = b, r(n), tmp(n);
    gemv(A, x.data(), tmp.data(), n);
    for (int i = 0; i < n; i++) r[i] = b[i] - tmp[i];
    double error = std::sqrt(sqnorm(r));
    if (error < tol) {
        std::memcpy(xout, x.data(), sizeof(cd) * n);
        *residual = error;
        return 0;
    }
    std::vector<cd> rhat = r, v(n, 0.0), p(n, 0.0), s(n), t(n), pi(n), vi(n), ph(n), sh(n);
    cd rho_prev = 1, alpha = 1, w_prev = 1;
    int it = 1;
    while (true) {
        cd rho = dotc(rhat, r);
        cd beta = (rho / rho_prev) * (alpha / w_prev);
        for (int i = 0; i < n; i++) pi[i] = r[i] + beta * (p[i] - w_prev * v[i]);
        apply_block_precond(Minv, bs, pi, ph);
        gemv(A, ph.data(), vi.data(), n);
        alpha = rho / dotc(rhat, vi);
        for (int i = 0; i < n; i++) s[i] = r[i] - alpha * vi[i];
        apply_block_precond(Minv, bs, s, sh);
        gemv(A, sh.data(), t.data(), n);
        cd w = dotc(t, s) / dotc(t, t);
        for (int i = 0; i < n; i++) {
            x[i] = x[i] + alpha * ph[i] + w * sh[i];
            r[i] = s[i] - w * t[i];
        }
        error = std::sqrt(sqnorm(r));
        if (error < tol || it > maxiter) break;
        rho_prev = rho;
        v = vi;
        p = pi;
        w_prev = w;
        it++;
    }
    std::memcpy(xout, x.data(), sizeof(cd) * n);
    *residual = error;
    return it;
}

int orc_bicgstab_custom(const double* A_, const double* b_, int n, int maxiter, double tol, double* x_,
                        double* residual) {
    return bicgstab_custom(reinterpret_cast<const cd*>(A_), reinterpret_cast<const cd*>(b_), n, nullptr, 1, maxiter,
                           tol, reinterpret_cast<cd*>(x_), residual);
}

int orc_bicgstab_custom_pc(const double* A_, const double* b_, int n, const double* Minv, int bs, int maxiter,
                           double tol, double* x_, double* residual) {
    return bicgstab_custom(reinterpret_cast<const cd*>(A_), reinterpret_cast<const cd*>(b_), n,
                           reinterpret_cast<const cd*>(Minv), bs, maxiter, tol, reinterpret_cast<cd*>(x_), residual);
}

// Partial-pivot LU solve (Eigen partialPivLu at cpp/src/interactions.cpp:357).
// Returns 0, or 1 for an exactly singular pivot.
int orc_lu_solve(const double* A_, const double* b_, int n, double* x_) {
    const cd* Ain = reinterpret_cast<const cd*>(A_);
    std::vector<cd> A(Ain, Ain + (size_t)n * n);
    std::vector<cd> b(reinterpret_cast<const cd*>(b_), reinterpret_cast<const cd*>(b_) + n);
    for (int c = 0; c < n; c++) {
        int piv = c;
        double best = std::abs(A[(size_t)c * n + c]);
        for (int r = c + 1; r < n; r++) {
            double v = std::abs(A[(size_t)r * n + c]);
            if (v > best) { best = v; piv = r; }
        }
        if (best == 0) return 1;
        if (piv != c) {
            for (int k = 0; k < n; k++) std::swap(A[(size_t)c * n + k], A[(size_t)piv * n + k]);
            std::swap(b[c], b[piv]);
        }
        cd d = A[(size_t)c * n + c];
#pragma omp parallel for schedule(static)
        for (int r = c + 1; r < n; r++) {
            cd f = A[(size_t)r * n + c] / d;
            A[(size_t)r * n + c] = f;
            for (int k = c + 1; k < n; k++) A[(size_t)r * n + k] -= f * A[(size_t)c * n + k];
            b[r] -= f * b[c];
        }
    }
    for (int r = n - 1; r >= 0; r--) {
        cd s = b[r];
        for (int k = r + 1; k < n; k++) s -= A[(size_t)r * n + k] * b[k];
        b[r] = s / A[(size_t)r * n + r];
    }
    std::memcpy(x_, b.data(), sizeof(cd) * n);
    return 0;
}

// cpp/src/interactions.cpp:347-368 (solve_linear_system): A = T + I, then
// method 0 = BiCGSTAB (Eigen restatement, tol/maxiter given; the reference
// hard-codes 1e-5 / 1000), method 1 = exact (LU).  Returns iterations (0 for
// LU) and writes ||r||/||b|| to *err for BiCGSTAB.
int orc_solve_linear_system(const double* T_, const double* p_src, int n, int method, double tol, int maxiter,
                            double* x, double* err) {
    const cd* T = reinterpret_cast<const cd*>(T_);
    std::vector<cd> A(T, T + (size_t)n * n);
    for (int i = 0; i < n; i++) A[(size_t)i * n + i] += 1.0;
    if (method == 1) {
        *err = 0;
        return orc_lu_solve(reinterpret_cast<double*>(A.data()), p_src, n, x) ? -1 : 0;
    }
    return orc_bicgstab_eigen(reinterpret_cast<double*>(A.data()), p_src, n, tol, maxiter, x, err);
}

// Dense y = A x (test helper for residuals).
void orc_gemv(const double* A, const double* x, double* y, int n) {
    gemv(reinterpret_cast<const cd*>(A), reinterpret_cast<const cd*>(x), reinterpret_cast<cd*>(y), n);
}

// ---- cross-sections --------------------------------------------------------

// cpp/src/flux.cpp:5-30.  Outputs are [2*lmax] each (index a*lmax + n-1).
void orc_particle_cross_sections(const double* p_scat_, const double* p_inc_, const double* p_src_, int rmax,
                                 double k, double* Cscat, double* Cabs, double* Cext) {
    const cd* ps = reinterpret_cast<const cd*>(p_scat_);
    const cd* pi = reinterpret_cast<const cd*>(p_inc_);
    const cd* pr = reinterpret_cast<const cd*>(p_src_);
    int lmax = rmax_to_lmax(rmax);
    double factor = 4 * PI / (k * k);
    for (int i = 0; i < 2 * lmax; i++) Cabs[i] = Cext[i] = 0;
    for (int n = 1; n <= lmax; n++)
        for (int m = -n; m <= n; m++)
            for (int a = 0; a < 2; a++) {
                int r = a * rmax + n * n + n + m - 1;
                int idx = a * lmax + n - 1;
                Cabs[idx] += factor * (std::real(std::conj(pi[r]) * ps[r]) - std::norm(ps[r]));
                Cext[idx] += factor * std::real(std::conj(pr[r]) * ps[r]);
            }
    for (int i = 0; i < 2 * lmax; i++) Cscat[i] = Cext[i] - Cabs[i];
}

// cpp/src/flux.cpp:32-57
void orc_cluster_cross_sections(const double* p_cluster_, const double* p_src_, int rmax, double k, double* Cscat,
                                double* Cabs, double* Cext) {
    const cd* pc = reinterpret_cast<const cd*>(p_cluster_);
    const cd* pr = reinterpret_cast<const cd*>(p_src_);
    int lmax = rmax_to_lmax(rmax);
    double factor = 4 * PI / (k * k);
    for (int i = 0; i < 2 * lmax; i++) Cscat[i] = Cext[i] = 0;
    for (int n = 1; n <= lmax; n++)
        for (int m = -n; m <= n; m++)
            for (int a = 0; a < 2; a++) {
                int r = a * rmax + n * n + n + m - 1;
                int idx = a * lmax + n - 1;
                Cscat[idx] += factor * std::norm(pc[r]);
                Cext[idx] += factor * std::real(std::conj(pr[r]) * pc[r]);
            }
    for (int i = 0; i < 2 * lmax; i++) Cabs[i] = Cext[i] - Cscat[i];
}

// ---- force and torque ------------------------------------------------------

// GSL_CONST_MKSA_VACUUM_PERMITTIVITY as compiled into GSL (cpp/src/forces.cpp:12).
const double ORC_EPS0 = 8.854187817e-12;

double orc_eps0() { return ORC_EPS0; }

// cpp/src/forces.cpp:7-86 (force).  p_scat, p_inc are [2*rmax]; F[3].
void orc_force(const double* p_scat_, const double* p_inc_, int rmax, double k, double eps_b, double mu_b,
               double* F) {
    (void)mu_b;
    const cd* p = reinterpret_cast<const cd*>(p_scat_);
    const cd* q = p + rmax;
    const cd* pi = reinterpret_cast<const cd*>(p_inc_);
    const cd* qi = pi + rmax;
    int lmax = rmax_to_lmax(rmax);
    cd Fxy = 0, Fz = 0;
    double Axy = PI / (k * k) * ORC_EPS0 * eps_b;
    double Az = -2 * Axy;
    for (int n = 1; n <= lmax; n++) {
        for (int m = -n; m <= n; m++) {
            int r = n * n + n + m - 1;
            if (m != n) {
                double factor = Axy * std::sqrt(double((n + m + 1) * (n - m))) / double(n * (n + 1));
                int r1 = r + 1;
                Fxy += factor * (2.0 * p[r] * std::conj(q[r1]) - p[r] * std::conj(qi[r1]) - pi[r] * std::conj(q[r1]) +
                                 2.0 * q[r] * std::conj(p[r1]) - q[r] * std::conj(pi[r1]) - qi[r] * std::conj(p[r1]));
            }
            double factor = Az * m / double(n * (n + 1));
            Fz += factor * (2.0 * p[r] * std::conj(q[r]) - p[r] * std::conj(qi[r]) - pi[r] * std::conj(q[r]));
            if (n < lmax) {
                factor = -Axy * std::sqrt((n + m + 2) * (n + m + 1) * n * (n + 2) / double((2 * n + 3) * (2 * n + 1))) /
                         (n + 1);
                int r1 = (n + 1) * (n + 1) + (n + 1) + m;
                Fxy += factor * (2.0 * p[r] * std::conj(p[r1]) - p[r] * std::conj(pi[r1]) - pi[r] * std::conj(p[r1]) +
                                 2.0 * q[r] * std::conj(q[r1]) - q[r] * std::conj(qi[r1]) - qi[r] * std::conj(q[r1]));
                factor = Axy * std::sqrt((n - m + 1) * (n - m + 2) * n * (n + 2) / double((2 * n + 3) * (2 * n + 1))) /
                         (n + 1);
                r1 = (n + 1) * (n + 1) + (n + 1) + m - 2;
                Fxy += factor * (2.0 * p[r1] * std::conj(p[r]) - p[r1] * std::conj(pi[r]) - pi[r1] * std::conj(p[r]) +
                                 2.0 * q[r1] * std::conj(q[r]) - q[r1] * std::conj(qi[r]) - qi[r1] * std::conj(q[r]));
                factor = Az * std::sqrt((n - m + 1) * (n + m + 1) * n * (n + 2) / double((2 * n + 3) * (2 * n + 1))) /
                         (n + 1);
                r1 = (n + 1) * (n + 1) + (n + 1) + m - 1;
                Fz += factor * (2.0 * p[r1] * std::conj(p[r]) - p[r1] * std::conj(pi[r]) - pi[r1] * std::conj(p[r]) +
                                2.0 * q[r1] * std::conj(q[r]) - q[r1] * std::conj(qi[r]) - qi[r1] * std::conj(q[r]));
            }
        }
    }
    F[0] = Fxy.real();
    F[1] = Fxy.imag();
    F[2] = Fz.real();
}

// cpp/src/forces.cpp:88-137 (torque)
void orc_torque(const double* p_scat_, const double* p_inc_, int rmax, double k, double eps_b, double mu_b,
                double* T) {
    (void)mu_b;
    const cd* p = reinterpret_cast<const cd*>(p_scat_);
    const cd* q = p + rmax;
    const cd* pi = reinterpret_cast<const cd*>(p_inc_);
    const cd* qi = pi + rmax;
    int lmax = rmax_to_lmax(rmax);
    double A = -2 * PI / (k * k * k) * ORC_EPS0 * eps_b;
    T[0] = T[1] = T[2] = 0;
    for (int n = 1; n <= lmax; n++) {
        for (int m = -n; m <= n; m++) {
            int r = n * n + n + m - 1;
            if (m != n) {
                double factor = -A * std::sqrt(double((n - m) * (n + m + 1)));
                int r1 = r + 1;
                T[0] += factor * std::real(p[r] * std::conj(p[r1]) + q[r] * std::conj(q[r1]) -
                                           0.5 * (p[r1] * std::conj(pi[r]) + p[r] * std::conj(pi[r1]) +
                                                  q[r1] * std::conj(qi[r]) + q[r] * std::conj(qi[r1])));
                T[1] += factor * std::imag(p[r] * std::conj(p[r1]) + q[r] * std::conj(q[r1]) +
                                           0.5 * (p[r1] * std::conj(pi[r]) - p[r] * std::conj(pi[r1]) +
                                                  q[r1] * std::conj(qi[r]) - q[r] * std::conj(qi[r1])));
            }
            double factor = A * m;
            T[2] += factor * (std::norm(p[r]) + std::norm(q[r]) - std::real(p[r] * std::conj(pi[r]) + q[r] * std::conj(qi[r])));
        }
    }
}

// cpp/src/forces.cpp:139-154 / 156-171: per-particle loops over [N, 2rmax].
void orc_force_all(const double* p_scat_all, const double* p_inc_all, int N, int rmax, double k, double eps_b,
                   double mu_b, double* F) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; i++)
        orc_force(p_scat_all + (size_t)i * 4 * rmax, p_inc_all + (size_t)i * 4 * rmax, rmax, k, eps_b, mu_b, F + 3 * i);
}

void orc_torque_all(const double* p_scat_all, const double* p_inc_all, int N, int rmax, double k, double eps_b,
                    double mu_b, double* T) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; i++)
        orc_torque(p_scat_all + (size_t)i * 4 * rmax, p_inc_all + (size_t)i * 4 * rmax, rmax, k, eps_b, mu_b, T + 3 * i);
}

int orc_max_threads() { return omp_get_max_threads(); }
void orc_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }

}  // extern "C"
