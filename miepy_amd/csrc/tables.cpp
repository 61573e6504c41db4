// Host-side construction of the per-lmax VSH translation coefficient tables.
//
// The coefficients follow the reference's vsh_cache (cpp/src/vsh_translation.cpp:12-34)
// and Gaunt functions a_q, b_q (cpp/src/special.cpp:276-296, 365-440):
//   f    = 1/2 (-1)^m' sqrt[(2v+1)(2n+1)(v-u)!(n-m')! / (v(v+1)n(n+1)(v+u)!(n+m')!)]
//   A(q) = a_q i^p [n(n+1) + v(v+1) - p(p+1)],                      p = n+v-2q
//   B(q) = b_q i^{p+1} sqrt[((p+1)^2-(n-v)^2)((n+v+1)^2-(p+1)^2)]
// with m' = -m (the row order enters negated, vsh_translation.cpp:178).
// Unlike the reference's per-call std::map, one flat table per lmax is built
// once per process and device, uploaded, and read by every kernel.  Wigner 3j
// symbols come from the Racah formula summed in x87 extended precision.
// Negative Legendre orders are folded into the coefficients:
//   P_p^{-mu} = (-1)^mu (p-mu)!/(p+mu)! P_p^{mu}  (cpp/src/special.cpp:193).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "common.hpp"

namespace {

thread_local std::string g_last_error;

long double lfact(int k) {
    static long double cache[64];
    static bool init = false;
    if (!init) {
        cache[0] = 1.0L;
        for (int i = 1; i < 64; i++) cache[i] = cache[i - 1] * i;
        init = true;
    }
    return cache[k];
}

// Racah formula (Edmonds eq. 3.7.3)
long double w3j(int j1, int j2, int j3, int m1, int m2, int m3) {
    if (m1 + m2 + m3 != 0) return 0;
    if (std::abs(m1) > j1 || std::abs(m2) > j2 || std::abs(m3) > j3) return 0;
    if (j3 < std::abs(j1 - j2) || j3 > j1 + j2) return 0;
    long double tri = lfact(j1 + j2 - j3) * lfact(j1 - j2 + j3) * lfact(-j1 + j2 + j3) / lfact(j1 + j2 + j3 + 1);
    long double pre = std::sqrt(tri * lfact(j1 + m1) * lfact(j1 - m1) * lfact(j2 + m2) * lfact(j2 - m2) *
                                lfact(j3 + m3) * lfact(j3 - m3));
    int kmin = std::max({0, j2 - j3 - m1, j1 - j3 + m2});
    int kmax = std::min({j1 + j2 - j3, j1 - m1, j2 + m2});
    long double s = 0;
    for (int k = kmin; k <= kmax; k++) {
        long double d = lfact(k) * lfact(j3 - j2 + k + m1) * lfact(j3 - j1 + k - m2) * lfact(j1 + j2 - j3 - k) *
                        lfact(j1 - k - m1) * lfact(j2 - k + m2);
        s += (k & 1 ? -1.0L : 1.0L) / d;
    }
    int ph = j1 - j2 - m3;
    return ((ph % 2 + 2) % 2 ? -1.0L : 1.0L) * pre * s;
}

// a_q and b_q Gaunt coefficients (special.cpp:276-296) for p = n+v-2q
long double gaunt_a(int m, int n, int u, int v, int p) {
    long double num = lfact(n + m) * lfact(v + u) * lfact(p - m - u);
    long double den = lfact(n - m) * lfact(v - u) * lfact(p + m + u);
    long double sign = ((m + u) % 2 + 2) % 2 ? -1.0L : 1.0L;
    return sign * (2 * p + 1) * std::sqrt(num / den) * w3j(n, v, p, 0, 0, 0) * w3j(n, v, p, m, u, -m - u);
}
long double gaunt_b(int m, int n, int u, int v, int p) {
    long double num = lfact(n + m) * lfact(v + u) * lfact(p - m - u + 1);
    long double den = lfact(n - m) * lfact(v - u) * lfact(p + m + u + 1);
    long double sign = ((m + u) % 2 + 2) % 2 ? -1.0L : 1.0L;
    return sign * (2 * p + 3) * std::sqrt(num / den) * w3j(n, v, p, 0, 0, 0) * w3j(n, v, p + 1, m, u, -m - u);
}

struct Key {
    int device, lmax;
    bool operator<(const Key& o) const { return device != o.device ? device < o.device : lmax < o.lmax; }
};

std::mutex g_mutex;
std::map<Key, TransTable*> g_tables;

// i^p as (re, im)
void ipow(int p, long double& re, long double& im) {
    switch (((p % 4) + 4) % 4) {
        case 0: re = 1; im = 0; break;
        case 1: re = 0; im = 1; break;
        case 2: re = -1; im = 0; break;
        default: re = 0; im = -1; break;
    }
}

TransTable* build_table(int lmax) {
    const int rmax = lmax_to_rmax(lmax);
    const int qa_slots = lmax + 1;
    const int coef_slots = 2 * lmax + 1;
    std::vector<int4> meta(rmax * rmax);
    std::vector<double2> coef((size_t)rmax * rmax * coef_slots, make_double2(0, 0));
    for (int n = 1; n <= lmax; n++)
        for (int m = -n; m <= n; m++)
            for (int v = 1; v <= lmax; v++)
                for (int u = -v; u <= v; u++) {
                    const int t = mode_r(n, m) * rmax + mode_r(v, u);
                    const int mp = -m;  // row order enters negated
                    const int mu = u + mp;
                    const int qA = std::min({n, v, (n + v - std::abs(mu)) / 2});
                    const int qB = std::min({n, v, (n + v + 1 - std::abs(mu)) / 2});
                    long double f = 0.5L * (((mp % 2) + 2) % 2 ? -1.0L : 1.0L) *
                                    std::sqrt((long double)(2 * v + 1) * (2 * n + 1) * lfact(v - u) * lfact(n - mp) /
                                              ((long double)v * (v + 1) * n * (n + 1) * lfact(v + u) * lfact(n + mp)));
                    auto legsign = [&](int p) -> long double {
                        if (mu >= 0) return 1.0L;
                        const int a = -mu;
                        if (a > p) return 0.0L;
                        return ((a % 2) ? -1.0L : 1.0L) * lfact(p - a) / lfact(p + a);
                    };
                    double2* c = &coef[(size_t)t * coef_slots];
                    for (int q = 0; q <= qA; q++) {
                        const int p = n + v - 2 * q;
                        long double a = gaunt_a(mp, n, u, v, p);
                        long double re, im;
                        ipow(p, re, im);
                        long double mag = f * a * (n * (n + 1) + v * (v + 1) - p * (p + 1)) * legsign(p);
                        c[q] = make_double2((double)(mag * re), (double)(mag * im));
                    }
                    for (int q = 1; q <= qB; q++) {
                        const int p = n + v - 2 * q;
                        long double b = gaunt_b(mp, n, u, v, p);
                        long double re, im;
                        ipow(p + 1, re, im);
                        long double g = std::sqrt((long double)((p + 1) * (p + 1) - (n - v) * (n - v)) *
                                                  (long double)((n + v + 1) * (n + v + 1) - (p + 1) * (p + 1)));
                        long double mag = -f * b * g * legsign(p + 1);
                        c[qa_slots + q - 1] = make_double2((double)(mag * re), (double)(mag * im));
                    }
                    meta[t] = make_int4(qA + 1, qB, mu, n + v);
                }
    // compact form: only the nonzero-range terms, with their HL indices
    std::vector<int2> cmeta(rmax * rmax);
    std::vector<double2> ccoef;
    std::vector<unsigned short> clidx;
    for (int t = 0; t < rmax * rmax; t++) {
        const int4 mt = meta[t];
        const int amu = std::abs(mt.z);
        cmeta[t] = make_int2((int)ccoef.size(), mt.x | (mt.y << 8) | ((mt.z + 64) << 16));
        for (int q = 0; q < mt.x; q++) {
            ccoef.push_back(coef[(size_t)t * coef_slots + q]);
            clidx.push_back((unsigned short)leg_index(mt.w - 2 * q, amu));
        }
        for (int q = 1; q <= mt.y; q++) {
            ccoef.push_back(coef[(size_t)t * coef_slots + qa_slots + q - 1]);
            clidx.push_back((unsigned short)leg_index(mt.w - 2 * q + 1, amu));
        }
    }
    // contraction form for the factorized matrix-free operator: entries
    // e = (r, kind, mu, p) in the canonical order (r, kind A then B, mu from
    // -2 lmax, p ascending; matfree.hip enumerates the same order at compile
    // time), each with the (s, coefficient) terms that map to it:
    // X_kind[r,s] = sum over those terms of coef * h_p P_p^|mu| e^{i mu phi}.
    std::vector<int2> wmeta;
    std::vector<int> wsrc;
    std::vector<double2> wcoef;
    // and, per (row mode r, HLE index h = p(p+1) + mu), its A and B entries (or -1)
    const int nh = (2 * lmax + 1) * (2 * lmax + 1);
    std::vector<int2> hmap((size_t)rmax * nh, make_int2(-1, -1));
    for (int r = 0; r < rmax; r++)
        for (int kind = 0; kind < 2; kind++)
            for (int mu = -2 * lmax; mu <= 2 * lmax; mu++)
                for (int p = std::abs(mu); p <= 2 * lmax; p++) {
                    const int start = (int)wsrc.size();
                    for (int sidx = 0; sidx < rmax; sidx++) {
                        const int t = r * rmax + sidx;
                        const int4 mt = meta[t];
                        if (mt.z != mu) continue;
                        if (kind == 0) {
                            for (int q = 0; q < mt.x; q++)
                                if (mt.w - 2 * q == p) {
                                    wsrc.push_back(sidx);
                                    wcoef.push_back(coef[(size_t)t * coef_slots + q]);
                                }
                        } else {
                            for (int q = 1; q <= mt.y; q++)
                                if (mt.w - 2 * q + 1 == p) {
                                    wsrc.push_back(sidx);
                                    wcoef.push_back(coef[(size_t)t * coef_slots + qa_slots + q - 1]);
                                }
                        }
                    }
                    const int cnt = (int)wsrc.size() - start;
                    if (cnt) {
                        const int hh = p * (p + 1) + mu;
                        int2& slot = hmap[(size_t)r * nh + hh];
                        (kind == 0 ? slot.x : slot.y) = (int)wmeta.size();
                        wmeta.push_back(make_int2(start, cnt | (kind << 16)));
                    }
                }
    TransTable* T = new TransTable();
    T->lmax = lmax;
    T->rmax = rmax;
    T->qa_slots = qa_slots;
    T->coef_slots = coef_slots;
    if (hipMalloc(&T->meta, meta.size() * sizeof(int4)) != hipSuccess ||
        hipMalloc(&T->coef, coef.size() * sizeof(double2)) != hipSuccess) {
        mpa_set_error("hipMalloc for the lmax=%d translation table failed", lmax);
        delete T;
        return nullptr;
    }
    T->nterms = (int)ccoef.size();
    if (hipMalloc(&T->cmeta, cmeta.size() * sizeof(int2)) != hipSuccess ||
        hipMalloc(&T->ccoef, ccoef.size() * sizeof(double2)) != hipSuccess ||
        hipMalloc(&T->clidx, clidx.size() * sizeof(unsigned short)) != hipSuccess) {
        mpa_set_error("hipMalloc for the lmax=%d translation table failed", lmax);
        delete T;
        return nullptr;
    }
    T->nentries = (int)wmeta.size();
    T->nh = nh;
    if (hipMalloc(&T->hmap, hmap.size() * sizeof(int2)) != hipSuccess ||
        hipMemcpy(T->hmap, hmap.data(), hmap.size() * sizeof(int2), hipMemcpyHostToDevice) != hipSuccess) {
        mpa_set_error("contraction map for lmax=%d failed", lmax);
        delete T;
        return nullptr;
    }
    {
        std::vector<int2> hlist((size_t)rmax * nh, make_int2(0, 0));
        std::vector<int> hsrc;
        std::vector<double2> hcoef;
        for (int r = 0; r < rmax; r++)
            for (int h = 0; h < nh; h++) {
                const int2 es = hmap[(size_t)r * nh + h];
                hlist[(size_t)r * nh + h].x = (int)hsrc.size();
                for (int kind = 0; kind < 2; kind++) {
                    const int e = kind ? es.y : es.x;
                    if (e < 0) continue;
                    const int2 m = wmeta[e];
                    for (int k = m.x; k < m.x + (m.y & 0xffff); k++) {
                        hsrc.push_back(wsrc[k] | (kind << 8));
                        hcoef.push_back(wcoef[k]);
                    }
                }
                hlist[(size_t)r * nh + h].y = (int)hsrc.size() - hlist[(size_t)r * nh + h].x;
            }
        T->nhterms = (int)hsrc.size();
        if (hipMalloc(&T->hlist, hlist.size() * sizeof(int2)) != hipSuccess ||
            hipMalloc(&T->hterm_src, std::max<size_t>(1, hsrc.size()) * sizeof(int)) != hipSuccess ||
            hipMalloc(&T->hterm_coef, std::max<size_t>(1, hcoef.size()) * sizeof(double2)) != hipSuccess ||
            hipMemcpy(T->hlist, hlist.data(), hlist.size() * sizeof(int2), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(T->hterm_src, hsrc.data(), hsrc.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(T->hterm_coef, hcoef.data(), hcoef.size() * sizeof(double2), hipMemcpyHostToDevice) !=
                hipSuccess) {
            mpa_set_error("flattened contraction table for lmax=%d failed", lmax);
            delete T;
            return nullptr;
        }
    }
    if (hipMalloc(&T->wmeta, wmeta.size() * sizeof(int2)) != hipSuccess ||
        hipMalloc(&T->wsrc, wsrc.size() * sizeof(int)) != hipSuccess ||
        hipMalloc(&T->wcoef, wcoef.size() * sizeof(double2)) != hipSuccess ||
        hipMemcpy(T->wmeta, wmeta.data(), wmeta.size() * sizeof(int2), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(T->wsrc, wsrc.data(), wsrc.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(T->wcoef, wcoef.data(), wcoef.size() * sizeof(double2), hipMemcpyHostToDevice) != hipSuccess) {
        mpa_set_error("contraction table for lmax=%d failed", lmax);
        delete T;
        return nullptr;
    }
    if (hipMemcpy(T->meta, meta.data(), meta.size() * sizeof(int4), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(T->coef, coef.data(), coef.size() * sizeof(double2), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(T->cmeta, cmeta.data(), cmeta.size() * sizeof(int2), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(T->ccoef, ccoef.data(), ccoef.size() * sizeof(double2), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(T->clidx, clidx.data(), clidx.size() * sizeof(unsigned short), hipMemcpyHostToDevice) !=
            hipSuccess) {
        mpa_set_error("upload of the lmax=%d translation table failed", lmax);
        delete T;
        return nullptr;
    }
    return T;
}

}  // namespace

void mpa_set_error(const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
}

const TransTable* mpa_get_table(int lmax) {
    if (lmax < 1 || lmax > MPA_LMAX_MAX) {
        mpa_set_error("lmax=%d outside [1, %d]", lmax, MPA_LMAX_MAX);
        return nullptr;
    }
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        mpa_set_error("hipGetDevice failed");
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_mutex);
    Key key{dev, lmax};
    auto it = g_tables.find(key);
    if (it != g_tables.end()) return it->second;
    TransTable* T = build_table(lmax);
    if (T) g_tables[key] = T;
    return T;
}

extern "C" {

const char* mpa_last_error(void) { return g_last_error.c_str(); }

const char* mpa_version(void) { return "miepy_amd 0.1 (gfx950)"; }

// Host copy of the table (for tests): fills meta [rmax*rmax*4] ints and coef
// [rmax*rmax*(2lmax+1)] complex doubles; returns MPA_OK.
int mpa_translation_table(int lmax, int* meta_out, double* coef_out) {
    const TransTable* T = mpa_get_table(lmax);
    if (!T) return MPA_ERR_ARG;
    const int rmax = T->rmax;
    if (hipMemcpy(meta_out, T->meta, sizeof(int4) * rmax * rmax, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(coef_out, T->coef, sizeof(double2) * (size_t)rmax * rmax * T->coef_slots, hipMemcpyDeviceToHost) !=
            hipSuccess) {
        mpa_set_error("table download failed");
        return MPA_ERR_HIP;
    }
    return MPA_OK;
}

}  // extern "C"
