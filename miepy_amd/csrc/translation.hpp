// Per-pair special functions and the translation-coefficient sums shared by
// the assembly and the matrix-free operator.
//
// For a pair at separation d = r_i - r_j (cpp/src/interactions.cpp:404) and
// the compact table of tables.cpp, the two translation blocks are
//   X_A[r,s] = e^{i mu phi} sum_q cA[q] h_p(k|d|) P_p^{|mu|}(cos theta),
//   X_B[r,s] = e^{i mu phi} sum_q cB[q] h_{p'}(k|d|) P_{p'}^{|mu|}(cos theta)
// (cpp/src/vsh_translation.cpp:171-206 with m' = -m).  Per pair, the
// products HL[p][mu] = h_p P_p^mu and e^{i mu phi} are formed once (one
// thread per Legendre order mu) and kept in LDS; the sums then cost one
// complex multiply-add per table term.
#pragma once

#include "common.hpp"

namespace mpa {

// Legendre/Hankel products per pair: (pmax+1)(pmax+2)/2 entries, pmax = 2 lmax + 1.
__host__ __device__ inline int hl_count(int lmax) {
    const int pmax = 2 * lmax + 1;
    return (pmax + 1) * (pmax + 2) / 2;
}

// Thread (pair, mu): HL[leg_index(l, mu)] = h_l(k|d|) P_l^mu(cos theta) for
// l = mu..pmax, and eph[mu] = e^{i mu phi} for mu <= 2 lmax.  Spherical
// Hankel h_l = j_l + i y_l by upward recurrence from j_0 = sin z / z,
// y_0 = -cos z / z (cpp/src/special.cpp:90-112); P_l^mu without the
// Condon-Shortley phase by the standard three-term recurrence in l
// (cpp/src/special.cpp:180-204, GSL_SF_LEGENDRE_NONE).  zkind selects the
// radial function of the VSH mode (cpp/src/vsh_functions.cpp:42-54): 0 h_l^(1)
// = j_l + i y_l (outgoing), 1 h_l^(2) = j_l - i y_l (ingoing), 2 j_l alone
// (incident and interior; spherical_jn's upward recurrence, special.cpp:20-44).
__device__ inline void pair_hl(double2* HL, double2* eph, int mu, int lmax, double dx, double dy, double dz,
                               double k, int zkind = 0) {
    const int pmax = 2 * lmax + 1;
    const double rad = sqrt(dx * dx + dy * dy + dz * dz);
    const double x = dz / rad;
    const double s = sqrt(fmax(0.0, (1.0 - x) * (1.0 + x)));
    double pmm = 1.0;
    for (int i = 1; i <= mu; i++) pmm *= (2 * i - 1) * s;
    const double z = k * rad;
    double sz, cz;
    sincos(z, &sz, &cz);
    double ja = sz / z, ya = -cz / z;           // h_{l-2}
    double jb = ja / z + ya, yb = ya / z - ja;  // h_{l-1}
    double p2 = 0.0, p1 = 0.0;
    for (int l = 0; l <= pmax; l++) {
        double jl, yl;
        if (l == 0) {
            jl = ja;
            yl = ya;
        } else if (l == 1) {
            jl = jb;
            yl = yb;
        } else {
            jl = (2 * l - 1) / z * jb - ja;
            yl = (2 * l - 1) / z * yb - ya;
            ja = jb;
            jb = jl;
            ya = yb;
            yb = yl;
        }
        if (l < mu) continue;
        double P;
        if (l == mu)
            P = pmm;
        else if (l == mu + 1)
            P = x * (2 * mu + 1) * pmm;
        else
            P = ((2 * l - 1) * x * p1 - (l + mu - 1) * p2) / (l - mu);
        p2 = p1;
        p1 = P;
        HL[leg_index(l, mu)] = zkind == 2 ? make_double2(jl * P, 0.0)
                                          : cscale(make_double2(jl, zkind == 1 ? -yl : yl), P);
    }
    if (mu <= 2 * lmax) {
        double sn, cs;
        sincos(mu * atan2(dy, dx), &sn, &cs);
        eph[mu] = make_double2(cs, sn);
    }
}

// Compact translation table (device view, in global memory or LDS):
//   meta[t] = {first term, nA | nB << 8 | (mu + 64) << 16} for tuple
//   t = r(n,m) rmax + r(v,u); the tuple's nA A-terms and then nB B-terms are
//   coef[first ..] with HL indices lidx[first ..].
struct TabDev {
    const int2* meta;
    const double2* coef;
    const unsigned short* lidx;
};

// LDS bytes of the compact table and its layout (coef, meta, lidx).
__host__ __device__ inline size_t tab_lds_bytes(int rmax, int nterms) {
    return (size_t)nterms * sizeof(double2) + (size_t)rmax * rmax * sizeof(int2) +
           (((size_t)nterms * sizeof(unsigned short) + 15) & ~size_t(15));
}

// Copy the table into LDS at `dst` (16-byte aligned); all threads of the
// block take part; the caller synchronizes before use.
__device__ inline TabDev stage_table(const TabDev& g, int rmax, int nterms, char* dst) {
    double2* c = (double2*)dst;
    int2* m = (int2*)(c + nterms);
    unsigned short* li = (unsigned short*)(m + rmax * rmax);
    for (int u = threadIdx.x; u < nterms; u += blockDim.x) {
        c[u] = g.coef[u];
        li[u] = g.lidx[u];
    }
    for (int u = threadIdx.x; u < rmax * rmax; u += blockDim.x) m[u] = g.meta[u];
    return TabDev{m, c, li};
}

// X_A and X_B of tuple t for one pair.
__device__ inline void tuple_xab(const TabDev& tab, int t, const double2* hl, const double2* eph, double2& xa,
                                 double2& xb) {
    const int2 m = tab.meta[t];
    const int nA = m.y & 255, nB = (m.y >> 8) & 255, mu = (m.y >> 16) - 64;
    const double2* c = tab.coef + m.x;
    const unsigned short* li = tab.lidx + m.x;
    xa = make_double2(0, 0);
    xb = make_double2(0, 0);
    for (int q = 0; q < nA; q++) cfma(xa, c[q], hl[li[q]]);
    for (int q = nA; q < nA + nB; q++) cfma(xb, c[q], hl[li[q]]);
    double2 e = eph[mu < 0 ? -mu : mu];
    if (mu < 0) e.y = -e.y;
    xa = cmul(e, xa);
    xb = cmul(e, xb);
}

}  // namespace mpa
