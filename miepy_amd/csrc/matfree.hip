// Matrix-free (I + T) x for sphere clusters: the translation blocks are
// recomputed on every product and never stored.
//
// Same operator as the assembled matrix of assemble.hip (reference
// cpp/src/interactions.cpp:370-415, cpp/src/vsh_translation.cpp:161-237),
// applied as
//   y_i[a, r] = x_i[a, r] + sum_{j != i} sum_s ( X_A^{ij}[r,s] z_j[a,s] + X_B^{ij}[r,s] z_j[1-a,s] ),
//   z_j[b, s] = mie_j[b, v(s)] x_j[b, s],
// because T_ij[(a,r),(b,s)] = X_{(a+b)%2}[r,s] mie_j[b,v(s)].  Memory is O(n)
// instead of O(n^2): no assembly pass, no matrix in HBM, and clusters whose
// matrix would not fit in HBM can still be solved.  The kernel is bound by
// FP64 VALU and LDS issue, not by HBM.
//
// Work decomposition: a workgroup owns block row i and a chunk of column
// particles and walks the chunk in batches of JB pairs.  Per batch, one
// thread per pair forms the pair geometry (cos theta, sin theta, k|d|,
// sin/cos k|d|, e^{i phi}); then one thread per (pair, Legendre order mu)
// forms the products HL[p][mu] = h_p(k|d|) P_p^mu(cos theta) and
// e^{i mu phi}, and the batch's z_j, all in LDS.  Phase 2 gives each lane
// one (row mode r, pair): lanes that share r sit in the same wave, so the
// coefficient table is read as a broadcast and the q-loops of a wave
// diverge over only a few tuples; the per-pair LDS arrays have odd strides,
// so the lanes' reads hit distinct banks.  A lane walks the column modes s,
// forms X_A[r,s] and X_B[r,s] and accumulates both polarizations of y_i for
// SG right-hand sides in registers.  At the end the JB partial rows meet in
// LDS in a fixed order and the chunk's partial row is written; a second
// kernel adds the chunks in a fixed order plus the identity term, so the
// result is deterministic.

#include <algorithm>
#include <utility>

#include "matfree.hpp"
#include "translation.hpp"

namespace mpa {
namespace {

// Workgroup id -> work id so that consecutive work ids run on one XCD:
// workgroups are dealt round-robin to the 8 XCDs (id % 8), so XCD x gets the
// contiguous work range [x G8 / 8, (x + 1) G8 / 8) of the first G8 = G & ~7
// workgroups (the tail keeps its id).  With work id = chunk * groups + row
// group, the row groups that read the same column chunk's operands share an
// XCD's L2.
__device__ __forceinline__ int xcd_swizzle(int b, int G) {
    const int full = G & ~7;
    return b < full ? (b & 7) * (full >> 3) + (b >> 3) : b;
}


constexpr int kJB = 16;   // most pairs per batch
constexpr int kSMax = 4;  // right-hand sides per pass (largest SG)

__host__ __device__ inline int pairs_per_batch(int rmax) {
    const int jb = kThreads / rmax;
    return jb > kJB ? kJB : jb;
}
constexpr int kTabLdsLmax = 3;  // largest lmax whose table is staged in LDS

// Odd per-pair stride of the HL array (conflict-free 16-byte LDS reads).
__host__ __device__ inline int hl_stride(int lmax) { return hl_count(lmax) | 1; }

// Right-hand sides per pass for S right-hand sides.
inline int sg_for(int S) { return S == 1 ? 1 : (S == 2 ? 2 : kSMax); }

// Per-pair geometry staged for phase 1.
struct PairGeo {
    double x, s, z, sz, cz, ex, ey, pad;
};

// LDS bytes of one workgroup: pair geometry, HL, e^{i mu phi}, z /
// reduction buffer, then (lmax <= kTabLdsLmax) the compact coefficient table.
inline size_t lds_bytes(int lmax, int SG, int nterms) {
    const int rmax = lmax_to_rmax(lmax);
    const int JB = pairs_per_batch(rmax);
    const size_t b = sizeof(PairGeo) * JB +
                     sizeof(double2) * ((size_t)JB * hl_stride(lmax) + (size_t)JB * (2 * lmax + 1) +
                                        (size_t)SG * 2 * JB * rmax);
    return b + (lmax <= kTabLdsLmax ? tab_lds_bytes(rmax, nterms) : 0);
}

// Chunks of column particles per block row: enough workgroups to fill the
// chip for small clusters, one chunk per row for large ones.
inline int chunks_for(int N) {
    int c = (2048 + N - 1) / N;
    if (c > 64) c = 64;
    if (c > N) c = N;
    return c < 1 ? 1 : c;
}

__device__ inline bool job_done(const SysDev& d) {
    for (int s = 0; s < d.S; s++)
        if (!d.scal[s].done) return false;
    return true;
}

// Pair geometry for d = r_i - r_j: cos/sin theta, k|d| with its sine and
// cosine, and e^{i phi} = (dx + i dy) / rho (1 on the z axis, where every
// P_p^mu with mu > 0 vanishes).
__device__ inline PairGeo pair_geo(double dx, double dy, double dz, double k) {
    PairGeo g;
    const double rad = sqrt(dx * dx + dy * dy + dz * dz);
    g.x = dz / rad;
    g.s = sqrt(fmax(0.0, (1.0 - g.x) * (1.0 + g.x)));
    g.z = k * rad;
    sincos(g.z, &g.sz, &g.cz);
    const double rho = sqrt(dx * dx + dy * dy);
    g.ex = rho > 0 ? dx / rho : 1.0;
    g.ey = rho > 0 ? dy / rho : 0.0;
    g.pad = 0;
    return g;
}

// Thread (pair, mu): HL[leg_index(l, mu)] = h_l(z) P_l^mu(x), l = mu..pmax,
// with the recurrences of pair_hl, and eph[mu] = (e^{i phi})^mu.
__device__ inline void pair_hl_geo(double2* HL, double2* eph, int mu, int lmax, const PairGeo& g) {
    const int pmax = 2 * lmax + 1;
    double pmm = 1.0;
    for (int i = 1; i <= mu; i++) pmm *= (2 * i - 1) * g.s;
    const double rz = 1.0 / g.z;
    double ja = g.sz * rz, ya = -g.cz * rz;
    double jb = ja * rz + ya, yb = ya * rz - ja;
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
            jl = (2 * l - 1) * rz * jb - ja;
            yl = (2 * l - 1) * rz * yb - ya;
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
            P = g.x * (2 * mu + 1) * pmm;
        else
            P = ((2 * l - 1) * g.x * p1 - (l + mu - 1) * p2) / (l - mu);
        p2 = p1;
        p1 = P;
        HL[leg_index(l, mu)] = cscale(make_double2(jl, yl), P);
    }
    if (mu <= 2 * lmax) {
        double2 e = make_double2(1.0, 0.0);
        const double2 e1 = make_double2(g.ex, g.ey);
        for (int i = 0; i < mu; i++) e = cmul(e, e1);
        eph[mu] = e;
    }
}

// Partial rows: part[chunk][rhs][row].  SG right-hand sides per pass;
// TAB_LDS: stage the table in LDS.
template <int SG, bool TAB_LDS>
__global__ __launch_bounds__(kThreads) void k_matfree(MfJob J, const SysDev* __restrict__ sys, int nchunks, int chunk,
                                                      TabDev gtab, int nterms) {
    extern __shared__ double2 lds[];
    if (J.skip >= 0 && job_done(sys[J.skip])) return;
    const int lmax = J.lmax;
    const int rmax = lmax_to_rmax(lmax);
    const int bs = 2 * rmax;
    const int n = bs * J.N;
    const int pmax = 2 * lmax + 1;
    const int Lp = hl_stride(lmax);
    const int nmu = 2 * lmax + 1;
    const int JB = pairs_per_batch(rmax);
    const int zs = JB * rmax;  // z / reduction stride per (rhs, polarization)
    const double2* zpre = J.zpre;  // z = Tm x (general T-matrices) or null
    PairGeo* geo = (PairGeo*)lds;
    double2* HL = (double2*)(geo + JB);
    double2* eph = HL + JB * Lp;
    double2* zr = eph + JB * nmu;
    TabDev tab = gtab;
    if (TAB_LDS) tab = stage_table(gtab, rmax, nterms, (char*)(zr + SG * 2 * zs));

    const int i = blockIdx.x / nchunks;
    const int c = blockIdx.x % nchunks;
    const int jbeg = c * chunk, jend = min(J.N, jbeg + chunk);
    const int t = threadIdx.x;
    const int r = t / JB, jme = t - r * JB;  // lanes sharing r are adjacent
    const bool worker = r < rmax;
    const double xi = J.pos[3 * i], yi = J.pos[3 * i + 1], zi = J.pos[3 * i + 2];

    for (int s0 = 0; s0 < J.S; s0 += SG) {
        const int sg = min(SG, J.S - s0);
        double2 acc[SG][2];
#pragma unroll
        for (int q = 0; q < SG; q++) acc[q][0] = acc[q][1] = make_double2(0, 0);
        for (int j0 = jbeg; j0 < jend; j0 += JB) {
            const int jc = min(JB, jend - j0);
            __syncthreads();
            if (t < jc) {
                const int j = j0 + t;
                geo[t] = pair_geo(xi - J.pos[3 * j], yi - J.pos[3 * j + 1], zi - J.pos[3 * j + 2], J.k);
            }
            for (int u = t; u < jc * sg * bs; u += kThreads) {
                const int jj = u / (sg * bs), rest = u - jj * (sg * bs);
                const int q = rest / bs, e = rest - q * bs;
                const int b = e >= rmax ? 1 : 0, s = e - b * rmax;
                const int v = (int)sqrtf((float)(s + 1) + 0.5f);
                const int j = j0 + jj;
                const size_t xo = (size_t)(s0 + q) * n + (size_t)j * bs + e;
                zr[(q * 2 + b) * zs + jj * rmax + s] = zpre ? zpre[xo] : cmul(J.mie[((size_t)j * 2 + b) * lmax + v - 1], J.x[xo]);
            }
            __syncthreads();
            for (int u = t; u < jc * (pmax + 1); u += kThreads) {
                const int jj = u / (pmax + 1), mu = u - jj * (pmax + 1);
                if (j0 + jj != i) pair_hl_geo(HL + jj * Lp, eph + jj * nmu, mu, lmax, geo[jj]);
            }
            __syncthreads();
            if (worker && jme < jc && j0 + jme != i) {
                const double2* hl = HL + jme * Lp;
                const double2* ep = eph + jme * nmu;
                const double2* zj = zr + jme * rmax;
                for (int s = 0; s < rmax; s++) {
                    double2 xa, xb;
                    tuple_xab(tab, r * rmax + s, hl, ep, xa, xb);
#pragma unroll
                    for (int q = 0; q < SG; q++) {
                        if (q < sg) {
                            const double2 z0 = zj[(q * 2) * zs + s], z1 = zj[(q * 2 + 1) * zs + s];
                            cfma(acc[q][0], xa, z0);
                            cfma(acc[q][0], xb, z1);
                            cfma(acc[q][1], xa, z1);
                            cfma(acc[q][1], xb, z0);
                        }
                    }
                }
            }
        }
        // the JB partial rows meet in LDS and are summed in a fixed order
        __syncthreads();
        if (worker) {
#pragma unroll
            for (int q = 0; q < SG; q++)
                if (q < sg) {
                    zr[(q * 2) * zs + jme * rmax + r] = acc[q][0];
                    zr[(q * 2 + 1) * zs + jme * rmax + r] = acc[q][1];
                }
        }
        __syncthreads();
        for (int u = t; u < sg * bs; u += kThreads) {
            const int q = u / bs, e = u - q * bs;
            const int a = e >= rmax ? 1 : 0, rr = e - a * rmax;
            double2 sum = make_double2(0, 0);
            for (int jj = 0; jj < JB; jj++) sum = cadd(sum, zr[(q * 2 + a) * zs + jj * rmax + rr]);
            J.part[((size_t)c * J.S + s0 + q) * n + (size_t)i * bs + e] = sum;
        }
    }
}

// ---- factorized operator (lmax <= 3, one lane per row particle) -------------
//
// With HLE_ij[p(p+1) + mu] = h_p(k|d|) P_p^|mu|(cos theta) e^{i mu phi} for
// signed mu, every block row is a contraction over entries e = (r, kind, mu, p):
//   y_i[a, r] = sum_j sum_{e of row r} HLE_ij[p, mu] W_j[e][a],
//   W_j[e][a] = sum over the table terms (s, c) of e of c z_j[a', s]
// (a' = a for A entries, 1 - a for B entries), because each q-term of
// X_A[r,s] / X_B[r,s] is a coefficient times one HLE value.  W_j depends on
// x but not on i, so it is formed once per matvec (O(N) work, k_matfree_w).
// In the pair kernel a lane owns one row particle i and walks the column
// particles j of its chunk: it forms its pair's HLE in registers (no LDS),
// and the fully unrolled entry loop multiplies them by W_j, whose address is
// the same for the whole wave and is read through the scalar cache; the
// accumulators of the 2 rmax rows of y_i stay in registers.  The entry list
// is enumerated at compile time in the order tables.cpp stores W's terms.

typedef float f16v __attribute__((ext_vector_type(16)));

// element of the real expansion X' of a complex B operand (see gemv.hip) for
// K parity `odd` (the Im part of the A value) and output part `part`
template <typename Rt>
__device__ inline Rt xprime(Rt re, Rt im, int odd, int part) {
    return odd ? (part ? re : -im) : (part ? im : re);
}

constexpr int kFacMaxLmax = 3;

// Matrix-core contraction variants (mpa_autotune_matfree): [c64, c128].
// c64: 0 one accumulator chain per tile, 1 two (K steps alternate), 2 and 3
// one chain with the next column's special functions interleaved with the
// MFMAs (8 or 12 VALU instructions per MFMA; FP16-split kernel; the FP32
// kernel treats 2, 3 as 0); c128: 0 and 1 the slot-mapped kernel (k_matfree_mf64s) with one or two
// accumulator chains (one is the faster at rand1000: 0.34 ms per product against 0.43 ms for the
// best per-lane-select variant; at lmax 4, 0 is the single-buffer form k_matfree_mf64k at two
// workgroups per CU), 2 and 3 the per-lane-select kernel (k_matfree_mf64), two or one.
// complex64: 4, 5 = column pairs (lmax <= 3), KS 1 and 2; 6 = sources sharing the special functions,
// many jobs per launch (k_matfree_mf16s); 7 (default) = per job: 6 for three or more sources or
// lmax 4, else 4.  The choice depends on the job alone, never on what shares its launch.
int g_mf_variant[2] = {7, 0};  // FP64: slot-mapped KS 1
constexpr int kMfVariants[2] = {8, 4};
constexpr int kMfShared = 6, kMfAuto = 7;

// Whether a complex64 job runs the shared kernel under the current variant choice.
bool mf_shared_for(int S, int lmax) {
    if (g_mf_variant[0] == kMfShared) return true;
    return g_mf_variant[0] == kMfAuto && (S >= 3 || lmax == 4);
}
constexpr int kFacMinN = 128;  // smaller clusters keep the LDS kernel (too few lanes)

constexpr int fac_n_of(int r) {
    int n = 1;
    while ((n + 1) * (n + 1) - 1 <= r) n++;
    return n;
}
constexpr int fac_m_of(int r) { return r - (fac_n_of(r) * fac_n_of(r) + fac_n_of(r) - 1); }
constexpr int fac_min(int a, int b) { return a < b ? a : b; }
constexpr int fac_abs(int a) { return a < 0 ? -a : a; }

template <int L>
constexpr bool fac_has_entry(int kind, int r, int mu, int p) {
    const int n = fac_n_of(r), m = fac_m_of(r);
    for (int v = 1; v <= L; v++)
        for (int u = -v; u <= v; u++) {
            if (u - m != mu) continue;
            const int qa = fac_min(fac_min(n, v), (n + v - fac_abs(mu)) / 2);
            const int qb = fac_min(fac_min(n, v), (n + v + 1 - fac_abs(mu)) / 2);
            if (kind == 0) {
                for (int q = 0; q <= qa; q++)
                    if (n + v - 2 * q == p) return true;
            } else {
                for (int q = 1; q <= qb; q++)
                    if (n + v - 2 * q + 1 == p) return true;
            }
        }
    return false;
}

template <int L>
struct FacEntries {
    static constexpr int R = L * (L + 2);
    static constexpr int MAXE = 2 * R * (4 * L + 1) * (2 * L + 1);
    int count;
    int r[MAXE];
    int h[MAXE];
    constexpr FacEntries() : count(0), r(), h() {
        for (int rr = 0; rr < R; rr++)
            for (int kind = 0; kind < 2; kind++)
                for (int mu = -2 * L; mu <= 2 * L; mu++)
                    for (int p = fac_abs(mu); p <= 2 * L; p++)
                        if (fac_has_entry<L>(kind, rr, mu, p)) {
                            r[count] = rr;
                            h[count] = p * (p + 1) + mu;
                            count++;
                        }
    }
};
template <int L>
constexpr FacEntries<L> kFacEntries{};

template <int L, int I>
__device__ __forceinline__ void fac_step(const double2 (&hle)[(2 * L + 1) * (2 * L + 1)],
                                        const double2* __restrict__ W, double2 (&acc)[2][L * (L + 2)]) {
    constexpr int r = kFacEntries<L>.r[I];
    constexpr int hi = kFacEntries<L>.h[I];
    const double2 w0 = W[2 * I], w1 = W[2 * I + 1];
    cfma(acc[0][r], hle[hi], w0);
    cfma(acc[1][r], hle[hi], w1);
}

template <int L, int... Is>
__device__ __forceinline__ void fac_entries(const double2 (&hle)[(2 * L + 1) * (2 * L + 1)],
                                           const double2* __restrict__ W, double2 (&acc)[2][L * (L + 2)],
                                           std::integer_sequence<int, Is...>) {
    (fac_step<L, Is>(hle, W, acc), ...);
}

// W[s][j][e] = (w0, w1) for every RHS s, particle j and entry e.
__global__ __launch_bounds__(kThreads) void k_matfree_w(MfJob J, const SysDev* __restrict__ sys, int ne,
                                                        const int2* __restrict__ wmeta, const int* __restrict__ wsrc,
                                                        const double2* __restrict__ wcoef, double2* __restrict__ W) {
    if (J.skip >= 0 && job_done(sys[J.skip])) return;
    const int lmax = J.lmax, rmax = lmax_to_rmax(lmax), bs = 2 * rmax;
    const size_t n = (size_t)bs * J.N;
    const size_t total = (size_t)J.S * J.N * ne;
    for (size_t idx = (size_t)blockIdx.x * kThreads + threadIdx.x; idx < total; idx += (size_t)gridDim.x * kThreads) {
        const size_t sj = idx / ne;
        const int e = (int)(idx - sj * ne);
        const size_t s = sj / J.N, j = sj - s * J.N;
        const int2 m = wmeta[e];
        const int cnt = m.y & 0xffff, kind = m.y >> 16;
        const double2* xj = J.x + s * n + j * bs;
        const double2* zj = J.zpre ? J.zpre + s * n + j * bs : nullptr;
        double2 w0 = make_double2(0, 0), w1 = make_double2(0, 0);
        for (int k = m.x; k < m.x + cnt; k++) {
            const int sidx = wsrc[k];
            const int v = (int)sqrtf((float)(sidx + 1) + 0.5f);
            // z_j[b, s] = mie_j[b, v] x_j[b, s] (spheres) or (Tm_j x_j)[b, s]
            const double2 z0 = zj ? zj[sidx] : cmul(J.mie[(j * 2 + 0) * lmax + v - 1], xj[sidx]);
            const double2 z1 = zj ? zj[rmax + sidx] : cmul(J.mie[(j * 2 + 1) * lmax + v - 1], xj[rmax + sidx]);
            cfma(w0, wcoef[k], kind ? z1 : z0);
            cfma(w1, wcoef[k], kind ? z0 : z1);
        }
        W[2 * idx] = w0;
        W[2 * idx + 1] = w1;
    }
}

// Partial rows part[chunk][s][row] of y = T x; blockIdx.y is the RHS.
template <int L>
__global__ __launch_bounds__(kThreads) void k_matfree_fac(MfJob J, const SysDev* __restrict__ sys, int nchunks,
                                                          int chunk, const double2* __restrict__ Wall) {
    constexpr int R = L * (L + 2);
    constexpr int PM = 2 * L;
    constexpr int NE = kFacEntries<L>.count;
    if (J.skip >= 0 && job_done(sys[J.skip])) return;
    const int N = J.N;
    const int s = blockIdx.y;
    const int c = blockIdx.x % nchunks;
    const int i = (blockIdx.x / nchunks) * kThreads + threadIdx.x;
    const bool live = i < N;
    const int ii = live ? i : N - 1;
    const double xi = J.pos[3 * ii], yi = J.pos[3 * ii + 1], zi = J.pos[3 * ii + 2];
    const double2* W = Wall + (size_t)s * N * NE * 2;
    double2 acc[2][R];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int q = 0; q < R; q++) acc[a][q] = make_double2(0, 0);
    const int jbeg = c * chunk, jend = min(N, jbeg + chunk);
    for (int j = jbeg; j < jend; j++) {
        if (j == ii) continue;
        const double dx = xi - J.pos[3 * j], dy = yi - J.pos[3 * j + 1], dz = zi - J.pos[3 * j + 2];
        const double rad = sqrt(dx * dx + dy * dy + dz * dz);
        const double ct = dz / rad, st = sqrt(fmax(0.0, (1.0 - ct) * (1.0 + ct)));
        const double zz = J.k * rad, rz = 1.0 / zz;
        double sz, cz;
        sincos(zz, &sz, &cz);
        const double rho = sqrt(dx * dx + dy * dy);
        const double2 e1 = rho > 0 ? make_double2(dx / rho, dy / rho) : make_double2(1.0, 0.0);
        double2 h[PM + 1];
        h[0] = make_double2(sz * rz, -cz * rz);
        h[1] = make_double2(h[0].x * rz + h[0].y, h[0].y * rz - h[0].x);
#pragma unroll
        for (int l = 2; l <= PM; l++)
            h[l] = make_double2((2 * l - 1) * rz * h[l - 1].x - h[l - 2].x, (2 * l - 1) * rz * h[l - 1].y - h[l - 2].y);
        double2 e[PM + 1];
        e[0] = make_double2(1.0, 0.0);
#pragma unroll
        for (int m = 1; m <= PM; m++) e[m] = cmul(e[m - 1], e1);
        double2 hle[(2 * L + 1) * (2 * L + 1)];
#pragma unroll
        for (int m = 0; m <= PM; m++) {
            double pmm = 1.0;
#pragma unroll
            for (int q = 1; q <= PM; q++)
                if (q <= m) pmm *= (2 * q - 1) * st;
            double p2 = 0.0, p1 = 0.0;
#pragma unroll
            for (int l = 0; l <= PM; l++) {
                if (l < m) continue;
                const double P = (l == m)       ? pmm
                                 : (l == m + 1) ? ct * (2 * m + 1) * pmm
                                                : ((2 * l - 1) * ct * p1 - (l + m - 1) * p2) * (1.0 / (l - m));  // l - m constant: a multiply
                p2 = p1;
                p1 = P;
                const double2 hp = cscale(h[l], P);
                hle[l * (l + 1) + m] = cmul(hp, e[m]);
                if (m > 0) hle[l * (l + 1) - m] = cmul(hp, cconj(e[m]));
            }
        }
        fac_entries<L>(hle, W + (size_t)j * NE * 2, acc, std::make_integer_sequence<int, NE>{});
    }
    if (live) {
        const size_t n = (size_t)2 * R * N;
        double2* out = J.part + ((size_t)c * J.S + s) * n + (size_t)i * 2 * R;
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
            for (int q = 0; q < R; q++) out[a * R + q] = acc[a][q];
    }
}

// FP64 kernel choice for lmax <= 3 (MPA_OPT_MATFREE_FP64_KERNEL): measured on
// MI355X at n = 30000, one RHS: matrix cores 0.45 ms (default), LDS kernel
// 1.17 ms, factorized register-resident kernel 1.57 ms (one wave per SIMD at
// 256 VGPRs leaves the scalar loads of W_j uncovered).
inline bool fac_path(int N, int lmax) {
    return lmax <= kFacMaxLmax && N >= kFacMinN && mpa_option(MPA_OPT_MATFREE_FP64_KERNEL) == 2;
}
constexpr int kMf64MaxLmax = 5;  // FP64 matrix-core kernels: lmax <= 4 two LDS buffers (128 KB at lmax 4),
                                 // lmax 5 one (136 KB)
inline bool mf64_path(int lmax) { return lmax <= kMf64MaxLmax && mpa_option(MPA_OPT_MATFREE_FP64_KERNEL) == 0; }
inline bool fac_possible(int N, int lmax) { return lmax <= kFacMaxLmax && N >= kFacMinN; }

// ---- complex64 operator on the FP32 matrix cores (mixed-precision inner solves)
//
// The same contraction as the factorized FP64 operator, grouped by HLE index:
//   y_i[a,r] = sum_j sum_h HLE_ij[h] Wm_j[h][a,r],  Wm_j[h][a,r] = sum of W_j over
// the (at most two: A and B) entries of (r, h),
// i.e. per column particle j a complex product [rows i] x [nh] times [nh] x [2 rmax].
// With the real expansion used by the matvec GEMM (a complex HLE becomes the
// two K positions (Re, Im); Wm the matching B columns X'), each j is nh
// v_mfma_f32_32x32x2f32 steps per 16 complex output columns for 32 row
// particles.  Lanes l and l + 32 of a wave own the same row particle and
// supply its Re and Im HLE values (computed in FP64 by both, rounded to FP32);
// Wm_j (FP32, formed once per matvec) is staged in LDS per workgroup and
// read as the B operand; the accumulators stay in registers across the
// workgroup's chunk of column particles.

// Wm[s][j][h][cc], cc = a rmax + r: thread per entry.  T: element type of x
// and Wm (float: complex64 operator, double: FP64 matrix-core operator).
template <typename T>
__global__ __launch_bounds__(kThreads) void k_matfree_wm(MfJob J, const SysDev* __restrict__ sys, int nh,
                                                         const int2* __restrict__ hmap,
                                                         const int2* __restrict__ wmeta, const int* __restrict__ wsrc,
                                                         const double2* __restrict__ wcoef, cplx_t<T>* __restrict__ Wm) {
    using C = cplx_t<T>;
    if (J.skip >= 0 && job_done(sys[J.skip])) return;
    const int lmax = J.lmax, rmax = lmax_to_rmax(lmax), bs = 2 * rmax;
    const size_t n = (size_t)bs * J.N;
    const size_t total = (size_t)J.S * J.N * nh * bs;
    const C* X = (const C*)J.x;
    for (size_t idx = (size_t)blockIdx.x * kThreads + threadIdx.x; idx < total; idx += (size_t)gridDim.x * kThreads) {
        const int cc = (int)(idx % bs);
        const size_t rest = idx / bs;
        const int h = (int)(rest % nh);
        const size_t sj = rest / nh;
        const size_t s = sj / J.N, j = sj - s * J.N;
        const int a = cc >= rmax ? 1 : 0, r = cc - a * rmax;
        const int2 es = hmap[(size_t)r * nh + h];
        const C* xj = X + s * n + j * bs;
        double2 w = make_double2(0, 0);
        for (int kind = 0; kind < 2; kind++) {
            const int e = kind ? es.y : es.x;
            if (e < 0) continue;
            const int2 m = wmeta[e];
            const int cnt = m.y & 0xffff;
            const int src_pol = kind ? 1 - a : a;  // B entries couple the other polarization
            for (int k = m.x; k < m.x + cnt; k++) {
                const int sidx = wsrc[k];
                const int v = (int)sqrtf((float)(sidx + 1) + 0.5f);
                const C xv = xj[src_pol * rmax + sidx];
                const double2 z = J.zpre ? J.zpre[s * n + j * bs + src_pol * rmax + sidx]
                                         : cmul(J.mie[(j * 2 + src_pol) * lmax + v - 1], make_double2(xv.x, xv.y));
                cfma(w, wcoef[k], z);
            }
        }
        Wm[idx] = cmake<T>((T)w.x, (T)w.y);
    }
}

// B operands of the complex64 kernel, expanded per lane: Wx[s][j][tile][q][lane]
// (4 floats: h = 4q + e, e = 0..3) = X'[k = lane / 32][col = lane % 32] of
// the 32-column tile, i.e. each lane's B value of the MFMA step of HLE
// index h, zero where h >= nh or the complex column tile * 16 + col / 2 is
// past 2 rmax.  One thread per complex (h, column) writes its four lanes.
template <int NT, int NQ>
__global__ __launch_bounds__(kThreads) void k_matfree_wx(MfJob J, const SysDev* __restrict__ sys, int nh,
                                                         const int2* __restrict__ hmap,
                                                         const int2* __restrict__ wmeta, const int* __restrict__ wsrc,
                                                         const double2* __restrict__ wcoef, float* __restrict__ Wx) {
    if (J.skip >= 0 && job_done(sys[J.skip])) return;
    const int lmax = J.lmax, rmax = lmax_to_rmax(lmax), bs = 2 * rmax;
    const size_t n = (size_t)bs * J.N;
    constexpr int NCP = NT * 16, NHP = NQ * 4;
    const size_t total = (size_t)J.S * J.N * NHP * NCP;
    const float2* X = (const float2*)J.x;
    for (size_t idx = (size_t)blockIdx.x * kThreads + threadIdx.x; idx < total; idx += (size_t)gridDim.x * kThreads) {
        const int cc = (int)(idx % NCP);
        const size_t rest = idx / NCP;
        const int h = (int)(rest % NHP);
        const size_t sj = rest / NHP;
        const size_t s = sj / J.N, j = sj - s * J.N;
        double2 w = make_double2(0, 0);
        if (h < nh && cc < bs) {
            const int a = cc >= rmax ? 1 : 0, r = cc - a * rmax;
            const int2 es = hmap[(size_t)r * nh + h];
            const float2* xj = X + s * n + j * bs;
            for (int kind = 0; kind < 2; kind++) {
                const int e = kind ? es.y : es.x;
                if (e < 0) continue;
                const int2 m = wmeta[e];
                const int cnt = m.y & 0xffff;
                const int src_pol = kind ? 1 - a : a;  // B entries couple the other polarization
                for (int k = m.x; k < m.x + cnt; k++) {
                    const int sidx = wsrc[k];
                    const int v = (int)sqrtf((float)(sidx + 1) + 0.5f);
                    const float2 xv = xj[src_pol * rmax + sidx];
                    cfma(w, wcoef[k], cmul(J.mie[(j * 2 + src_pol) * lmax + v - 1], make_double2(xv.x, xv.y)));
                }
            }
        }
        const float re = (float)w.x, im = (float)w.y;
        const int tile = cc / 16, c2 = 2 * (cc % 16);
        float* dst = Wx + (sj * NT + tile) * (size_t)NQ * 256 + (size_t)(h / 4) * 256 + (h % 4);
        dst[4 * c2] = re;             // k = 0, part 0
        dst[4 * (c2 + 1)] = im;       // k = 0, part 1
        dst[4 * (32 + c2)] = -im;     // k = 1, part 0
        dst[4 * (32 + c2 + 1)] = re;  // k = 1, part 1
    }
}

typedef float f4v __attribute__((ext_vector_type(4)));

// sin and cos of 0 <= x < 1e6 without branches: Cody-Waite reduction by pi/2
// (two-term constant, fused multiply-adds), Taylor polynomials through r^13
// and r^14 on |r| <= pi/4 (truncation below 3e-14), quadrant by selects.
__device__ __forceinline__ void sincos_reduced(double x, double* sp, double* cp) {
    const double q = rint(x * 0.63661977236758134308);
    double r = fma(-q, 1.5707963267948966, x);
    r = fma(-q, 6.123233995736766e-17, r);
    const double r2 = r * r;
    double ps = -1.0 / 6227020800.0;
    ps = fma(ps, r2, 1.0 / 39916800.0);
    ps = fma(ps, r2, -1.0 / 362880.0);
    ps = fma(ps, r2, 1.0 / 5040.0);
    ps = fma(ps, r2, -1.0 / 120.0);
    ps = fma(ps, r2, 1.0 / 6.0);
    const double S = fma(-ps * r2, r, r);
    double pc = 1.0 / 87178291200.0;
    pc = fma(pc, r2, -1.0 / 479001600.0);
    pc = fma(pc, r2, 1.0 / 3628800.0);
    pc = fma(pc, r2, -1.0 / 40320.0);
    pc = fma(pc, r2, 1.0 / 720.0);
    pc = fma(pc, r2, -1.0 / 24.0);
    pc = fma(pc, r2, 0.5);
    const double C = fma(-pc, r2, 1.0);
    const int quad = (int)q & 3;
    const double s0 = (quad & 1) ? C : S, c0 = (quad & 1) ? S : C;
    *sp = (quad & 2) ? -s0 : s0;
    *cp = ((quad + 1) & 2) ? -c0 : c0;
}

// The real part this lane supplies (Re for kh = 0, Im for kh = 1) of the
// HLE values of pair (i, j), FP64 rounded to FP32; af[h] = 0 for h >= nh or
// when !on.  With hp = h_p(kd) = (a, b) and e^{i m phi} = (c, s):
//   Re HLE[p, +-m] = P (a c -+ b s),  Im HLE[p, +-m] = P (b c +- a s),
// so each lane forms (alpha c + beta s'), (alpha c - beta s') with
// (alpha, beta, s') = (a, b, -s) or (b, a, s).
template <int L, int NHP>
__device__ __forceinline__ void hle_lane(double xi, double yi, double zi, const double* __restrict__ pos, int j,
                                         double k, int kh, bool on, float (&af)[NHP]) {
    constexpr int PM = 2 * L;
    constexpr int NH = (2 * L + 1) * (2 * L + 1);
#pragma unroll
    for (int t = NH; t < NHP; t++) af[t] = 0.f;
    const double dx = xi - pos[3 * j], dy = yi - pos[3 * j + 1], dz = zi - pos[3 * j + 2];
    const double rho2 = dx * dx + dy * dy;
    const double rad = sqrt(rho2 + dz * dz);
    const double rinv = 1.0 / rad;
    const double ct = dz * rinv, st = sqrt(fmax(0.0, (1.0 - ct) * (1.0 + ct)));
    const double zz = k * rad, rz = rinv / k;
    double sz, cz;
    sincos_reduced(zz, &sz, &cz);
    const double rho = sqrt(rho2);
    const double rhoinv = 1.0 / fmax(rho, 1e-300);
    const double2 e1 = rho > 0 ? make_double2(dx * rhoinv, dy * rhoinv) : make_double2(1.0, 0.0);
    double2 hh[PM + 1];
    hh[0] = make_double2(sz * rz, -cz * rz);
    hh[1] = make_double2(hh[0].x * rz + hh[0].y, hh[0].y * rz - hh[0].x);
#pragma unroll
    for (int l = 2; l <= PM; l++)
        hh[l] = make_double2((2 * l - 1) * rz * hh[l - 1].x - hh[l - 2].x,
                             (2 * l - 1) * rz * hh[l - 1].y - hh[l - 2].y);
    double al[PM + 1], be[PM + 1];
#pragma unroll
    for (int l = 0; l <= PM; l++) {
        al[l] = kh ? hh[l].y : hh[l].x;
        be[l] = kh ? hh[l].x : hh[l].y;
    }
    double2 e = make_double2(1.0, 0.0);
    double pmm = 1.0;
#pragma unroll
    for (int m = 0; m <= PM; m++) {
        if (m > 0) {
            e = cmul(e, e1);
            pmm *= (2 * m - 1) * st;
        }
        const double sp = kh ? e.y : -e.y;
        double p2 = 0.0, p1 = 0.0;
#pragma unroll
        for (int l = 0; l <= PM; l++) {
            if (l < m) continue;
            const double P = (l == m)       ? pmm
                             : (l == m + 1) ? ct * (2 * m + 1) * pmm
                                            : ((2 * l - 1) * ct * p1 - (l + m - 1) * p2) * (1.0 / (l - m));
            p2 = p1;
            p1 = P;
            const double ac = al[l] * e.x, bs = be[l] * sp;
            af[l * (l + 1) + m] = on ? (float)(P * (ac + bs)) : 0.f;
            if (m > 0) af[l * (l + 1) - m] = on ? (float)(P * (ac - bs)) : 0.f;
        }
    }
}

template <int PER, int PFN>
__device__ __forceinline__ void wx_fetch(f4v (&pf)[PFN], const f4v* __restrict__ src, int tid) {
#pragma unroll
    for (int q = 0; q < PFN; q++) {
        const int u = tid + q * kThreads;
        pf[q] = src[u < PER ? u : 0];
    }
}
template <int PER, int PFN>
__device__ __forceinline__ void wx_put(const f4v (&pf)[PFN], f4v* dst, int tid) {
#pragma unroll
    for (int q = 0; q < PFN; q++) {
        const int u = tid + q * kThreads;
        if (q < PFN - 1 || u < PER) dst[u] = pf[q];
    }
}

// Partial rows part32[chunk][s][row] (complex64); blockIdx.y is the RHS.
// Software-pipelined over the workgroup's columns j: the HLE values of
// column j + 1 (VALU) and the prefetch of its Wx block (global -> registers)
// are issued between the MFMAs of column j; the Wx block goes to the other
// LDS buffer before the one barrier per column.  KS accumulator chains per
// tile (K steps t % KS).
template <int L, int KS>
__global__ __launch_bounds__(kThreads) void k_matfree_mf32(MfJob J, const SysDev* __restrict__ sys, int nchunks,
                                                           int chunk, const f4v* __restrict__ Wx,
                                                           float2* __restrict__ part32) {
    constexpr int R = L * (L + 2);
    constexpr int NC = 2 * R;              // complex output columns per particle
    constexpr int NT = (NC + 15) / 16;     // 32-column MFMA tiles
    constexpr int NH = (2 * L + 1) * (2 * L + 1);
    constexpr int NQ = (NH + 3) / 4;
    constexpr int PER = NT * NQ * 64;      // f4v per column particle
    constexpr int PFN = (PER + kThreads - 1) / kThreads;
    __shared__ f4v Ws[2][PER];
    if (J.skip >= 0 && job_done(sys[J.skip])) return;
    const int N = J.N;
    const int s = blockIdx.y;
    const int groups = (N + 127) / 128;
    const int v = xcd_swizzle(blockIdx.x, gridDim.x);
    const int c = v / groups;
    const int group = v - c * groups;
    const int jbeg = c * chunk, jend = min(N, jbeg + chunk);
    if (jbeg >= jend) return;  // uniform over the workgroup
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int kh = lane >> 5, col = lane & 31, part = col & 1;
    const int i = group * 128 + w * 32 + col;
    const bool live = i < N;
    const int ii = live ? i : N - 1;
    const double xi = J.pos[3 * ii], yi = J.pos[3 * ii + 1], zi = J.pos[3 * ii + 2];
    const f4v* Wall = Wx + (size_t)s * N * PER;
    f16v acc[NT][KS];
#pragma unroll
    for (int t = 0; t < NT; t++)
#pragma unroll
        for (int k = 0; k < KS; k++)
#pragma unroll
            for (int q = 0; q < 16; q++) acc[t][k][q] = 0.f;
    f4v pf[PFN];
    wx_fetch<PER, PFN>(pf, Wall + (size_t)jbeg * PER, tid);
    wx_put<PER, PFN>(pf, Ws[0], tid);
    float af[NQ * 4];
    hle_lane<L, NQ * 4>(xi, yi, zi, J.pos, jbeg, J.k, kh, live && jbeg != ii, af);
    __syncthreads();
    for (int j = jbeg; j < jend; j++) {
        const int buf = (j - jbeg) & 1;
        const int jn = min(j + 1, jend - 1);
        wx_fetch<PER, PFN>(pf, Wall + (size_t)jn * PER, tid);
        float afn[NQ * 4];
        hle_lane<L, NQ * 4>(xi, yi, zi, J.pos, jn, J.k, kh, live && jn != ii, afn);
#pragma unroll
        for (int q = 0; q < NQ; q++)
#pragma unroll
            for (int tile = 0; tile < NT; tile++) {
                const f4v b = Ws[buf][(tile * NQ + q) * 64 + lane];
#pragma unroll
                for (int e4 = 0; e4 < 4; e4++) {
                    const int t = 4 * q + e4;
                    if (t < NH)
                        acc[tile][t % KS] =
                            __builtin_amdgcn_mfma_f32_32x32x2f32(af[t], b[e4], acc[tile][t % KS], 0, 0, 0);
                }
            }
        // scheduling hint: one MFMA, then up to five VALU instructions of the next column's HLE values
#pragma unroll
        for (int t = 0; t < NH * NT; t++) {
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
            __builtin_amdgcn_sched_group_barrier(0x002, 5, 0);
        }
        wx_put<PER, PFN>(pf, Ws[buf ^ 1], tid);
#pragma unroll
        for (int t = 0; t < NQ * 4; t++) af[t] = afn[t];
        __syncthreads();
    }
    const size_t n = (size_t)NC * N;
#pragma unroll
    for (int tile = 0; tile < NT; tile++)
#pragma unroll
        for (int q = 0; q < 16; q++) {
            const int ip = group * 128 + w * 32 + 8 * (q >> 2) + 4 * kh + (q & 3);
            const int cc = tile * 16 + (col >> 1);
            if (ip < N && cc < NC) {
                float v = acc[tile][0][q];
#pragma unroll
                for (int k = 1; k < KS; k++) v += acc[tile][k][q];
                ((float*)part32)[2 * (((size_t)c * J.S + s) * n + (size_t)ip * NC + cc) + part] = v;
            }
        }
}

// ---- complex64 operator on the FP16 matrix cores with hi/lo split operands
//
// The contraction transposed, D[real output column][row particle i] =
// sum_k A[col][k] B[k][i], on v_mfma_f32_32x32x16_f16: B = the pair's HLE
// values (each lane's column is its own particle i, so per-pair factors are
// undone lane-locally), A = W_j from LDS.  K slots: the (p, m >= 0) combos
// u in m-major order; lane half 0 supplies HLE[p, +m], half 1 HLE[p, -m]
// (zero for m = 0), each as (Re, Im) -> 2 K positions; 4 slots per 16-K
// block.  Scaling keeps every FP16 value within [-2^12, 2^12]:
//   B' = HLE c_pm g_ij,  c_pm = sqrt((p-m)!/(p+m)!),  g_ij = 1 / max_p |h_p(k d_ij)|
//   A' = W / (c_pm tau_j),  tau_j = 2^-12 x the power of two at or above max |W_j / c_pm|
// and each pair's block sum is scaled by tau_j / g_ij into the FP32 total.
// (A' reaches up to 2^12 so that its small entries keep a normal lo part;
// k_matfree_mf16s scales B' the same way.)
// Each operand is split v = hi + lo (hi = v rounded to FP16, lo = v - hi
// rounded to FP16); the products hi*hi + hi*lo + lo*hi leave a relative
// error of about 2^-21 per term.

typedef _Float16 h8v __attribute__((ext_vector_type(8)));
typedef _Float16 h2v __attribute__((ext_vector_type(2)));
typedef float f2v __attribute__((ext_vector_type(2)));

template <int L>
struct F16Shape {
    static constexpr int PM = 2 * L;
    static constexpr int NCOMBO = (PM + 1) * (PM + 2) / 2;  // (p, m >= 0) combos
    static constexpr int NKB = (NCOMBO + 3) / 4;             // 16-K blocks
    static constexpr int NC = 2 * L * (L + 2);               // complex output columns
    static constexpr int NT = (2 * NC + 31) / 32;            // 32-row tiles of real output columns
    static constexpr int PER = 2 * NT * NKB * 64;            // 16-byte vectors per column particle (hi, lo)
};

__device__ __forceinline__ void split_hilo(f2v v, h2v& hi, h2v& lo) {
    hi = __builtin_convertvector(v, h2v);  // round to nearest
    // lo = v - hi, the FP16 hi read back by the mixed-precision FMA (one instruction per value)
    const unsigned hb = __builtin_bit_cast(unsigned, hi);
    float l0, l1;
    asm("v_fma_mix_f32 %0, -%1, 1.0, %2 op_sel_hi:[1,0,0]" : "=v"(l0) : "v"(hb), "v"(v.x));
    asm("v_fma_mix_f32 %0, -%1, 1.0, %2 op_sel:[1,0,0] op_sel_hi:[1,0,0]" : "=v"(l1) : "v"(hb), "v"(v.y));
    const f2v lv = {l0, l1};
    lo = __builtin_convertvector(lv, h2v);
}

// combo u (m-major: m = 0, p = 0..PM; m = 1, p = 1..PM; ...) -> (p, m)
__device__ __forceinline__ void combo_pm(int u, int PM, int& p, int& m) {
    m = 0;
    int base = 0;
    while (u >= base + PM + 1 - m) {
        base += PM + 1 - m;
        m++;
    }
    p = m + (u - base);
}

// Per group of WhShape<L>::JB column particles j (blockIdx.x), every source
// s in [s0, s1) in turn: W_j / c_pm into LDS, tau_j = its largest modulus,
// then the A operands (hi and lo) in lane order Wh[s][j][hl][tile][kb][lane]
// (8 halves each) and tau[s][j].  The flattened term lists of every (r, h)
// are staged in LDS once per workgroup and serve all of its sources (packed:
// start | count << 16 per list, 9-bit sources, FP32 coefficients), so each
// entry's terms are LDS reads rather than chains of dependent global loads.
// W is formed in FP32 (its terms are few and the operands keep 22 bits).

// column particles per workgroup, the LDS capacity of the term lists (lmax 3 has 789 terms, lmax 4 2452)
// and the threads per workgroup (lmax 4: 512, so the 64 KB of LDS per workgroup still leaves 4 waves per SIMD;
// lmax 5, 6151 terms: 150 KB of LDS, one workgroup of 1024 threads per CU)
template <int L>
struct WhShape {
    static constexpr int JB = L <= 3 ? 2 : 1;
    static constexpr int TERMS = L <= 3 ? 1024 : (L == 4 ? 2560 : 6400);
    static constexpr int TH = L <= 3 ? 256 : (L == 4 ? 512 : 1024);
};

// The operand kernels' term capacity at lmax (for the launch-time check).
inline int wh_term_capacity(int lmax) {
    return lmax <= 3 ? WhShape<3>::TERMS : (lmax == 4 ? WhShape<4>::TERMS : WhShape<5>::TERMS);
}

// The smallest power of two >= t (t > 0 and normal): a power-of-two tau_j
// makes every rescale by tau exact.
__device__ __forceinline__ float pow2_ceil(float t) {
    unsigned b = __builtin_bit_cast(unsigned, t);
    if (b & 0x7fffffu) b = (b & 0xff800000u) + 0x800000u;
    return __builtin_bit_cast(float, b & 0xff800000u);
}

template <int L>
__device__ __forceinline__ void wh_body(const MfJob& J, int j0, int s0, int s1, int nterms,
                                        const int2* __restrict__ hlist, const int* __restrict__ hsrc,
                                        const double2* __restrict__ hcoef, h8v* __restrict__ Wh,
                                        float* __restrict__ tau) {
    using Sh = F16Shape<L>;
    constexpr int NHm = (2 * L + 1) * (2 * L + 1);
    constexpr int NC = Sh::NC, NT = Sh::NT, NKB = Sh::NKB, JB = WhShape<L>::JB, RM = L * (L + 2);
    constexpr int TERMS = WhShape<L>::TERMS, TH = WhShape<L>::TH, NW = TH / 64;
    static_assert(JB <= 2 && TH % 64 == 0, "the item loop assumes at most two columns and whole waves");
    __shared__ float2 Wn[JB][NHm][NC];
    __shared__ float2 zs[JB][NC];  // Mie coefficient times x of column j, per (polarization, r)
    __shared__ float cinv_s[NHm];  // 1 / c_pm = sqrt((p+m)!/(p-m)!) per HLE index
    __shared__ unsigned hl_s[RM * NHm];
    __shared__ int4 slot_h[2][NKB];  // HLE indices of the 4 K slots of block kb for lane half kh, -1: zero slot
    __shared__ unsigned short hs_s[TERMS];
    __shared__ float2 hc_s[TERMS];
    __shared__ float red[JB][NW];
    const int tid = threadIdx.x;
    const int lmax = J.lmax;
    const size_t n = (size_t)NC * J.N;
    for (int q = tid; q < RM * NHm; q += TH) {
        const int2 se = hlist[q];
        hl_s[q] = (unsigned)se.x | ((unsigned)se.y << 16);
    }
    for (int q = tid; q < nterms; q += TH) {
        const double2 c = hcoef[q];
        hs_s[q] = (unsigned short)hsrc[q];
        hc_s[q] = make_float2((float)c.x, (float)c.y);
    }
    if (tid < NHm) {
        const int p = (int)sqrtf((float)tid + 0.5f), mu = tid - p * (p + 1), m = mu < 0 ? -mu : mu;
        double prod = 1.0;
        for (int q = p - m + 1; q <= p + m; q++) prod *= q;
        cinv_s[tid] = (float)sqrt(prod);
    }
    if (tid < 2 * NKB * 4) {
        const int kh = tid / (NKB * 4), u = tid - kh * (NKB * 4);
        int h = -1;
        if (u < Sh::NCOMBO) {
            int p, m;
            combo_pm(u, Sh::PM, p, m);
            if (!(kh && m == 0)) h = p * (p + 1) + (kh ? -m : m);
        }
        ((int*)&slot_h[kh][0])[u] = h;
    }
    // this thread's output lane is fixed (whole waves per round): lane half, real row parity, row pair
    const int lane = tid & 63, kh = lane >> 5, part = lane & 1, c0 = (lane & 31) >> 1;
    for (int s = s0; s < s1; s++) {
        if (tid < JB * NC) {
            const int b = tid / NC, q = tid - b * NC, j = j0 + b;
            double2 z = make_double2(0, 0);
            if (j < J.N) {
                const float2 xv = ((const float2*)J.x)[s * n + (size_t)j * NC + q];
                const int a = q >= RM ? 1 : 0, r = q - a * RM;
                const int v = (int)sqrtf((float)(r + 1) + 0.5f);
                z = cmul(J.mie[((size_t)j * 2 + a) * lmax + v - 1], make_double2(xv.x, xv.y));
            }
            zs[b][q] = make_float2((float)z.x, (float)z.y);
        }
        __syncthreads();  // zs (and the staged tables at the first source)
        float mx[JB];
#pragma unroll
        for (int b = 0; b < JB; b++) mx[b] = 0.f;
        // entries (h, cc) = idx in steps of TH, the indices advanced without divisions
        int h = tid / NC, cc = tid - (tid / NC) * NC;
        for (int idx = tid; idx < NHm * NC; idx += TH) {
            const int a = cc >= RM ? 1 : 0, r = cc - a * RM;
            const unsigned se = hl_s[r * NHm + h];
            const int k0 = (int)(se & 0xffffu), k1 = k0 + (int)(se >> 16);
            float2 w[JB];
#pragma unroll
            for (int b = 0; b < JB; b++) w[b] = make_float2(0.f, 0.f);
            for (int k = k0; k < k1; k++) {
                const int src = hs_s[k];
                const int zi = ((src >> 8) ? 1 - a : a) * RM + (src & 0xff);  // B terms couple the other polarization
                const float2 cf = hc_s[k];
#pragma unroll
                for (int b = 0; b < JB; b++) cfma(w[b], cf, zs[b][zi]);
            }
            const float cinv = cinv_s[h];
#pragma unroll
            for (int b = 0; b < JB; b++) {
                const float2 wn = make_float2(w[b].x * cinv, w[b].y * cinv);
                Wn[b][h][cc] = wn;
                mx[b] = fmaxf(mx[b], fmaxf(fabsf(wn.x), fabsf(wn.y)));
            }
            cc += TH % NC;
            h += TH / NC;
            if (cc >= NC) {
                cc -= NC;
                h++;
            }
        }
#pragma unroll
        for (int b = 0; b < JB; b++) {
            for (int off = 32; off > 0; off >>= 1) mx[b] = fmaxf(mx[b], __shfl_xor(mx[b], off));
            if (lane == 0) red[b][tid >> 6] = mx[b];
        }
        __syncthreads();
        // per column: a power of two 2^12 times smaller than its largest |W / c_pm| or larger (the A
        // operands reach up to 2^12, high in the FP16 range, so the lo part of small entries stays normal)
        // and its exact reciprocal
        float tinv[JB];
#pragma unroll
        for (int b = 0; b < JB; b++) {
            float t = red[b][0];
#pragma unroll
            for (int q = 1; q < NW; q++) t = fmaxf(t, red[b][q]);
            t = t > 1e-30f ? pow2_ceil(t) * 0x1p-12f : 1.f;
            tinv[b] = __builtin_bit_cast(float, (254u << 23) - __builtin_bit_cast(unsigned, t));
            if (tid == 0 && j0 + b < J.N) tau[(size_t)s * J.N + j0 + b] = t;
        }
        for (int tkb = tid >> 6; tkb < JB * NT * NKB; tkb += NW) {
            const int b = (JB == 2 && tkb >= NT * NKB) ? 1 : 0;
            const int j = j0 + b;
            if (j >= J.N) break;  // items run in b order: the rest are past N too
            const int tk = tkb - b * (NT * NKB);
            const int tile = tk / NKB, kb = tk - tile * NKB;
            const int ccl = 16 * tile + c0;  // this lane's real rows 2 ccl + part
            const float ti = (JB == 2 && b) ? tinv[JB - 1] : tinv[0];
            const int4 hq = slot_h[kh][kb];
            const int hh[4] = {hq.x, hq.y, hq.z, hq.w};
            h8v hv, lv;
#pragma unroll
            for (int e2 = 0; e2 < 4; e2++) {
                const float2 wv = (ccl < NC && hh[e2] >= 0) ? Wn[b][hh[e2]][ccl] : make_float2(0.f, 0.f);
                // A[col = 2cc + part][K = (Re, Im) of the slot]
                f2v av;
                av.x = (part ? wv.y : wv.x) * ti;
                av.y = (part ? wv.x : -wv.y) * ti;
                h2v hi, lo;
                split_hilo(av, hi, lo);
                hv[2 * e2] = hi.x;
                hv[2 * e2 + 1] = hi.y;
                lv[2 * e2] = lo.x;
                lv[2 * e2 + 1] = lo.y;
            }
            h8v* dst = Wh + ((size_t)s * J.N + j) * Sh::PER;
            // streaming stores: the blocks are read back by the contraction, not by this kernel
            __builtin_nontemporal_store(hv, &dst[tk * 64 + lane]);
            __builtin_nontemporal_store(lv, &dst[(NT * NKB + tk) * 64 + lane]);
        }
        __syncthreads();  // Wn, zs and red are rewritten for the next source
    }
}

// Per group of WhShape<L>::JB column particles (blockIdx.x), every source.
template <int L>
__global__ __launch_bounds__(WhShape<L>::TH) void k_matfree_wh(MfJob J, const SysDev* __restrict__ sys, int nterms, const int2* __restrict__ hlist,
                                                         const int* __restrict__ hsrc,
                                                         const double2* __restrict__ hcoef, h8v* __restrict__ Wh,
                                                         float* __restrict__ tau) {
    if (J.skip >= 0 && job_done(sys[J.skip])) return;
    wh_body<L>(J, blockIdx.x * WhShape<L>::JB, 0, J.S, nterms, hlist, hsrc, hcoef, Wh, tau);
}

// This lane's B operands (hi, lo) for pair (i, j): slots of lane half kh,
// and ginv = 1 / g_ij (0 when !on; the geometry is replaced by a unit
// distance there so nothing non-finite arises).  FP64: the separation, k d
// and its reduction modulo pi/2 (Cody-Waite, two-term constant); FP32: sin
// and cos of the reduced argument (Taylor through r^9 / r^10, truncation
// below 2e-9), the Hankel recurrence, and the normalized Legendre functions
//   Pn_m^m = Pn_{m-1}^{m-1} st sqrt((2m-1)/(2m)),  Pn_{m+1}^m = ct sqrt(2m+1) Pn_m^m,
//   Pn_l^m = (2l-1)/sqrt((l-m)(l+m)) ct Pn_{l-1}^m
//            - sqrt((l+m-1)(l-m-1)/((l-m)(l+m))) Pn_{l-2}^m,
// Pn = c_lm P (the standard recurrence of P scaled by c_lm).  g_ij = 1 / max
// of |Re|, |Im| of h_PM, the largest order.
// POW2: the scale is the power of two g = 2^-E with g max|h_PM| in
// [2^11, 2^12) (E returned; ginv unset), so that rescaling by it is exact.
template <int L, bool POW2 = false>
__device__ __forceinline__ void hle_f16(double xi, double yi, double zi, const double* __restrict__ pos, int j, double k,
                                        float kinv, int kh, bool on, h8v (&bh)[F16Shape<L>::NKB],
                                        h8v (&bl)[F16Shape<L>::NKB], float& ginv, int* E = nullptr) {
    using Sh = F16Shape<L>;
    constexpr int PM = Sh::PM, NKB = Sh::NKB;
    double dx = xi - pos[3 * j], dy = yi - pos[3 * j + 1], dz = zi - pos[3 * j + 2];
    dx = on ? dx : 1.0;
    dy = on ? dy : 0.0;
    dz = on ? dz : 0.0;
    const double rad = sqrt(dx * dx + dy * dy + dz * dz);
    const double zz = k * rad;
    const double qd = rint(zz * 0.63661977236758134308);
    double rr = fma(-qd, 1.5707963267948966, zz);
    rr = fma(-qd, 6.123233995736766e-17, rr);
    const float r = (float)rr, r2 = r * r;
    const float S = r + r * r2 * (-1.f / 6.f + r2 * (1.f / 120.f + r2 * (-1.f / 5040.f + r2 * (1.f / 362880.f))));
    const float C = 1.f + r2 * (-0.5f + r2 * (1.f / 24.f + r2 * (-1.f / 720.f + r2 * (1.f / 40320.f -
                                                                                       r2 * (1.f / 3628800.f)))));
    const int quad = (int)qd & 3;
    const float s0 = (quad & 1) ? C : S, c0 = (quad & 1) ? S : C;
    const float sz = (quad & 2) ? -s0 : s0, cz = ((quad + 1) & 2) ? -c0 : c0;
    const float fx = (float)dx, fy = (float)dy, fz = (float)dz;
    const float rinv = __builtin_amdgcn_rcpf((float)rad);
    const float rz = rinv * kinv;
    float hx[PM + 1], hy[PM + 1];
    hx[0] = sz * rz;
    hy[0] = -cz * rz;
    hx[1] = hx[0] * rz + hy[0];
    hy[1] = hy[0] * rz - hx[0];
#pragma unroll
    for (int l = 2; l <= PM; l++) {
        hx[l] = (2 * l - 1) * rz * hx[l - 1] - hx[l - 2];
        hy[l] = (2 * l - 1) * rz * hy[l - 1] - hy[l - 2];
    }
    const float hmax = fmaxf(fabsf(hx[PM]), fabsf(hy[PM]));
    float g;
    if (POW2) {
        // g max|h_PM| in [2^11, 2^12): the block's largest values sit high in the FP16 range, so the
        // lo part of elements down to 2^-15 of the largest is still a normal number
        const int e = (int)((__builtin_bit_cast(unsigned, hmax) >> 23) & 0xff) - 126 - 12;
        *E = e;
        g = on ? __builtin_bit_cast(float, (unsigned)(127 - e) << 23) : 0.f;
    } else {
        g = on ? __builtin_amdgcn_rcpf(hmax) : 0.f;
        ginv = on ? hmax : 0.f;
    }
#pragma unroll
    for (int l = 0; l <= PM; l++) {
        hx[l] *= g;
        hy[l] *= g;
    }
    const float rho2 = fx * fx + fy * fy;
    const float rhoinv = rho2 > 0.f ? __builtin_amdgcn_rsqf(rho2) : 0.f;
    const float ct = fz * rinv;
    const float st = rho2 * rhoinv * rinv;
    const float e1x = rho2 > 0.f ? fx * rhoinv : 1.f;
    const float e1y = (kh ? -fy : fy) * rhoinv;  // e^{-i phi} on lane half 1
    // each (Re, Im) pair is split into hi and lo as soon as it is formed (no array of all values)
#pragma unroll
    for (int kb = 0; kb < NKB; kb++)
#pragma unroll
        for (int e = 0; e < 8; e++) bh[kb][e] = bl[kb][e] = (_Float16)0.f;
    float ex = kh ? 0.f : 1.f, ey = 0.f;  // m = 0 slots belong to lane half 0
    float pmm = 1.f;
    int u = 0;
#pragma unroll
    for (int m = 0; m <= PM; m++) {
        if (m > 0) {
            const float bx = m == 1 ? 1.f : ex, by = m == 1 ? 0.f : ey;
            ex = bx * e1x - by * e1y;
            ey = bx * e1y + by * e1x;
            pmm *= st * (float)__builtin_sqrt((2.0 * m - 1) / (2.0 * m));
        }
        float p2 = 0.f, p1 = 0.f;
#pragma unroll
        for (int l = 0; l <= PM; l++) {
            if (l < m) continue;
            float P;
            if (l == m) {
                P = pmm;
            } else if (l == m + 1) {
                P = ct * (float)__builtin_sqrt(2.0 * m + 1) * pmm;
            } else {
                const float A = (float)((2.0 * l - 1) / __builtin_sqrt((double)(l - m) * (l + m)));
                const float B = (float)__builtin_sqrt((double)(l + m - 1) * (l - m - 1) / ((double)(l - m) * (l + m)));
                P = A * (ct * p1) - B * p2;
            }
            p2 = p1;
            p1 = P;
            const float ar = hx[l] * P, ai = hy[l] * P;
            const f2v v = {ar * ex - ai * ey, ar * ey + ai * ex};
            h2v hi, lo;
            split_hilo(v, hi, lo);
            bh[u >> 2][2 * (u & 3)] = hi.x;
            bh[u >> 2][2 * (u & 3) + 1] = hi.y;
            bl[u >> 2][2 * (u & 3)] = lo.x;
            bl[u >> 2][2 * (u & 3) + 1] = lo.y;
            u++;
        }
    }
}

// The B operands of a column pair (j, j + 1) with the special functions of
// each pair computed once: lane half 0 evaluates its row particle against
// column j, half 1 against column j + 1 (`jown`), each forming both the +m
// slots (half 0's share of a column) and the -m slots (half 1's share, zero
// for m = 0) of its own pair; one v_permlane32_swap per dword then gives the
// lanes of each half their share of both columns: (P, M) -> (column j:
// P of half 0 | M of half 0, column j + 1: P of half 1 | M of half 1).  The
// same for g_ij.  Same values as hle_f16 for each column.
typedef unsigned u4v __attribute__((ext_vector_type(4)));

template <int L>
__device__ __forceinline__ void hle_f16_pair(double xi, double yi, double zi, const double* __restrict__ pos, int jown,
                                             double k, float kinv, bool on, h8v (&bh0)[F16Shape<L>::NKB],
                                             h8v (&bl0)[F16Shape<L>::NKB], h8v (&bh1)[F16Shape<L>::NKB],
                                             h8v (&bl1)[F16Shape<L>::NKB], float& ginv0, float& ginv1) {
    using Sh = F16Shape<L>;
    constexpr int PM = Sh::PM, NKB = Sh::NKB;
    double dx = xi - pos[3 * jown], dy = yi - pos[3 * jown + 1], dz = zi - pos[3 * jown + 2];
    dx = on ? dx : 1.0;
    dy = on ? dy : 0.0;
    dz = on ? dz : 0.0;
    const double rad = sqrt(dx * dx + dy * dy + dz * dz);
    const double zz = k * rad;
    const double qd = rint(zz * 0.63661977236758134308);
    double rr = fma(-qd, 1.5707963267948966, zz);
    rr = fma(-qd, 6.123233995736766e-17, rr);
    const float r = (float)rr, r2 = r * r;
    const float S = r + r * r2 * (-1.f / 6.f + r2 * (1.f / 120.f + r2 * (-1.f / 5040.f + r2 * (1.f / 362880.f))));
    const float C = 1.f + r2 * (-0.5f + r2 * (1.f / 24.f + r2 * (-1.f / 720.f + r2 * (1.f / 40320.f -
                                                                                       r2 * (1.f / 3628800.f)))));
    const int quad = (int)qd & 3;
    const float s0 = (quad & 1) ? C : S, c0 = (quad & 1) ? S : C;
    const float sz = (quad & 2) ? -s0 : s0, cz = ((quad + 1) & 2) ? -c0 : c0;
    const float fx = (float)dx, fy = (float)dy, fz = (float)dz;
    const float rinv = __builtin_amdgcn_rcpf((float)rad);
    const float rz = rinv * kinv;
    float hx[PM + 1], hy[PM + 1];
    hx[0] = sz * rz;
    hy[0] = -cz * rz;
    hx[1] = hx[0] * rz + hy[0];
    hy[1] = hy[0] * rz - hx[0];
#pragma unroll
    for (int l = 2; l <= PM; l++) {
        hx[l] = (2 * l - 1) * rz * hx[l - 1] - hx[l - 2];
        hy[l] = (2 * l - 1) * rz * hy[l - 1] - hy[l - 2];
    }
    const float hmax = fmaxf(fabsf(hx[PM]), fabsf(hy[PM]));
    const float g = on ? __builtin_amdgcn_rcpf(hmax) : 0.f;
    const float ginv = on ? hmax : 0.f;
#pragma unroll
    for (int l = 0; l <= PM; l++) {
        hx[l] *= g;
        hy[l] *= g;
    }
    const float rho2 = fx * fx + fy * fy;
    const float rhoinv = rho2 > 0.f ? __builtin_amdgcn_rsqf(rho2) : 0.f;
    const float ct = fz * rinv;
    const float st = rho2 * rhoinv * rinv;
    const float e1x = rho2 > 0.f ? fx * rhoinv : 1.f;
    const float e1y = fy * rhoinv;
    float vp[NKB * 8], vm[NKB * 8];
#pragma unroll
    for (int t = 0; t < NKB * 8; t++) vp[t] = vm[t] = 0.f;
    float ex = 1.f, ey = 0.f;
    float pmm = 1.f;
    int u = 0;
#pragma unroll
    for (int m = 0; m <= PM; m++) {
        if (m > 0) {
            const float bx = m == 1 ? 1.f : ex, by = m == 1 ? 0.f : ey;
            ex = bx * e1x - by * e1y;
            ey = bx * e1y + by * e1x;
            pmm *= st * (float)__builtin_sqrt((2.0 * m - 1) / (2.0 * m));
        }
        float p2 = 0.f, p1 = 0.f;
#pragma unroll
        for (int l = 0; l <= PM; l++) {
            if (l < m) continue;
            float P;
            if (l == m) {
                P = pmm;
            } else if (l == m + 1) {
                P = ct * (float)__builtin_sqrt(2.0 * m + 1) * pmm;
            } else {
                const float A = (float)((2.0 * l - 1) / __builtin_sqrt((double)(l - m) * (l + m)));
                const float B = (float)__builtin_sqrt((double)(l + m - 1) * (l - m - 1) / ((double)(l - m) * (l + m)));
                P = A * (ct * p1) - B * p2;
            }
            p2 = p1;
            p1 = P;
            const float ar = hx[l] * P, ai = hy[l] * P;
            vp[2 * u] = ar * ex - ai * ey;
            vp[2 * u + 1] = ar * ey + ai * ex;
            if (m > 0) {  // e^{-i m phi} = conj(e^{i m phi})
                vm[2 * u] = ar * ex + ai * ey;
                vm[2 * u + 1] = ai * ex - ar * ey;
            }
            u++;
        }
    }
#pragma unroll
    for (int kb = 0; kb < NKB; kb++) {
#pragma unroll
        for (int e2 = 0; e2 < 4; e2++) {
            const f2v a = {vp[kb * 8 + 2 * e2], vp[kb * 8 + 2 * e2 + 1]};
            const f2v b = {vm[kb * 8 + 2 * e2], vm[kb * 8 + 2 * e2 + 1]};
            h2v ahi, alo, bhi, blo;
            split_hilo(a, ahi, alo);
            split_hilo(b, bhi, blo);
            bh0[kb][2 * e2] = ahi.x;
            bh0[kb][2 * e2 + 1] = ahi.y;
            bl0[kb][2 * e2] = alo.x;
            bl0[kb][2 * e2 + 1] = alo.y;
            bh1[kb][2 * e2] = bhi.x;
            bh1[kb][2 * e2 + 1] = bhi.y;
            bl1[kb][2 * e2] = blo.x;
            bl1[kb][2 * e2 + 1] = blo.y;
        }
        u4v ph = __builtin_bit_cast(u4v, bh0[kb]), mh = __builtin_bit_cast(u4v, bh1[kb]);
        u4v pl = __builtin_bit_cast(u4v, bl0[kb]), ml = __builtin_bit_cast(u4v, bl1[kb]);
#pragma unroll
        for (int d = 0; d < 4; d++) {
            const auto sh = __builtin_amdgcn_permlane32_swap(ph[d], mh[d], false, false);
            ph[d] = sh[0];
            mh[d] = sh[1];
            const auto sl = __builtin_amdgcn_permlane32_swap(pl[d], ml[d], false, false);
            pl[d] = sl[0];
            ml[d] = sl[1];
        }
        bh0[kb] = __builtin_bit_cast(h8v, ph);
        bh1[kb] = __builtin_bit_cast(h8v, mh);
        bl0[kb] = __builtin_bit_cast(h8v, pl);
        bl1[kb] = __builtin_bit_cast(h8v, ml);
    }
    const unsigned gb = __builtin_bit_cast(unsigned, ginv);
    const auto sg = __builtin_amdgcn_permlane32_swap(gb, gb, false, false);
    ginv0 = __builtin_bit_cast(float, (unsigned)sg[0]);
    ginv1 = __builtin_bit_cast(float, (unsigned)sg[1]);
}

// Column pairs: as k_matfree_mf16 with KS chains and no interleaving, the B
// operands of columns j and j + 1 formed together at even steps (the second
// kept in registers for the odd step), so each pair's special functions are
// evaluated once instead of once per lane half.
template <int L, int KS>
__global__ __launch_bounds__(kThreads, 2) void k_matfree_mf16p(MfJob J, const SysDev* __restrict__ sys, int nchunks,
                                                            int chunk, const f4v* __restrict__ Wh,
                                                            const float* __restrict__ tau,
                                                            float2* __restrict__ part32) {
    using Sh = F16Shape<L>;
    constexpr int NC = Sh::NC, NT = Sh::NT, NKB = Sh::NKB, PER = Sh::PER;
    constexpr int PFN = (PER + kThreads - 1) / kThreads;
    __shared__ f4v Ws[2][PER];
    if (J.skip >= 0 && job_done(sys[J.skip])) return;
    const int N = J.N;
    const int s = blockIdx.y;
    const int groups = (N + 127) / 128;
    const int v = xcd_swizzle(blockIdx.x, gridDim.x);
    const int c = v / groups;
    const int group = v - c * groups;
    const int jbeg = c * chunk, jend = min(N, jbeg + chunk);
    if (jbeg >= jend) return;  // uniform over the workgroup
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int kh = lane >> 5, col = lane & 31;
    const int i = group * 128 + w * 32 + col;
    const bool live = i < N;
    const int ii = live ? i : N - 1;
    const double xi = J.pos[3 * ii], yi = J.pos[3 * ii + 1], zi = J.pos[3 * ii + 2];
    const f4v* Wall = Wh + (size_t)s * N * PER;
    const float* tau_s = tau + (size_t)s * N;
    const float kinv = (float)(1.0 / J.k);
    f16v tot[NT];
#pragma unroll
    for (int t = 0; t < NT; t++)
#pragma unroll
        for (int q = 0; q < 16; q++) tot[t][q] = 0.f;
    f4v pf[PFN];
    wx_fetch<PER, PFN>(pf, Wall + (size_t)jbeg * PER, tid);
    wx_put<PER, PFN>(pf, Ws[0], tid);
    __syncthreads();
    h8v bh[NKB], bl[NKB], bh1[NKB], bl1[NKB];
    float ginv = 0.f, ginv1 = 0.f;
    for (int j = jbeg; j < jend; j++) {
        const int buf = (j - jbeg) & 1;
        const int jn = min(j + 1, jend - 1);
        wx_fetch<PER, PFN>(pf, Wall + (size_t)jn * PER, tid);
        if (buf == 0) {
            const int jo = j + kh;  // this half's column; past the chunk at an odd tail
            const bool has = jo < jend;
            const int jc = has ? jo : j;
            hle_f16_pair<L>(xi, yi, zi, J.pos, jc, J.k, kinv, live && has && jc != ii, bh, bl, bh1, bl1, ginv, ginv1);
        } else {
#pragma unroll
            for (int kb = 0; kb < NKB; kb++) {
                bh[kb] = bh1[kb];
                bl[kb] = bl1[kb];
            }
            ginv = ginv1;
        }
        f16v acc[NT][KS];
#pragma unroll
        for (int t = 0; t < NT; t++)
#pragma unroll
            for (int k = 0; k < KS; k++)
#pragma unroll
                for (int q = 0; q < 16; q++) acc[t][k][q] = 0.f;
#pragma unroll
        for (int kb = 0; kb < NKB; kb++)
#pragma unroll
            for (int t = 0; t < NT; t++) {
                const h8v ah = __builtin_bit_cast(h8v, Ws[buf][(t * NKB + kb) * 64 + lane]);
                const h8v al = __builtin_bit_cast(h8v, Ws[buf][((NT + t) * NKB + kb) * 64 + lane]);
                f16v& a = acc[t][kb % KS];
                a = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bh[kb], a, 0, 0, 0);
                a = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bl[kb], a, 0, 0, 0);
                a = __builtin_amdgcn_mfma_f32_32x32x16_f16(al, bh[kb], a, 0, 0, 0);
            }
        const float sc = tau_s[j] * ginv;
#pragma unroll
        for (int t = 0; t < NT; t++)
#pragma unroll
            for (int q = 0; q < 16; q++) {
                float a = acc[t][0][q];
#pragma unroll
                for (int k = 1; k < KS; k++) a += acc[t][k][q];
                tot[t][q] = fmaf(sc, a, tot[t][q]);
            }
        wx_put<PER, PFN>(pf, Ws[buf ^ 1], tid);
        __syncthreads();
    }
    if (live) {
        const size_t n = (size_t)NC * N;
        float* out = (float*)part32 + 2 * (((size_t)c * J.S + s) * n + (size_t)i * NC);
#pragma unroll
        for (int t = 0; t < NT; t++)
#pragma unroll
            for (int q = 0; q < 16; q++) {
                const int Rr = 32 * t + 8 * (q >> 2) + 4 * kh + (q & 3);  // real output column 2 cc + part
                if (Rr < 2 * NC) out[Rr] = tot[t][q];
            }
    }
}

// Partial rows part32[chunk][s][row] (complex64); blockIdx.y is the RHS.  Per
// column j: prefetch of the next Wh block (global -> registers), the pair's
// B operands, 3 NKB NT FP16 MFMAs into KS chains per tile, the block sum
// scaled by tau_j / g_ij into the total, the prefetch to the other LDS
// buffer, one barrier.
template <int L, int KS, int SGB>
__global__ __launch_bounds__(kThreads) void k_matfree_mf16(MfJob J, const SysDev* __restrict__ sys, int nchunks,
                                                           int chunk, const f4v* __restrict__ Wh,
                                                           const float* __restrict__ tau, float2* __restrict__ part32) {
    using Sh = F16Shape<L>;
    constexpr int NC = Sh::NC, NT = Sh::NT, NKB = Sh::NKB, PER = Sh::PER;
    constexpr int PFN = (PER + kThreads - 1) / kThreads;
    __shared__ f4v Ws[2][PER];
    if (J.skip >= 0 && job_done(sys[J.skip])) return;
    const int N = J.N;
    const int s = blockIdx.y;
    const int groups = (N + 127) / 128;
    const int v = xcd_swizzle(blockIdx.x, gridDim.x);
    const int c = v / groups;
    const int group = v - c * groups;
    const int jbeg = c * chunk, jend = min(N, jbeg + chunk);
    if (jbeg >= jend) return;  // uniform over the workgroup
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int kh = lane >> 5, col = lane & 31;
    const int i = group * 128 + w * 32 + col;
    const bool live = i < N;
    const int ii = live ? i : N - 1;
    const double xi = J.pos[3 * ii], yi = J.pos[3 * ii + 1], zi = J.pos[3 * ii + 2];
    const f4v* Wall = Wh + (size_t)s * N * PER;
    const float* tau_s = tau + (size_t)s * N;
    const float kinv = (float)(1.0 / J.k);
    f16v tot[NT];
#pragma unroll
    for (int t = 0; t < NT; t++)
#pragma unroll
        for (int q = 0; q < 16; q++) tot[t][q] = 0.f;
    f4v pf[PFN];
    wx_fetch<PER, PFN>(pf, Wall + (size_t)jbeg * PER, tid);
    wx_put<PER, PFN>(pf, Ws[0], tid);
    h8v bh[NKB], bl[NKB];
    float ginv;
    if (SGB) hle_f16<L>(xi, yi, zi, J.pos, jbeg, J.k, kinv, kh, live && jbeg != ii, bh, bl, ginv);
    __syncthreads();
    for (int j = jbeg; j < jend; j++) {
        const int buf = (j - jbeg) & 1;
        const int jn = min(j + 1, jend - 1);
        wx_fetch<PER, PFN>(pf, Wall + (size_t)jn * PER, tid);
        // SGB > 0: the next column's B operands are computed here and their
        // VALU work is interleaved with this column's MFMAs (SGB VALU
        // instructions per MFMA); SGB = 0: this column's, before its MFMAs
        h8v bhn[NKB], bln[NKB];
        float ginvn;
        if (SGB) hle_f16<L>(xi, yi, zi, J.pos, jn, J.k, kinv, kh, live && jn != ii, bhn, bln, ginvn);
        else hle_f16<L>(xi, yi, zi, J.pos, j, J.k, kinv, kh, live && j != ii, bh, bl, ginv);
        f16v acc[NT][KS];
#pragma unroll
        for (int t = 0; t < NT; t++)
#pragma unroll
            for (int k = 0; k < KS; k++)
#pragma unroll
                for (int q = 0; q < 16; q++) acc[t][k][q] = 0.f;
#pragma unroll
        for (int kb = 0; kb < NKB; kb++)
#pragma unroll
            for (int t = 0; t < NT; t++) {
                const h8v ah = __builtin_bit_cast(h8v, Ws[buf][(t * NKB + kb) * 64 + lane]);
                const h8v al = __builtin_bit_cast(h8v, Ws[buf][((NT + t) * NKB + kb) * 64 + lane]);
                f16v& a = acc[t][kb % KS];
                a = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bh[kb], a, 0, 0, 0);
                a = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bl[kb], a, 0, 0, 0);
                a = __builtin_amdgcn_mfma_f32_32x32x16_f16(al, bh[kb], a, 0, 0, 0);
            }
        if (SGB) {
#pragma unroll
            for (int t = 0; t < 3 * NKB * NT; t++) {
                __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
                __builtin_amdgcn_sched_group_barrier(0x002, SGB, 0);
            }
        }
        const float sc = tau_s[j] * ginv;
#pragma unroll
        for (int t = 0; t < NT; t++)
#pragma unroll
            for (int q = 0; q < 16; q++) {
                float a = acc[t][0][q];
#pragma unroll
                for (int k = 1; k < KS; k++) a += acc[t][k][q];
                tot[t][q] = fmaf(sc, a, tot[t][q]);
            }
        wx_put<PER, PFN>(pf, Ws[buf ^ 1], tid);
        if (SGB) {
#pragma unroll
            for (int kb = 0; kb < NKB; kb++) {
                bh[kb] = bhn[kb];
                bl[kb] = bln[kb];
            }
            ginv = ginvn;
        }
        __syncthreads();
    }
    if (live) {
        const size_t n = (size_t)NC * N;
        float* out = (float*)part32 + 2 * (((size_t)c * J.S + s) * n + (size_t)i * NC);
#pragma unroll
        for (int t = 0; t < NT; t++)
#pragma unroll
            for (int q = 0; q < 16; q++) {
                const int Rr = 32 * t + 8 * (q >> 2) + 4 * kh + (q & 3);  // real output column 2 cc + part
                if (Rr < 2 * NC) out[Rr] = tot[t][q];
            }
    }
}

// ---- complex64 operator: sources sharing each pair's special functions, many jobs per launch
//
// k_matfree_mf16s is the contraction of k_matfree_mf16 with SG right-hand
// sides per workgroup: the B operands (the pair's scaled HLE values, split
// hi/lo) are formed once per pair and feed the 3 NKB NT MFMAs of each of the
// SG sources, whose A operands (W_j of that source, from k_matfree_wh) are
// staged in LDS side by side.  Each source keeps its running total R in the
// scale of the current pair: the pair's factor F = tau_j 2^E_ij (tau_j a
// power of two from k_matfree_wh, 2^-E_ij the power-of-two scale of the
// pair's B operands, 2^-E max|h_PM| in [2^11, 2^12)); the pair's block sum is
// accumulated from zero in one accumulator set shared by the sources in
// turn, and R = R 2^(E_prev - E_new) + acc (the scale is exact), so the total
// needs one register set per source instead of two, with the roundings of
// k_matfree_mf16; the result is R 2^E at the end.
//
// Jobs (one cluster operator product each) travel by value in the kernel
// arguments, up to kMfBatchMax per launch; each job keeps its own
// decomposition (chunks of columns x 128-row groups x source groups), the
// same whatever else shares the launch, so a system's arithmetic does not
// depend on its batch.  Work ids are dealt to the jobs in order and
// XCD-swizzled as in the one-job kernels.
constexpr int kMfBatchMax = 16;

struct Mf16Job {
    MfJob J;
    const f4v* Wh;      // [S][N][PER] operands of k_matfree_wh
    const float* tau;   // [S][N] power-of-two column scales
    float2* part32;     // [nch][S][n] chunk partial rows
    int nch, chunk;     // column chunks and columns per chunk
    int groups, nsg;    // 128-row groups, source groups
};

struct Mf16Batch {
    int count;
    int first[kMfBatchMax + 1];  // prefix of each job's workgroups (contraction, operands or sums)
    Mf16Job job[kMfBatchMax];
};

__device__ __forceinline__ int batch_job(const Mf16Batch& B, int v) {
    int q = 0;
    while (q + 1 < B.count && v >= B.first[q + 1]) q++;
    return q;
}

// Operands of every job: work id = column group of the job, every source in turn.
template <int L>
__global__ __launch_bounds__(WhShape<L>::TH) void k_matfree_wh_b(Mf16Batch B, const SysDev* __restrict__ sys, int nterms, const int2* __restrict__ hlist,
                                                           const int* __restrict__ hsrc,
                                                           const double2* __restrict__ hcoef) {
    const int v = blockIdx.x;
    const int q = batch_job(B, v);
    const Mf16Job& Q = B.job[q];
    if (Q.J.skip >= 0 && job_done(sys[Q.J.skip])) return;
    constexpr int JB = WhShape<L>::JB;
    const int jg = v - B.first[q];
    wh_body<L>(Q.J, jg * JB, 0, Q.J.S, nterms, hlist, hsrc, hcoef, (h8v*)Q.Wh, (float*)Q.tau);
}

typedef __attribute__((address_space(3))) void* lds_ptr_t;

template <int L, int SG>
__global__ __launch_bounds__(kThreads, 2) void k_matfree_mf16s(Mf16Batch B, const SysDev* __restrict__ sys) {
    using Sh = F16Shape<L>;
    constexpr int NC = Sh::NC, NT = Sh::NT, NKB = Sh::NKB, PER = Sh::PER;
    static_assert(PER % kThreads == 0, "operand blocks are whole workgroup rounds");
    // one operand buffer (lmax 3, SG 2: 57 KB), so two workgroups fit a CU and one's special functions
    // overlap the other's MFMAs
    __shared__ f4v Ws[SG * PER];
    const int v0 = xcd_swizzle(blockIdx.x, gridDim.x);
    const int q = batch_job(B, v0);
    const Mf16Job& Q = B.job[q];
    const MfJob& J = Q.J;
    if (J.skip >= 0 && job_done(sys[J.skip])) return;
    int v = v0 - B.first[q];
    const int group = v % Q.groups;
    v /= Q.groups;
    const int c = v % Q.nch;
    const int s0 = (v / Q.nch) * SG;
    const int N = J.N, S = J.S;
    const int nsg = min(SG, S - s0);
    const int jbeg = c * Q.chunk, jend = min(N, jbeg + Q.chunk);
    if (jbeg >= jend) return;  // uniform over the workgroup
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int kh = lane >> 5, col = lane & 31;
    const int i = group * 128 + w * 32 + col;
    const bool live = i < N;
    // a wave whose 32 rows are all past N only helps load the operands: its SIMD goes to the other
    // workgroup of the CU (the 128-row groups pad lmax-4 clusters of 125-377 spheres by 19 %)
    const bool wlive = group * 128 + w * 32 < N;
    const int ii = live ? i : N - 1;
    const double xi = J.pos[3 * ii], yi = J.pos[3 * ii + 1], zi = J.pos[3 * ii + 2];
    const float kinv = (float)(1.0 / J.k);
    const f4v* __restrict__ Wh = Q.Wh;
    const float* __restrict__ tau = Q.tau;
    f16v R[SG][NT];
    int cur[SG];
#pragma unroll
    for (int g = 0; g < SG; g++) {
        cur[g] = 0;
#pragma unroll
        for (int t = 0; t < NT; t++)
#pragma unroll
            for (int e = 0; e < 16; e++) R[g][t][e] = 0.f;
    }
    // the SG sources' operand blocks of column j, side by side (entry u = g PER + e), loaded straight
    // into LDS: each wave fills 64 consecutive entries per round
    auto fetch = [&](f4v* dst, int j) {
#pragma unroll
        for (int g = 0; g < SG; g++) {
            if (g >= nsg) break;  // uniform
            const f4v* src = Wh + ((size_t)(s0 + g) * N + j) * PER + w * 64 + lane;  // wave-uniform base + lane
#pragma unroll
            for (int p = 0; p < PER / kThreads; p++)
                __builtin_amdgcn_global_load_lds((const void*)(src + p * kThreads),
                                                 (lds_ptr_t)(dst + g * PER + p * kThreads + w * 64), 16, 0, 0);
        }
    };
    h8v bh[NKB], bl[NKB];
    float unused;
    // Per column: the pair's B operands (VALU, while the DMA of this column's operands, issued at the
    // end of the previous column, lands), a wait and a barrier (every wave's DMA is in LDS), the rescale
    // and the MFMAs, a barrier (every wave has read the buffer), then the DMA of the next column.
    fetch(Ws, jbeg);
    // lmax 4 skips the idle waves' work inside the loop below (measured faster there); lmax <= 3 in a loop
    // of their own (inside, the condition costs the SG = 2 kernel registers and it spills)
    if (L <= 3 && !wlive) {  // wave-uniform: the same DMA share and barriers per column as the working waves
        for (int j = jbeg; j < jend; j++) {
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
            __syncthreads();
            __syncthreads();
            if (j + 1 < jend) fetch(Ws, j + 1);
        }
        return;
    }
    for (int j = jbeg; j < jend; j++) {
        const bool on = live && j != ii;
        int E = 0;
        if (wlive) hle_f16<L, true>(xi, yi, zi, J.pos, j, J.k, kinv, kh, on, bh, bl, unused, &E);
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __syncthreads();
        // per source: the pair's block sum in an accumulator that starts at zero (its roundings stay at
        // the pair's own magnitude), then R = R 2^(E_prev - E_new) + acc in one FMA (the scale is exact)
#pragma unroll
        for (int g = 0; g < SG; g++) {
            if (g >= nsg || !wlive) continue;  // uniform: the last source group of a job may be short
            const float tj = tau[(size_t)(s0 + g) * N + j];
            const int nw = (int)((__builtin_bit_cast(unsigned, tj) >> 23) & 0xff) - 127 + E;
            const float scale = __builtin_amdgcn_ldexpf(1.f, on ? cur[g] - nw : 0);
            cur[g] = on ? nw : cur[g];
            f16v acc[NT];
#pragma unroll
            for (int t = 0; t < NT; t++)
#pragma unroll
                for (int e = 0; e < 16; e++) acc[t][e] = 0.f;
#pragma unroll
            for (int kb = 0; kb < NKB; kb++)
#pragma unroll
                for (int t = 0; t < NT; t++) {
                    const h8v ah = __builtin_bit_cast(h8v, Ws[g * PER + (t * NKB + kb) * 64 + lane]);
                    const h8v al = __builtin_bit_cast(h8v, Ws[g * PER + ((NT + t) * NKB + kb) * 64 + lane]);
                    f16v& a = acc[t];
                    a = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bh[kb], a, 0, 0, 0);
                    a = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bl[kb], a, 0, 0, 0);
                    a = __builtin_amdgcn_mfma_f32_32x32x16_f16(al, bh[kb], a, 0, 0, 0);
                }
#pragma unroll
            for (int t = 0; t < NT; t++)
#pragma unroll
                for (int e = 0; e < 16; e++) R[g][t][e] = fmaf(R[g][t][e], scale, acc[t][e]);
        }
        __syncthreads();
        if (j + 1 < jend) fetch(Ws, j + 1);
    }
    if (!live) return;
    const size_t n = (size_t)NC * N;
    const bool direct = Q.nch == 1;  // one chunk: the final rows (identity term included) straight into y
#pragma unroll
    for (int g = 0; g < SG; g++) {
        if (g >= nsg) continue;
        const size_t row = (size_t)(s0 + g) * n + (size_t)i * NC;
        float* out = direct ? (float*)J.y + 2 * row : (float*)Q.part32 + 2 * ((size_t)c * S * n + row);
        const float* xin = (const float*)J.x + 2 * row;
#pragma unroll
        for (int t = 0; t < NT; t++)
#pragma unroll
            for (int e = 0; e < 16; e++) {
                const int Rr = 32 * t + 8 * (e >> 2) + 4 * kh + (e & 3);  // real output column 2 cc + part
                if (Rr < 2 * NC) {
                    float val = __builtin_amdgcn_ldexpf(R[g][t][e], cur[g]);
                    if (direct && J.add_identity) val += xin[Rr];
                    out[Rr] = val;
                }
            }
    }
}

// hle_f16<L, true> in two parts, so that a kernel can form the B operands of
// one K slice at a time (the same values): hle_geom evaluates the pair's
// geometry and scaled Hankel values; hle_part<L, U0, U1> forms the combos
// u in [U0, U1) into bh/bl[u / 4 - U0 / 4] (U0 a multiple of 4), resuming the
// Legendre and azimuth recurrences from `st`, which the slices carry in order.
struct HleGeom {
    float ct, st, e1x, e1y;
};
struct HleRec {
    float ex, ey, pmm, p1, p2;
};

template <int L>
__device__ __forceinline__ void hle_geom(double xi, double yi, double zi, const double* __restrict__ pos, int j,
                                         double k, float kinv, int kh, bool on, float (&hx)[2 * L + 1],
                                         float (&hy)[2 * L + 1], HleGeom& G, HleRec& Rc, int& E) {
    constexpr int PM = 2 * L;
    double dx = xi - pos[3 * j], dy = yi - pos[3 * j + 1], dz = zi - pos[3 * j + 2];
    dx = on ? dx : 1.0;
    dy = on ? dy : 0.0;
    dz = on ? dz : 0.0;
    const double rad = sqrt(dx * dx + dy * dy + dz * dz);
    const double zz = k * rad;
    const double qd = rint(zz * 0.63661977236758134308);
    double rr = fma(-qd, 1.5707963267948966, zz);
    rr = fma(-qd, 6.123233995736766e-17, rr);
    const float r = (float)rr, r2 = r * r;
    const float S = r + r * r2 * (-1.f / 6.f + r2 * (1.f / 120.f + r2 * (-1.f / 5040.f + r2 * (1.f / 362880.f))));
    const float C = 1.f + r2 * (-0.5f + r2 * (1.f / 24.f + r2 * (-1.f / 720.f + r2 * (1.f / 40320.f -
                                                                                       r2 * (1.f / 3628800.f)))));
    const int quad = (int)qd & 3;
    const float s0 = (quad & 1) ? C : S, c0 = (quad & 1) ? S : C;
    const float sz = (quad & 2) ? -s0 : s0, cz = ((quad + 1) & 2) ? -c0 : c0;
    const float fx = (float)dx, fy = (float)dy, fz = (float)dz;
    const float rinv = __builtin_amdgcn_rcpf((float)rad);
    const float rz = rinv * kinv;
    hx[0] = sz * rz;
    hy[0] = -cz * rz;
    hx[1] = hx[0] * rz + hy[0];
    hy[1] = hy[0] * rz - hx[0];
#pragma unroll
    for (int l = 2; l <= PM; l++) {
        hx[l] = (2 * l - 1) * rz * hx[l - 1] - hx[l - 2];
        hy[l] = (2 * l - 1) * rz * hy[l - 1] - hy[l - 2];
    }
    // g max|h_PM| in [2^11, 2^12), as hle_f16<L, true>
    const float hmax = fmaxf(fabsf(hx[PM]), fabsf(hy[PM]));
    E = (int)((__builtin_bit_cast(unsigned, hmax) >> 23) & 0xff) - 126 - 12;
    const float g = on ? __builtin_bit_cast(float, (unsigned)(127 - E) << 23) : 0.f;
#pragma unroll
    for (int l = 0; l <= PM; l++) {
        hx[l] *= g;
        hy[l] *= g;
    }
    const float rho2 = fx * fx + fy * fy;
    const float rhoinv = rho2 > 0.f ? __builtin_amdgcn_rsqf(rho2) : 0.f;
    G.ct = fz * rinv;
    G.st = rho2 * rhoinv * rinv;
    G.e1x = rho2 > 0.f ? fx * rhoinv : 1.f;
    G.e1y = (kh ? -fy : fy) * rhoinv;  // e^{-i phi} on lane half 1
    Rc.ex = kh ? 0.f : 1.f;  // m = 0 slots belong to lane half 0
    Rc.ey = 0.f;
    Rc.pmm = 1.f;
    Rc.p1 = Rc.p2 = 0.f;
}

template <int L, int U0, int U1>
__device__ __forceinline__ void hle_part(const float (&hx)[2 * L + 1], const float (&hy)[2 * L + 1],
                                         const HleGeom& G, HleRec& Rc, h8v* __restrict__ bh,
                                         h8v* __restrict__ bl) {
    constexpr int PM = 2 * L;
    static_assert(U0 % 4 == 0, "slices start at a K block");
#pragma unroll
    for (int kb = 0; kb < (U1 - U0 + 3) / 4; kb++)
#pragma unroll
        for (int e = 0; e < 8; e++) bh[kb][e] = bl[kb][e] = (_Float16)0.f;
    int u = 0;
#pragma unroll
    for (int m = 0; m <= PM; m++) {
        if (m > 0 && u >= U0 && u < U1) {  // the group of order m starts in this slice
            const float bx = m == 1 ? 1.f : Rc.ex, by = m == 1 ? 0.f : Rc.ey;
            Rc.ex = bx * G.e1x - by * G.e1y;
            Rc.ey = bx * G.e1y + by * G.e1x;
            Rc.pmm *= G.st * (float)__builtin_sqrt((2.0 * m - 1) / (2.0 * m));
        }
#pragma unroll
        for (int l = 0; l <= PM; l++) {
            if (l < m) continue;
            if (u >= U0 && u < U1) {
                float P;
                if (l == m) {
                    P = Rc.pmm;
                } else if (l == m + 1) {
                    P = G.ct * (float)__builtin_sqrt(2.0 * m + 1) * Rc.pmm;
                } else {
                    const float A = (float)((2.0 * l - 1) / __builtin_sqrt((double)(l - m) * (l + m)));
                    const float B =
                        (float)__builtin_sqrt((double)(l + m - 1) * (l - m - 1) / ((double)(l - m) * (l + m)));
                    P = A * (G.ct * Rc.p1) - B * Rc.p2;
                }
                Rc.p2 = Rc.p1;
                Rc.p1 = P;
                const float ar = hx[l] * P, ai = hy[l] * P;
                const f2v v = {ar * Rc.ex - ai * Rc.ey, ar * Rc.ey + ai * Rc.ex};
                h2v hi, lo;
                split_hilo(v, hi, lo);
                const int w = u - U0;
                bh[w >> 2][2 * (w & 3)] = hi.x;
                bh[w >> 2][2 * (w & 3) + 1] = hi.y;
                bl[w >> 2][2 * (w & 3)] = lo.x;
                bl[w >> 2][2 * (w & 3) + 1] = lo.y;
            }
            u++;
        }
    }
}

// Compile-time loop over the K slices: f(std::integral_constant<int, P>) for P = 0 .. KP - 1.
template <int P, int KP>
struct SliceLoop {
    template <class F>
    __device__ __forceinline__ static void run(F& f) {
        f(std::integral_constant<int, P>{});
        SliceLoop<P + 1, KP>::run(f);
    }
};
template <int KP>
struct SliceLoop<KP, KP> {
    template <class F>
    __device__ __forceinline__ static void run(F&) {}
};

// ---- lmax 5: the operand block of a column (174 KB) exceeds the LDS
//
// k_matfree_mf16k2 is the contraction of k_matfree_mf16s (one source per
// workgroup, the same per-pair power-of-two scales and running totals) with
// the column's operand block taken in KP slices of K blocks (lmax 5, KP 6:
// 3, 3, 3, 3, 3 and 2 of the 17) through two 30 KB LDS buffers, each filled
// by DMA straight from global memory, and the B operands formed slice by
// slice (hle_geom, then hle_part resuming the recurrences), so that a wave
// fits in the 256 registers of two waves per SIMD and two workgroups share
// a CU.  Slice q + 1's DMA is issued at
// the start of slice q, ahead of slice q's B operands (VALU); each wave waits
// for its own share before reading slice q's operands, so by the barrier that
// ends slice q the whole of slice q + 1 has landed.  The other workgroup on
// the CU covers these waits.  (The earlier form, one workgroup per CU with
// register-staged 61 KB slices, took 7.4 ms at N = 3000 with two sources
// against 6.0 ms here.)
template <int L, int KP>
__global__ __launch_bounds__(kThreads, 2) void k_matfree_mf16k2(Mf16Batch B, const SysDev* __restrict__ sys) {
    using Sh = F16Shape<L>;
    constexpr int NC = Sh::NC, NT = Sh::NT, NKB = Sh::NKB, PER = Sh::PER;
    constexpr int KBP = (NKB + KP - 1) / KP;  // K blocks per slice
    constexpr int SL = 2 * NT * KBP * 64;     // f4v entries of a slice buffer: [hl][tile][kbi][lane]
    __shared__ f4v Ws[2][SL];
    const int v0 = xcd_swizzle(blockIdx.x, gridDim.x);
    const int q = batch_job(B, v0);
    const Mf16Job& Q = B.job[q];
    const MfJob& J = Q.J;
    if (J.skip >= 0 && job_done(sys[J.skip])) return;
    int v = v0 - B.first[q];
    const int group = v % Q.groups;
    v /= Q.groups;
    const int c = v % Q.nch;
    const int s = v / Q.nch;  // one source per workgroup
    const int N = J.N, S = J.S;
    const int jbeg = c * Q.chunk, jend = min(N, jbeg + Q.chunk);
    if (jbeg >= jend) return;  // uniform over the workgroup
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int kh = lane >> 5, col = lane & 31;
    const int i = group * 128 + w * 32 + col;
    const bool live = i < N;
    const bool wlive = group * 128 + w * 32 < N;  // a wave past N only loads its share of the slices
    const int ii = live ? i : N - 1;
    const double xi = J.pos[3 * ii], yi = J.pos[3 * ii + 1], zi = J.pos[3 * ii + 2];
    const float kinv = (float)(1.0 / J.k);
    const f4v* __restrict__ Wsrc = Q.Wh + (size_t)s * N * PER;
    const float* __restrict__ tau = Q.tau + (size_t)s * N;
    f16v R[NT];
#pragma unroll
    for (int t = 0; t < NT; t++)
#pragma unroll
        for (int e = 0; e < 16; e++) R[t][e] = 0.f;
    int cur = 0;
    // slice p of column j by DMA: runs of 64 entries (hl, tile, kbi) dealt to the 4 waves; buffer entry
    // ((hl NT + t) KBP + kbi) 64 + lane <- the column's block entry ((hl NT + t) NKB + p KBP + kbi) 64 + lane
    auto dma = [&](int buf, int j, int p) {
        const int kb0 = p * KBP, nk = min(KBP, NKB - kb0);
        const f4v* src = Wsrc + (size_t)j * PER;
        for (int r = w; r < 2 * NT * nk; r += 4) {  // wave-uniform
            const int ht = r / nk, kbi = r - ht * nk;
            __builtin_amdgcn_global_load_lds((const void*)(src + ((size_t)ht * NKB + kb0 + kbi) * 64 + lane),
                                             (lds_ptr_t)(&Ws[buf][(ht * KBP + kbi) * 64]), 16, 0, 0);
        }
    };
    dma(0, jbeg, 0);
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    int buf = 0;
    for (int j = jbeg; j < jend; j++) {
        const bool on = live && j != ii;
        int E = 0;
        float hx[2 * L + 1], hy[2 * L + 1];
        HleGeom G;
        HleRec Rc;
        if (wlive) hle_geom<L>(xi, yi, zi, J.pos, j, J.k, kinv, kh, on, hx, hy, G, Rc, E);
        const float tj = tau[j];
        const int nw = (int)((__builtin_bit_cast(unsigned, tj) >> 23) & 0xff) - 127 + E;
        const float scale = __builtin_amdgcn_ldexpf(1.f, on ? cur - nw : 0);
        cur = on ? nw : cur;
        f16v acc[NT];
#pragma unroll
        for (int t = 0; t < NT; t++)
#pragma unroll
            for (int e = 0; e < 16; e++) acc[t][e] = 0.f;
        auto step = [&](auto pc) {
            constexpr int p = decltype(pc)::value;
            constexpr int U0 = 4 * KBP * p;
            constexpr int U1 = (4 * KBP * (p + 1) < Sh::NCOMBO) ? 4 * KBP * (p + 1) : Sh::NCOMBO;
            constexpr int NKS = (U1 - U0 + 3) / 4;  // K blocks of this slice
            if (p + 1 < KP) dma(buf ^ 1, j, p + 1);
            else if (j + 1 < jend) dma(buf ^ 1, j + 1, 0);
            if (wlive) {
                h8v bh[NKS], bl[NKS];
                hle_part<L, U0, U1>(hx, hy, G, Rc, bh, bl);
#pragma unroll
                for (int kbi = 0; kbi < NKS; kbi++)
#pragma unroll
                    for (int t = 0; t < NT; t++) {
                        const h8v ah = __builtin_bit_cast(h8v, Ws[buf][((0 * NT + t) * KBP + kbi) * 64 + lane]);
                        const h8v al = __builtin_bit_cast(h8v, Ws[buf][((1 * NT + t) * KBP + kbi) * 64 + lane]);
                        f16v& a = acc[t];
                        a = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bh[kbi], a, 0, 0, 0);
                        a = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bl[kbi], a, 0, 0, 0);
                        a = __builtin_amdgcn_mfma_f32_32x32x16_f16(al, bh[kbi], a, 0, 0, 0);
                    }
            }
            // this wave's share of the next slice, then the barrier: the next slice is whole, this one read
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
            __syncthreads();
            buf ^= 1;
        };
        SliceLoop<0, KP>::run(step);
#pragma unroll
        for (int t = 0; t < NT; t++)
#pragma unroll
            for (int e = 0; e < 16; e++) R[t][e] = fmaf(R[t][e], scale, acc[t][e]);
    }
    if (!live) return;
    const size_t n = (size_t)NC * N;
    const bool direct = Q.nch == 1;  // one chunk: the final rows (identity term included) straight into y
    const size_t row = (size_t)s * n + (size_t)i * NC;
    float* out = direct ? (float*)J.y + 2 * row : (float*)Q.part32 + 2 * ((size_t)c * S * n + row);
    const float* xin = (const float*)J.x + 2 * row;
#pragma unroll
    for (int t = 0; t < NT; t++)
#pragma unroll
        for (int e = 0; e < 16; e++) {
            const int Rr = 32 * t + 8 * (e >> 2) + 4 * kh + (e & 3);  // real output column 2 cc + part
            if (Rr < 2 * NC) {
                float val = __builtin_amdgcn_ldexpf(R[t][e], cur);
                if (direct && J.add_identity) val += xin[Rr];
                out[Rr] = val;
            }
        }
}

// y = (add_identity ? x : 0) + the chunk partials in chunk order, for every
// job with more than one chunk: work id = (job, block of 64 elements x 4
// waves) as in k_matfree_sum32.
__global__ __launch_bounds__(kThreads) void k_matfree_sum32_b(Mf16Batch B, const SysDev* __restrict__ sys) {
    __shared__ double2 red[kThreads / 64 - 1][64];
    const int q = batch_job(B, blockIdx.x);
    const Mf16Job& Q = B.job[q];
    const MfJob& J = Q.J;
    if (J.skip >= 0 && job_done(sys[J.skip])) return;
    const int nchunks = Q.nch;
    const size_t total = (size_t)J.S * 2 * lmax_to_rmax(J.lmax) * J.N;
    const float2* X = (const float2*)J.x;
    float2* Y = (float2*)J.y;
    const float2* __restrict__ part32 = Q.part32;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int per = (nchunks + kThreads / 64 - 1) / (kThreads / 64);
    const int c0 = w * per, c1 = min(nchunks, c0 + per);
    const size_t idx = (size_t)(blockIdx.x - B.first[q]) * 64 + lane;
    double ax = 0, ay = 0;
    if (idx < total) {
        int cc = c0;
        for (; cc + 8 <= c1; cc += 8) {
            float2 vv[8];
#pragma unroll
            for (int p = 0; p < 8; p++) vv[p] = part32[(size_t)(cc + p) * total + idx];
#pragma unroll
            for (int p = 0; p < 8; p++) {
                ax += vv[p].x;
                ay += vv[p].y;
            }
        }
        for (; cc < c1; cc++) {
            const float2 vv = part32[(size_t)cc * total + idx];
            ax += vv.x;
            ay += vv.y;
        }
    }
    if (w > 0) red[w - 1][lane] = make_double2(ax, ay);
    __syncthreads();
    if (w == 0 && idx < total) {
#pragma unroll
        for (int p = 0; p < kThreads / 64 - 1; p++) {
            ax += red[p][lane].x;
            ay += red[p][lane].y;
        }
        if (J.add_identity) {
            ax += X[idx].x;
            ay += X[idx].y;
        }
        Y[idx] = make_float2((float)ax, (float)ay);
    }
}

// FP64 variant on the FP64 matrix cores (v_mfma_f64_16x16x4f64): 16 row
// particles per wave; the 98 real K positions of a pair's HLE values are
// spread over the four 16-lane groups g (position 4t + g: HLE index
// h = 2t + g/2, Re for even g, Im for odd g), so each lane keeps 25 of them;
// 4 output tiles of 8 complex columns; FP64 accumulation.
typedef double d4v __attribute__((ext_vector_type(4)));

template <int L, int KS>
__global__ __launch_bounds__(kThreads) void k_matfree_mf64(MfJob J, const SysDev* __restrict__ sys, int nchunks,
                                                           int chunk, const double2* __restrict__ Wm) {
    constexpr int R = L * (L + 2);
    constexpr int NC = 2 * R;               // complex output columns per particle
    constexpr int NT = (NC + 7) / 8;        // 16-column MFMA tiles
    constexpr int PM = 2 * L;
    constexpr int NH = (2 * L + 1) * (2 * L + 1);
    constexpr int KT = (2 * NH + 3) / 4;    // K steps of 4 real positions
    constexpr int HP = 2 * KT;              // HLE rows staged (zero-padded past NH)
    __shared__ double2 Ws[2][HP][NT * 8 + 1];
    if (J.skip >= 0 && job_done(sys[J.skip])) return;
    const int N = J.N;
    const int s = blockIdx.y;
    const int groups = (N + 63) / 64;
    const int v = xcd_swizzle(blockIdx.x, gridDim.x);
    const int c = v / groups;
    const int group = v - c * groups;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int g = lane >> 4, col = lane & 15, part = col & 1;
    const int i = group * 64 + w * 16 + (lane & 15);
    const bool live = i < N;
    const int ii = live ? i : N - 1;
    const double xi = J.pos[3 * ii], yi = J.pos[3 * ii + 1], zi = J.pos[3 * ii + 2];
    const double2* Wall = Wm + (size_t)s * N * NH * NC;
    const int jbeg = c * chunk, jend = min(N, jbeg + chunk);
    d4v acc[NT][KS];  // KS partial sums per tile (K steps t % KS): independent MFMA chains
#pragma unroll
    for (int t = 0; t < NT; t++)
#pragma unroll
        for (int k = 0; k < KS; k++)
#pragma unroll
            for (int q = 0; q < 4; q++) acc[t][k][q] = 0.0;
    auto stage = [&](int j, int buf) {
        const double2* src = Wall + (size_t)j * NH * NC;
        for (int u = tid; u < HP * NT * 8; u += kThreads) {
            const int h = u / (NT * 8), cc = u - h * (NT * 8);
            Ws[buf][h][cc] = (h < NH && cc < NC) ? src[h * NC + cc] : make_double2(0.0, 0.0);
        }
    };
    if (jbeg < jend) stage(jbeg, 0);
    __syncthreads();
    const int hpar = g >> 1, im = g & 1;  // this lane keeps HLE values with h % 2 == hpar, part im
    for (int j = jbeg; j < jend; j++) {
        const int buf = (j - jbeg) & 1;
        if (j + 1 < jend) stage(j + 1, buf ^ 1);
        double af[KT];
#pragma unroll
        for (int t = 0; t < KT; t++) af[t] = 0.0;
        {
            const double dx = xi - J.pos[3 * j], dy = yi - J.pos[3 * j + 1], dz = zi - J.pos[3 * j + 2];
            const double rad = sqrt(dx * dx + dy * dy + dz * dz);
            const double ct = dz / rad, st = sqrt(fmax(0.0, (1.0 - ct) * (1.0 + ct)));
            const double zz = J.k * rad, rz = 1.0 / zz;
            double sz, cz;
            sincos(zz, &sz, &cz);
            const double rho = sqrt(dx * dx + dy * dy);
            const double2 e1 = rho > 0 ? make_double2(dx / rho, dy / rho) : make_double2(1.0, 0.0);
            double2 hh[PM + 1];
            hh[0] = make_double2(sz * rz, -cz * rz);
            hh[1] = make_double2(hh[0].x * rz + hh[0].y, hh[0].y * rz - hh[0].x);
#pragma unroll
            for (int l = 2; l <= PM; l++)
                hh[l] = make_double2((2 * l - 1) * rz * hh[l - 1].x - hh[l - 2].x,
                                     (2 * l - 1) * rz * hh[l - 1].y - hh[l - 2].y);
            double2 e = make_double2(1.0, 0.0);
            const bool on = live && j != ii;
#pragma unroll
            for (int m = 0; m <= PM; m++) {
                if (m > 0) e = cmul(e, e1);
                double pmm = 1.0;
#pragma unroll
                for (int q = 1; q <= PM; q++)
                    if (q <= m) pmm *= (2 * q - 1) * st;
                double p2 = 0.0, p1 = 0.0;
#pragma unroll
                for (int l = 0; l <= PM; l++) {
                    if (l < m) continue;
                    const double P = (l == m)       ? pmm
                                     : (l == m + 1) ? ct * (2 * m + 1) * pmm
                                                    : ((2 * l - 1) * ct * p1 - (l + m - 1) * p2) * (1.0 / (l - m));  // l - m constant: a multiply
                    p2 = p1;
                    p1 = P;
                    const double2 hp = cscale(hh[l], P);
                    const double2 vp = cmul(hp, e);
                    const int hpos = l * (l + 1) + m;
                    if ((hpos & 1) == hpar) af[hpos >> 1] = on ? (im ? vp.y : vp.x) : 0.0;
                    if (m > 0) {
                        const double2 vm = cmul(hp, cconj(e));
                        const int hneg = l * (l + 1) - m;
                        if ((hneg & 1) == hpar) af[hneg >> 1] = on ? (im ? vm.y : vm.x) : 0.0;
                    }
                }
            }
        }
#pragma unroll
        for (int t = 0; t < KT; t++) {
            const int h = 2 * t + hpar;
#pragma unroll
            for (int tile = 0; tile < NT; tile++) {
                const double2 wv = Ws[buf][h][tile * 8 + (col >> 1)];
                acc[tile][t % KS] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[t], xprime<double>(wv.x, wv.y, im, part),
                                                                         acc[tile][t % KS], 0, 0, 0);
            }
        }
        __syncthreads();
    }
    const size_t n = (size_t)NC * N;
#pragma unroll
    for (int tile = 0; tile < NT; tile++)
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int ip = group * 64 + w * 16 + 4 * q + g;
            const int cc = tile * 8 + (col >> 1);
            if (ip < N && cc < NC) {
                double v = acc[tile][0][q];
#pragma unroll
                for (int k = 1; k < KS; k++) v += acc[tile][k][q];
                ((double*)J.part)[2 * (((size_t)c * J.S + s) * n + (size_t)ip * NC + cc) + part] = v;
            }
        }
}

// Slot-mapped FP64 variant.  The four K positions of MFMA step t hold one
// (l, |m|) combination for all four 16-lane groups g: for m > 0, g = 0 and
// 1 the real and imaginary parts of h_l P_l^m e^{+i m phi} (HLE index
// l(l+1)+m), g = 2 and 3 those of h_l P_l^m e^{-i m phi} (index l(l+1)-m);
// the m = 0 orders go in pairs, g = 0, 1 order l and g = 2, 3 order l+1.
// Every lane therefore evaluates the same combinations -- the A value is
// P (h.x a + h.y b) with per-m lane constants (a, b) -- instead of all of
// its pair's values with a lane select per value; the B value is the real or
// imaginary part of W[h][c] picked by the LDS address, with a lane-constant
// sign (the -Im W of the product's real part).
template <int L>
struct Mf64Slots {
    static constexpr int PM = 2 * L;
    static constexpr int NM0 = (PM + 2) / 2;            // m = 0 pair steps
    static constexpr int KT = NM0 + PM * (PM + 1) / 2;  // = (2 NH + 3) / 4
    int ha[KT], hb[KT];                                 // HLE index of lane groups 0-1 and 2-3
    constexpr Mf64Slots() : ha(), hb() {
        int t = 0;
        for (int l = 0; l <= PM; l += 2) {
            ha[t] = l * (l + 1);
            hb[t] = l + 1 <= PM ? (l + 1) * (l + 2) : l * (l + 1);
            t++;
        }
        for (int m = 1; m <= PM; m++)
            for (int l = m; l <= PM; l++) {
                ha[t] = l * (l + 1) + m;
                hb[t] = l * (l + 1) - m;
                t++;
            }
    }
};

// The pair special functions of the FP64 matrix-core kernels for one lane, in
// two parts: mf64_pair_geometry (the pair's direction and its scaled Hankel
// values; zero off the pairs, on = false) and mf64_pair_steps<L, MA, MB>, the
// A values of the K steps of the azimuthal orders m in [MA, MB] (K order of
// Mf64Slots<L>; MA = 0 includes the m = 0 steps), written to af[0 ..), with
// the azimuth and P_m^m recurrences carried in G between parts.  odd, hi: the
// lane group's Re/Im and order choices, sgn the sign of Im e it sees.
template <int L>
struct Mf64Pair {
    double sz, cz, rz;  // sin, cos of kd and 1 / kd (0 off the pairs): the Hankel values are re-formed per part
    double ct, st;
    double2 e1, e;
    double pmm;
};

// h_l(kd) for l = 0 .. 2L by the upward recurrence
template <int L>
__device__ __forceinline__ void mf64_hankel(const Mf64Pair<L>& G, double2 (&hh)[2 * L + 1]) {
    const double rz = G.rz;
    hh[0] = make_double2(G.sz * rz, -G.cz * rz);
    hh[1] = make_double2(hh[0].x * rz + hh[0].y, hh[0].y * rz - hh[0].x);
#pragma unroll
    for (int l = 2; l <= 2 * L; l++)
        hh[l] = make_double2((2 * l - 1) * rz * hh[l - 1].x - hh[l - 2].x, (2 * l - 1) * rz * hh[l - 1].y - hh[l - 2].y);
}

template <int L>
__device__ __forceinline__ void mf64_pair_geometry(double xi, double yi, double zi, const double* __restrict__ pos,
                                                   int j, double k, bool on, Mf64Pair<L>& G) {
    constexpr int PM = 2 * L;
    double dx = xi - pos[3 * j], dy = yi - pos[3 * j + 1], dz = zi - pos[3 * j + 2];
    dx = on ? dx : 1.0;  // a harmless geometry for the zero entries
    dy = on ? dy : 0.0;
    dz = on ? dz : 0.0;
    const double rad = sqrt(dx * dx + dy * dy + dz * dz);
    G.ct = dz / rad;
    G.st = sqrt(fmax(0.0, (1.0 - G.ct) * (1.0 + G.ct)));
    const double zz = k * rad;
    G.rz = on ? 1.0 / zz : 0.0;  // zero Hankel values off the pairs
    sincos(zz, &G.sz, &G.cz);
    const double rho = sqrt(dx * dx + dy * dy);
    G.e1 = rho > 0 ? make_double2(dx / rho, dy / rho) : make_double2(1.0, 0.0);
    G.e = make_double2(1.0, 0.0);
    G.pmm = 1.0;
}

// K steps before the group of azimuthal order m (m = 0: the pair steps of the m = 0 orders)
template <int L>
constexpr int mf64_steps_before(int m) {
    int t = m > 0 ? Mf64Slots<L>::NM0 : 0;
    for (int q = 1; q < m; q++) t += 2 * L - q + 1;
    return t;
}

template <int L, int MA, int MB>
__device__ __forceinline__ void mf64_pair_steps(Mf64Pair<L>& G, int odd, int hi, double sgn, double* af) {
    constexpr int PM = 2 * L, NM0 = Mf64Slots<L>::NM0;
    double2 hh[PM + 1];
    mf64_hankel<L>(G, hh);
    int t = 0;
    if constexpr (MA == 0) {
        // m = 0: P_l^0 by the recurrence in l; steps of two orders
        double P0[PM + 1];
        P0[0] = 1.0;
        P0[1] = G.ct;
#pragma unroll
        for (int l = 2; l <= PM; l++) P0[l] = ((2 * l - 1) * G.ct * P0[l - 1] - (l - 1) * P0[l - 2]) * (1.0 / l);
        // (a0, b0) = (1, 0) or (0, 1) picks Re or Im exactly; fh = 0 or 1
        // picks the step's first or second order exactly
        const double a0 = odd ? 0.0 : 1.0, b0 = odd ? 1.0 : 0.0, fh = hi ? 1.0 : 0.0;
#pragma unroll
        for (int q = 0; q < NM0; q++) {
            const int l1 = 2 * q, l2 = 2 * q + 1;
            const double v1 = P0[l1] * fma(hh[l1].x, a0, hh[l1].y * b0);
            const double v2 = l2 <= PM ? P0[l2] * fma(hh[l2].x, a0, hh[l2].y * b0) : 0.0;
            af[t++] = fma(v2, fh, v1 * (1.0 - fh));
        }
    }
    // m > 0: one step per (l, m), P_l^m by the recurrence in l
#pragma unroll
    for (int m = MA > 0 ? MA : 1; m <= MB; m++) {
        G.e = cmul(G.e, G.e1);
        G.pmm *= (2 * m - 1) * G.st;
        const double sy = sgn * G.e.y;
        const double a = odd ? -sy : G.e.x, b = odd ? G.e.x : sy;
        double p2 = 0.0, p1 = 0.0;
#pragma unroll
        for (int l = m; l <= PM; l++) {
            const double P = (l == m)       ? G.pmm
                             : (l == m + 1) ? G.ct * (2 * m + 1) * G.pmm
                                            : ((2 * l - 1) * G.ct * p1 - (l + m - 1) * p2) * (1.0 / (l - m));
            p2 = p1;
            p1 = P;
            af[t++] = P * fma(hh[l].x, a, hh[l].y * b);
        }
    }
}

// every K step of the pair in one call
template <int L>
__device__ __forceinline__ void mf64_pair_values(double xi, double yi, double zi, const double* __restrict__ pos,
                                                 int j, double k, bool on, int odd, int hi, double sgn,
                                                 double (&af)[Mf64Slots<L>::KT]) {
    Mf64Pair<L> G;
    mf64_pair_geometry<L>(xi, yi, zi, pos, j, k, on, G);
    mf64_pair_steps<L, 0, 2 * L>(G, odd, hi, sgn, af);
}

template <int L, int KS>
__global__ __launch_bounds__(kThreads) void k_matfree_mf64s(MfJob J, const SysDev* __restrict__ sys, int nchunks,
                                                            int chunk, const double2* __restrict__ Wm) {
    constexpr int R = L * (L + 2);
    constexpr int NC = 2 * R;
    constexpr int NT = (NC + 7) / 8;
    constexpr int PM = 2 * L;
    constexpr int NH = (PM + 1) * (PM + 1);
    using SM = Mf64Slots<L>;
    constexpr int KT = SM::KT, NM0 = SM::NM0;
    static_assert(KT == (2 * NH + 3) / 4, "one K step per combination");
    constexpr SM slots{};
    __shared__ double2 Ws[2][NH][NT * 8 + 1];
    if (J.skip >= 0 && job_done(sys[J.skip])) return;
    const int N = J.N;
    const int s = blockIdx.y;
    const int groups = (N + 63) / 64;
    const int v = xcd_swizzle(blockIdx.x, gridDim.x);
    const int c = v / groups;
    const int group = v - c * groups;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int g = lane >> 4, col = lane & 15, part = col & 1;
    const int odd = g & 1, hi = g >> 1;
    const int comp = odd ^ part;  // B: Re W (0) or Im W (1)
    const unsigned long long bsign = (odd && !part) ? 0x8000000000000000ull : 0ull;
    const double sgn = hi ? 1.0 : -1.0;  // sign of Im e seen by this lane group
    const int i = group * 64 + w * 16 + col;
    const bool live = i < N;
    const int ii = live ? i : N - 1;
    const double xi = J.pos[3 * ii], yi = J.pos[3 * ii + 1], zi = J.pos[3 * ii + 2];
    const double2* Wall = Wm + (size_t)s * N * NH * NC;
    const int jbeg = c * chunk, jend = min(N, jbeg + chunk);
    d4v acc[NT][KS];
#pragma unroll
    for (int t = 0; t < NT; t++)
#pragma unroll
        for (int k = 0; k < KS; k++)
#pragma unroll
            for (int q = 0; q < 4; q++) acc[t][k][q] = 0.0;
    // W_j staging, software-pipelined: the next column's block is loaded into
    // registers at the top of an iteration and written to the other LDS
    // buffer after the MFMAs, so the load latency hides behind this column's work
    constexpr int PF = (NH * NT * 8 + kThreads - 1) / kThreads;  // elements per thread
    double2 pf[PF];
    auto fetch = [&](int j) {
        const double2* src = Wall + (size_t)j * NH * NC;
#pragma unroll
        for (int e = 0; e < PF; e++) {
            const int u = e * kThreads + tid;
            const int h = u / (NT * 8), cc = u - h * (NT * 8);
            pf[e] = (u < NH * NT * 8 && cc < NC) ? src[h * NC + cc] : make_double2(0.0, 0.0);
        }
    };
    auto put = [&](int buf) {
#pragma unroll
        for (int e = 0; e < PF; e++) {
            const int u = e * kThreads + tid;
            if (u < NH * NT * 8) Ws[buf][u / (NT * 8)][u % (NT * 8)] = pf[e];
        }
    };
    if (jbeg < jend) {
        fetch(jbeg);
        put(0);
    }
    __syncthreads();
    for (int j = jbeg; j < jend; j++) {
        const int buf = (j - jbeg) & 1;
        const bool more = j + 1 < jend;
        if (more) fetch(j + 1);
        double af[KT];
        mf64_pair_values<L>(xi, yi, zi, J.pos, j, J.k, live && j != ii, odd, hi, sgn, af);
        // the B values of K step t + 2 are read from LDS while the MFMAs of step t run (one wave per
        // SIMD: nothing else covers the LDS latency)
        double bq[3][NT];
        auto ldb = [&](int t, double (&b)[NT]) {
            const int h = hi ? slots.hb[t] : slots.ha[t];
#pragma unroll
            for (int tile = 0; tile < NT; tile++) b[tile] = ((const double*)&Ws[buf][h][tile * 8 + (col >> 1)])[comp];
        };
        ldb(0, bq[0]);
        if (KT > 1) ldb(1, bq[1]);
#pragma unroll
        for (int t = 0; t < KT; t++) {
            if (t + 2 < KT) ldb(t + 2, bq[(t + 2) % 3]);
            __builtin_amdgcn_sched_barrier(0);  // keep the reads ahead of this step's MFMAs
#pragma unroll
            for (int tile = 0; tile < NT; tile++) {
                const double bv =
                    __builtin_bit_cast(double, __builtin_bit_cast(unsigned long long, bq[t % 3][tile]) ^ bsign);
                acc[tile][t % KS] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[t], bv, acc[tile][t % KS], 0, 0, 0);
            }
        }
        if (more) put(buf ^ 1);
        __syncthreads();
    }
    const size_t n = (size_t)NC * N;
#pragma unroll
    for (int tile = 0; tile < NT; tile++)
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int ip = group * 64 + w * 16 + 4 * q + g;
            const int cc = tile * 8 + (col >> 1);
            if (ip < N && cc < NC) {
                double v = acc[tile][0][q];
#pragma unroll
                for (int k = 1; k < KS; k++) v += acc[tile][k][q];
                ((double*)J.part)[2 * (((size_t)c * J.S + s) * n + (size_t)ip * NC + cc) + part] = v;
            }
        }
}

// k_matfree_mf64k: the FP64 matrix-core contraction with one column's W block
// in a single LDS buffer, filled by DMA straight from Wm (the same layout)
// while the waves form the next pair's special functions, then one barrier;
// one accumulator chain per tile, the B values read a K step or two ahead.
// lmax 5, where the block (121 x 70 complex, 136 KB) fills the LDS: one
// workgroup per CU.  lmax <= 4: two workgroups per CU, which needs the 256
// registers of two waves per SIMD, so the A values are formed in two parts
// (the orders m <= 2, the rest after their MFMAs): a single wave per SIMD
// issues FP64 MFMAs at about 33-36 TFLOP/s, two at about 46.
template <int L>
__global__ __launch_bounds__(kThreads, L <= 4 ? 2 : 1) void k_matfree_mf64k(MfJob J, const SysDev* __restrict__ sys,
                                                                             int nchunks, int chunk,
                                                                             const double2* __restrict__ Wm) {
    constexpr int R = L * (L + 2);
    constexpr int NC = 2 * R;
    constexpr int NT = (NC + 7) / 8;
    constexpr int PM = 2 * L;
    constexpr int NH = (PM + 1) * (PM + 1);
    using SM = Mf64Slots<L>;
    constexpr int KT = SM::KT;
    static_assert(KT == (2 * NH + 3) / 4, "one K step per combination");
    constexpr SM slots{};
    constexpr int MS = L <= 4 ? 2 : PM;                // orders of the first part
    constexpr int TS = mf64_steps_before<L>(MS + 1);    // its K steps
    constexpr int TB = KT - TS > 0 ? KT - TS : 1;       // the second part's (array size)
    constexpr int PD = L <= 4 ? 2 : 3;                  // K steps of B values in registers
    constexpr int WB = NH * NC;  // double2 entries of a column's block
    __shared__ double2 Ws[WB];
    if (J.skip >= 0 && job_done(sys[J.skip])) return;
    const int N = J.N;
    const int s = blockIdx.y;
    const int groups = (N + 63) / 64;
    const int v = xcd_swizzle(blockIdx.x, gridDim.x);
    const int c = v / groups;
    const int group = v - c * groups;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int g = lane >> 4, col = lane & 15, part = col & 1;
    const int odd = g & 1, hi = g >> 1;
    const int comp = odd ^ part;  // B: Re W (0) or Im W (1)
    const unsigned long long bsign = (odd && !part) ? 0x8000000000000000ull : 0ull;
    const double sgn = hi ? 1.0 : -1.0;  // sign of Im e seen by this lane group
    const int i = group * 64 + w * 16 + col;
    const bool live = i < N;
    const int ii = live ? i : N - 1;
    const double xi = J.pos[3 * ii], yi = J.pos[3 * ii + 1], zi = J.pos[3 * ii + 2];
    const double2* Wall = Wm + (size_t)s * N * WB;
    const int jbeg = c * chunk, jend = min(N, jbeg + chunk);  // an empty chunk still writes its (zero) rows
    d4v acc[NT];
#pragma unroll
    for (int t = 0; t < NT; t++)
#pragma unroll
        for (int q = 0; q < 4; q++) acc[t][q] = 0.0;
    // column j's block into LDS: 16-byte entries dealt in rounds of 64 per wave
    auto dma = [&](int j) {
        // the lane's entry offset, re-derived per call (so the rounds' addresses are not all kept live)
        int le = w * 64 + lane;
        asm volatile("" : "+v"(le));
        const double2* src = Wall + (size_t)j * WB + le;
#pragma unroll
        for (int r = 0; r < (WB + kThreads - 1) / kThreads; r++) {
            if (r * kThreads + le < WB)
                __builtin_amdgcn_global_load_lds((const void*)(src + r * kThreads), (lds_ptr_t)(Ws + r * kThreads + w * 64),
                                                 16, 0, 0);
        }
    };
    Mf64Pair<L> G;
    double afa[TS], afb[TB];
    // the pair's geometry and the first part's A values for column j
    auto first_part = [&](int j) {
        mf64_pair_geometry<L>(xi, yi, zi, J.pos, j, J.k, live && j != ii, G);
        mf64_pair_steps<L, 0, MS>(G, odd, hi, sgn, afa);
    };
    if (jbeg < jend) {  // uniform
        dma(jbeg);
        first_part(jbeg);
    }
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    constexpr int LAST = NC - (NT - 1) * 8;  // columns of the last tile inside NC
    for (int j = jbeg; j < jend; j++) {
        // the lane's LDS offsets, re-derived every column (an empty asm hides that they do not change), so
        // the compiler forms each step's addresses in the step instead of keeping all of them in registers
        int lc = col >> 1, hv = hi;
        asm volatile("" : "+v"(lc), "+v"(hv));
        const bool in = lc < LAST;
        const double* pb = (const double*)Ws + comp + 2 * lc;                // tiles 0 .. NT - 2
        const double* pl = (const double*)Ws + comp + 2 * (in ? lc : 0);     // the last tile
        double bq[PD][NT];
        auto ldb = [&](int t, double (&b)[NT]) {
            const int h = hv ? slots.hb[t] : slots.ha[t];
            const double* ph = pb + 2 * NC * h;
#pragma unroll
            for (int tile = 0; tile < NT - 1; tile++) b[tile] = ph[16 * tile];
            // the last tile's columns past NC read zero (and no address past the block)
            const double wv = pl[2 * NC * h + 16 * (NT - 1)];
            b[NT - 1] = in ? wv : 0.0;
        };
        auto step = [&](int t, double a) {
            if (t + PD - 1 < KT) ldb(t + PD - 1, bq[(t + PD - 1) % PD]);
            __builtin_amdgcn_sched_barrier(0);  // keep the reads ahead of this step's MFMAs
#pragma unroll
            for (int tile = 0; tile < NT; tile++) {
                const double bv =
                    __builtin_bit_cast(double, __builtin_bit_cast(unsigned long long, bq[t % PD][tile]) ^ bsign);
                acc[tile] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bv, acc[tile], 0, 0, 0);
            }
        };
#pragma unroll
        for (int t = 0; t < PD - 1; t++) ldb(t, bq[t]);
#pragma unroll
        for (int t = 0; t < TS; t++) step(t, afa[t]);
        if constexpr (TS < KT) {
            mf64_pair_steps<L, MS + 1, PM>(G, odd, hi, sgn, afb);
#pragma unroll
            for (int t = TS; t < KT; t++) step(t, afb[t - TS]);
        }
        __syncthreads();  // every wave has read this column's block
        if (j + 1 < jend) {  // uniform
            dma(j + 1);
            first_part(j + 1);
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
            __syncthreads();
        }
    }
    const size_t n = (size_t)NC * N;
#pragma unroll
    for (int tile = 0; tile < NT; tile++)
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int ip = group * 64 + w * 16 + 4 * q + g;
            const int cc = tile * 8 + (col >> 1);
            if (ip < N && cc < NC)
                ((double*)J.part)[2 * (((size_t)c * J.S + s) * n + (size_t)ip * NC + cc) + part] = acc[tile][q];
        }
}

// Column chunks of the FP64 matrix-core kernel: about 1024 workgroups in total
// over the 64-row groups and the S sources (the sources are the grid's y), so
// several sources mean fewer, longer chunks and fewer partial rows to sum.
// The workspace is sized for S = 1, the most chunks.
inline int mf64_chunks(int N, int S = 1) {
    const int groups = (N + 63) / 64, gs = groups * (S < 1 ? 1 : S);
    int c = (1024 + gs - 1) / gs;
    if (c > N) c = N;
    return c < 1 ? 1 : c;
}

// y32 = (add_identity ? x32 : 0) + sum of the complex64 chunk partials (FP64
// sums).  A workgroup covers 64 consecutive elements: wave w sums its quarter
// of the chunks (coalesced loads, 8 in flight), and the four partial sums are
// added in wave order through LDS (fixed order: deterministic).
__global__ __launch_bounds__(kThreads) void k_matfree_sum32(MfJob J, const SysDev* __restrict__ sys, int nchunks,
                                                            const float2* __restrict__ part32) {
    __shared__ double2 red[kThreads / 64 - 1][64];
    if (J.skip >= 0 && job_done(sys[J.skip])) return;
    const size_t total = (size_t)J.S * 2 * lmax_to_rmax(J.lmax) * J.N;
    const float2* X = (const float2*)J.x;
    float2* Y = (float2*)J.y;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int per = (nchunks + kThreads / 64 - 1) / (kThreads / 64);
    const int c0 = w * per, c1 = min(nchunks, c0 + per);
    for (size_t base = (size_t)blockIdx.x * 64; base < total; base += (size_t)gridDim.x * 64) {
        const size_t idx = base + lane;
        double ax = 0, ay = 0;
        if (idx < total) {
            int cc = c0;
            for (; cc + 8 <= c1; cc += 8) {
                float2 v[8];
#pragma unroll
                for (int q = 0; q < 8; q++) v[q] = part32[(size_t)(cc + q) * total + idx];
#pragma unroll
                for (int q = 0; q < 8; q++) {
                    ax += v[q].x;
                    ay += v[q].y;
                }
            }
            for (; cc < c1; cc++) {
                const float2 v = part32[(size_t)cc * total + idx];
                ax += v.x;
                ay += v.y;
            }
        }
        if (w > 0) red[w - 1][lane] = make_double2(ax, ay);
        __syncthreads();
        if (w == 0 && idx < total) {
#pragma unroll
            for (int q = 0; q < kThreads / 64 - 1; q++) {
                ax += red[q][lane].x;
                ay += red[q][lane].y;
            }
            if (J.add_identity) {
                ax += X[idx].x;
                ay += X[idx].y;
            }
            Y[idx] = make_float2((float)ax, (float)ay);
        }
        __syncthreads();
    }
}

// Column chunks of the complex64 kernel: about 512 workgroups in total (two
// per CU at the kernel's register count, one wave of workgroups).
inline int mf32_chunks(int N) {
    const int groups = (N + 127) / 128;
    int c = (512 + groups - 1) / groups;
    if (c > N) c = N;
    return c < 1 ? 1 : c;
}

// Column chunks of the factorized kernel: about 512 workgroups in total.
inline int fac_chunks(int N) {
    const int groups = (N + kThreads - 1) / kThreads;
    int c = (512 + groups - 1) / groups;
    if (c > N) c = N;
    return c < 1 ? 1 : c;
}

// f4v entries of Wx per column particle: tiles x K quads x 64 lanes
inline int wx_per(int lmax) {
    const int nc = 2 * lmax_to_rmax(lmax), nh = (2 * lmax + 1) * (2 * lmax + 1);
    return ((nc + 15) / 16) * ((nh + 3) / 4) * 64;
}

// 16-byte vectors of Wh per column particle (FP16 kernel)
inline int wh_per(int lmax) {
    return lmax == 1 ? F16Shape<1>::PER
           : lmax == 2 ? F16Shape<2>::PER
           : lmax == 3 ? F16Shape<3>::PER
           : lmax == 4 ? F16Shape<4>::PER
                       : F16Shape<5>::PER;
}

inline int fac_entries_of(int lmax) {
    return lmax == 1 ? kFacEntries<1>.count : lmax == 2 ? kFacEntries<2>.count : kFacEntries<3>.count;
}

// y = (add_identity ? x : 0) + sum of the chunk partials in chunk order
__global__ __launch_bounds__(kThreads) void k_matfree_sum(MfJob J, const SysDev* __restrict__ sys, int nchunks) {
    if (J.skip >= 0 && job_done(sys[J.skip])) return;
    const size_t total = (size_t)J.S * 2 * lmax_to_rmax(J.lmax) * J.N;
    for (size_t idx = (size_t)blockIdx.x * kThreads + threadIdx.x; idx < total; idx += (size_t)gridDim.x * kThreads) {
        double2 acc = J.add_identity ? J.x[idx] : make_double2(0, 0);
        for (int c = 0; c < nchunks; c++) acc = cadd(acc, J.part[(size_t)c * total + idx]);
        J.y[idx] = acc;
    }
}

}  // namespace

// Workspace: partial rows [slots][S][n], then z = Tm x [S][n], then (the
// factorized path) W [S][N][entries][2].
inline int partial_slots(int N) { return std::max(std::max(chunks_for(N), fac_chunks(N)), mf64_chunks(N)); }

size_t matfree_workspace_bytes(int N, int lmax, int S) {
    const size_t n = (size_t)2 * lmax_to_rmax(lmax) * N;
    size_t b = sizeof(double2) * n * S * (partial_slots(N) + 1);
    if (lmax <= kMf64MaxLmax) {  // W of the factorized kernel or Wm of the matrix-core kernel
        const int nh = (2 * lmax + 1) * (2 * lmax + 1);
        size_t w = sizeof(double2) * (size_t)S * nh * n;
        if (fac_possible(N, lmax)) w = std::max(w, sizeof(double2) * 2 * (size_t)S * N * fac_entries_of(lmax));
        b += w;
    }
    if (matfree_c64_supported(N, lmax, false)) {  // the complex64 layout: partials, then Wx
        const size_t ops = std::max(sizeof(float) * 4 * (size_t)S * N * wx_per(lmax),
                                    (size_t)16 * S * N * wh_per(lmax) + sizeof(float) * S * N);
        const size_t b32 = sizeof(float2) * n * S * mf32_chunks(N) + ops;
        b = std::max(b, b32);
    }
    return b;
}

// lmax <= 3: FP16-split or FP32 matrix-core kernel; lmax 4: FP16-split kernel (whatever the option says);
// lmax 5: the FP16-split kernel with the operand block staged in K slices (k_matfree_mf16k2)
constexpr int kC64MaxLmax = 5;
bool matfree_c64_supported(int N, int lmax, bool tmatrix) { return lmax <= kC64MaxLmax && N >= 1 && !tmatrix; }

// Sources per workgroup of k_matfree_mf16s: its LDS holds two buffers of SG
// operand blocks (lmax 3: 2 x 57 KB; lmax 4: 2 x 74 KB at SG = 1).
constexpr int mf16s_sg(int L) { return L <= 2 ? 4 : (L == 3 ? 2 : 1); }

// Column chunks of one job in k_matfree_mf16s: about 2048 workgroups for
// the job alone (128-row groups x source groups x chunks: four rounds of two
// workgroups per CU, so the last, partly filled round costs little), but no
// chunk under 32 columns when that still gives the job about one round (460
// workgroups), else none under 16 (8 at lmax <= 2, whose columns are cheap),
// so a workgroup's set-up and partial rows stay small against its columns;
// never more chunks than the workspace holds partial rows for.  (Measured:
// batch256's batched product 3.26 -> 3.10 ms with 16 against 32 columns for
// its clusters of 125-377 spheres; single systems of 150-400 spheres at
// lmax 4 and 5 1.3-1.9x faster, of 1000 and more 0-5 % slower, hence the
// 32-column floor where the job fills a round on its own.)
inline int mf16s_chunks(int N, int S, int L) {
    const int groups = (N + 127) / 128, nsg = (S + mf16s_sg(L) - 1) / mf16s_sg(L), gs = groups * nsg;
    int c = (2048 + gs - 1) / gs;
    const int minc = gs * std::max(1, N / 32) >= 460 ? 32 : (L <= 2 ? 8 : 16);
    c = std::min(c, std::max(1, N / minc));
    c = std::min(c, mf32_chunks(N));
    return c < 1 ? 1 : c;
}

template <int L>
int launch_mf16s_pack(const MfJob* const* jobs, int count, const SysDev* sys, hipStream_t st, const TransTable* T) {
    constexpr int SG = mf16s_sg(L), JB = WhShape<L>::JB;
    Mf16Batch Bw{}, Bc{}, Bs{};
    Bw.count = Bc.count = count;
    for (int k = 0; k < count; k++) {
        const MfJob& J = *jobs[k];
        const size_t n = (size_t)2 * lmax_to_rmax(L) * J.N, total = (size_t)J.S * n;
        Mf16Job Q{};
        Q.J = J;
        Q.part32 = (float2*)J.part;
        Q.Wh = (const f4v*)(Q.part32 + (size_t)mf32_chunks(J.N) * total);
        Q.tau = (const float*)((const h8v*)Q.Wh + (size_t)J.S * J.N * F16Shape<L>::PER);
        Q.groups = (J.N + 127) / 128;
        Q.nsg = (J.S + SG - 1) / SG;
        Q.chunk = (J.N + mf16s_chunks(J.N, J.S, L) - 1) / mf16s_chunks(J.N, J.S, L);
        Q.nch = (J.N + Q.chunk - 1) / Q.chunk;  // every chunk non-empty: each writes its partial rows
        Bw.job[k] = Bc.job[k] = Q;
        Bw.first[k + 1] = Bw.first[k] + (J.N + JB - 1) / JB;
        Bc.first[k + 1] = Bc.first[k] + Q.groups * Q.nch * Q.nsg;
        if (Q.nch > 1) {
            Bs.job[Bs.count] = Q;
            Bs.first[Bs.count + 1] = Bs.first[Bs.count] + (int)((total + 63) / 64);
            Bs.count++;
        }
    }
    const dim3 block(kThreads);
    hipLaunchKernelGGL((k_matfree_wh_b<L>), dim3((unsigned)Bw.first[count]), dim3(WhShape<L>::TH), 0, st, Bw, sys,
                       T->nhterms, T->hlist, T->hterm_src, T->hterm_coef);
    if constexpr (L == 5)
        hipLaunchKernelGGL((k_matfree_mf16k2<5, 6>), dim3((unsigned)Bc.first[count]), block, 0, st, Bc, sys);
    else
        hipLaunchKernelGGL((k_matfree_mf16s<L, SG>), dim3((unsigned)Bc.first[count]), block, 0, st, Bc, sys);
    if (Bs.count)
        hipLaunchKernelGGL(k_matfree_sum32_b, dim3((unsigned)Bs.first[Bs.count]), block, 0, st, Bs, sys);
    MPA_LAUNCH_CHECK();
    return MPA_OK;
}

// complex64 products of sphere clusters (lmax <= 4) through k_matfree_mf16s:
// jobs grouped by lmax, kMfBatchMax per launch.
int matfree_launch_c64_shared(const MfJob* const* jobs, int count, const SysDev* sys, hipStream_t st) {
    ProfScope prof(5, st);
    for (int L = 1; L <= kC64MaxLmax; L++) {
        const TransTable* T = nullptr;
        std::vector<const MfJob*> sel;
        for (int k = 0; k < count; k++)
            if (jobs[k]->lmax == L) sel.push_back(jobs[k]);
        if (sel.empty()) continue;
        T = mpa_get_table(L);
        if (!T) return MPA_ERR_ARG;
        if (T->nhterms > wh_term_capacity(L)) {
            mpa_set_error("matfree: %d contraction terms exceed the operand kernel's %d", T->nhterms,
                          wh_term_capacity(L));
            return MPA_ERR_ARG;
        }
        for (size_t b = 0; b < sel.size(); b += kMfBatchMax) {
            const int cnt = (int)std::min<size_t>(kMfBatchMax, sel.size() - b);
            long long wg = 0;
            for (int k = 0; k < cnt; k++) {
                const MfJob& J = *sel[b + k];
                wg += (long long)J.S * J.N * 2 * lmax_to_rmax(L);
            }
            if (wg > 0x7fffffffLL) {
                mpa_set_error("matfree: batch too large for one launch");
                return MPA_ERR_ARG;
            }
            int rc;
            switch (L) {
                case 1: rc = launch_mf16s_pack<1>(sel.data() + b, cnt, sys, st, T); break;
                case 2: rc = launch_mf16s_pack<2>(sel.data() + b, cnt, sys, st, T); break;
                case 3: rc = launch_mf16s_pack<3>(sel.data() + b, cnt, sys, st, T); break;
                case 4: rc = launch_mf16s_pack<4>(sel.data() + b, cnt, sys, st, T); break;
                default: rc = launch_mf16s_pack<5>(sel.data() + b, cnt, sys, st, T); break;
            }
            if (rc) return rc;
        }
    }
    return MPA_OK;
}

int matfree_launch_c64(const MfJob& J, const SysDev* sys, hipStream_t st, const TransTable* T) {
    if (!matfree_c64_supported(J.N, J.lmax, J.tmat != nullptr)) {
        mpa_set_error("matfree: the complex64 operator covers spheres with lmax <= %d", kC64MaxLmax);
        return MPA_ERR_ARG;
    }
    if (J.lmax == 5 || (mf_shared_for(J.S, J.lmax) && mpa_option(MPA_OPT_MATFREE_C64_KERNEL) == 0)) {
        const MfJob* one = &J;  // lmax 5 has only the K-sliced kernel of the shared launches
        return matfree_launch_c64_shared(&one, 1, sys, st);
    }
    const int nh = T->nh;
    const size_t n = (size_t)2 * lmax_to_rmax(J.lmax) * J.N;
    const size_t total = (size_t)J.S * n;
    const int chunk = (J.N + mf32_chunks(J.N) - 1) / mf32_chunks(J.N);
    const int nch = (J.N + chunk - 1) / chunk;  // every chunk non-empty: each writes its partial rows
    float2* part32 = (float2*)J.part;
    f4v* Wx = (f4v*)(part32 + (size_t)nch * total);
    const int groups = (J.N + 127) / 128;
    ProfScope prof(5, st);
    const dim3 grid((unsigned)(groups * nch), (unsigned)J.S), block(kThreads);
    const bool ks2 = g_mf_variant[0] == 1;
    if (mpa_option(MPA_OPT_MATFREE_C64_KERNEL) == 0 || J.lmax > kFacMaxLmax) {
        if (T->nhterms > wh_term_capacity(J.lmax)) {
            mpa_set_error("matfree: %d contraction terms exceed the operand kernel's %d", T->nhterms,
                          wh_term_capacity(J.lmax));
            return MPA_ERR_ARG;
        }
        h8v* Whp = (h8v*)Wx;
        float* taup = (float*)(Whp + (size_t)J.S * J.N * wh_per(J.lmax));
        const int jb = J.lmax <= 3 ? WhShape<3>::JB : WhShape<4>::JB;
        const dim3 wgrid((unsigned)((J.N + jb - 1) / jb));
#define MPA_MF16(L_)                                                                                                \
    do {                                                                                                            \
        hipLaunchKernelGGL((k_matfree_wh<L_>), wgrid, dim3(WhShape<L_>::TH), 0, st, J, sys, T->nhterms, T->hlist,  \
                           T->hterm_src, T->hterm_coef, Whp, taup);                                                 \
        switch (g_mf_variant[0]) {                                                                                  \
            case 1: hipLaunchKernelGGL((k_matfree_mf16<L_, 2, 0>), grid, block, 0, st, J, sys, nch, chunk,           \
                                       (const f4v*)Whp, taup, part32); break;                                       \
            case 2: hipLaunchKernelGGL((k_matfree_mf16<L_, 1, 8>), grid, block, 0, st, J, sys, nch, chunk,           \
                                       (const f4v*)Whp, taup, part32); break;                                       \
            case 3: hipLaunchKernelGGL((k_matfree_mf16<L_, 1, 12>), grid, block, 0, st, J, sys, nch, chunk,          \
                                       (const f4v*)Whp, taup, part32); break;                                       \
            case 4:                                                                                         \
            case 5:                                                                                         \
            case 7:                                                                                         \
                if (L_ <= 3 && g_mf_variant[0] != 5)                                                        \
                    hipLaunchKernelGGL((k_matfree_mf16p<(L_ < 3 ? L_ : 3), 1>), grid, block, 0, st, J, sys, nch,     \
                                       chunk, (const f4v*)Whp, taup, part32);                                       \
                else if (L_ <= 3)                                                                           \
                    hipLaunchKernelGGL((k_matfree_mf16p<(L_ < 3 ? L_ : 3), 2>), grid, block, 0, st, J, sys, nch,     \
                                       chunk, (const f4v*)Whp, taup, part32);                                       \
                else /* lmax 4: the pair kernel's operands would not fit the registers */                   \
                    hipLaunchKernelGGL((k_matfree_mf16<L_, 1, 0>), grid, block, 0, st, J, sys, nch, chunk,   \
                                       (const f4v*)Whp, taup, part32);                                              \
                break;                                                                                      \
            default: hipLaunchKernelGGL((k_matfree_mf16<L_, 1, 0>), grid, block, 0, st, J, sys, nch, chunk,          \
                                        (const f4v*)Whp, taup, part32);                                             \
        }                                                                                                           \
    } while (0)
        if (J.lmax == 1) MPA_MF16(1);
        else if (J.lmax == 2) MPA_MF16(2);
        else if (J.lmax == 3) MPA_MF16(3);
        else MPA_MF16(4);
#undef MPA_MF16
        hipLaunchKernelGGL(k_matfree_sum32, dim3((unsigned)std::min<size_t>(4096, (total + 63) / 64)),
                           dim3(kThreads), 0, st, J, sys, nch, part32);
        MPA_LAUNCH_CHECK();
        return MPA_OK;
    }
#define MPA_MF32(L_)                                                                                               \
    do {                                                                                                           \
        constexpr int NT_ = (2 * L_ * (L_ + 2) + 15) / 16, NQ_ = ((2 * L_ + 1) * (2 * L_ + 1) + 3) / 4;            \
        const size_t wtot = (size_t)J.S * J.N * NQ_ * 4 * NT_ * 16;                                                \
        hipLaunchKernelGGL((k_matfree_wx<NT_, NQ_>),                                                               \
                           dim3((unsigned)std::min<size_t>(4096, (wtot + kThreads - 1) / kThreads)), dim3(kThreads), \
                           0, st, J, sys, nh, T->hmap, T->wmeta, T->wsrc, T->wcoef, (float*)Wx);                   \
        if (ks2) hipLaunchKernelGGL((k_matfree_mf32<L_, 2>), grid, block, 0, st, J, sys, nch, chunk, Wx, part32);  \
        else hipLaunchKernelGGL((k_matfree_mf32<L_, 1>), grid, block, 0, st, J, sys, nch, chunk, Wx, part32);      \
    } while (0)
    if (J.lmax == 1) MPA_MF32(1);
    else if (J.lmax == 2) MPA_MF32(2);
    else MPA_MF32(3);
#undef MPA_MF32
    hipLaunchKernelGGL(k_matfree_sum32, dim3((unsigned)std::min<size_t>(4096, (total + 63) / 64)),
                       dim3(kThreads), 0, st, J, sys, nch, part32);
    MPA_LAUNCH_CHECK();
    return MPA_OK;
}

}  // namespace mpa

int mpa_internal_tmatrix_apply(int N, int lmax, int nrhs, const double* tmatrix, const void* x, void* y,
                               hipStream_t st);

namespace mpa {

int matfree_launch(const MfJob& J0, const SysDev* sys, hipStream_t st) {
    const TransTable* T = mpa_get_table(J0.lmax);
    if (!T) return MPA_ERR_ARG;
    if (J0.c64) return matfree_launch_c64(J0, sys, st, T);
    MfJob J = J0;
    const size_t n = (size_t)2 * lmax_to_rmax(J.lmax) * J.N;
    const size_t total = (size_t)J.S * n;
    const unsigned sum_blocks = (unsigned)std::min<size_t>(2048, (total + kThreads - 1) / kThreads);
    double2* zbuf = J.part + (size_t)partial_slots(J.N) * J.S * n;
    J.zpre = nullptr;
    if (J.tmat) {
        const int rc = mpa_internal_tmatrix_apply(J.N, J.lmax, J.S, (const double*)J.tmat, J.x, zbuf, st);
        if (rc) return rc;
        J.zpre = zbuf;
    }
    if (mf64_path(J.lmax)) {
        const int nh = T->nh;
        double2* Wm = zbuf + total;
        const int nch = mf64_chunks(J.N, J.S);
        const int chunk = (J.N + nch - 1) / nch;
        const int groups = (J.N + 63) / 64;
        const size_t wtot = (size_t)J.S * J.N * nh * (n / J.N);
        ProfScope prof(4, st);
        hipLaunchKernelGGL(k_matfree_wm<double>,
                           dim3((unsigned)std::min<size_t>(4096, (wtot + kThreads - 1) / kThreads)), dim3(kThreads),
                           0, st, J, sys, nh, T->hmap, T->wmeta, T->wsrc, T->wcoef, Wm);
        const dim3 grid((unsigned)(groups * nch), (unsigned)J.S), block(kThreads);
        const int var = g_mf_variant[1];
#define MPA_MF64(L_)                                                                                          \
    do {                                                                                                      \
        if (var == 0) hipLaunchKernelGGL((k_matfree_mf64s<L_, 1>), grid, block, 0, st, J, sys, nch, chunk, Wm); \
        else if (var == 1) hipLaunchKernelGGL((k_matfree_mf64s<L_, 2>), grid, block, 0, st, J, sys, nch, chunk, Wm); \
        else if (var == 2) hipLaunchKernelGGL((k_matfree_mf64<L_, 2>), grid, block, 0, st, J, sys, nch, chunk, Wm); \
        else hipLaunchKernelGGL((k_matfree_mf64<L_, 1>), grid, block, 0, st, J, sys, nch, chunk, Wm);         \
    } while (0)
        if (J.lmax == 1) MPA_MF64(1);
        else if (J.lmax == 2) MPA_MF64(2);
        else if (J.lmax == 3) MPA_MF64(3);
        else if (J.lmax == 4 && var == 0)  // two workgroups per CU: 9.86 against 10.49 ms for mf64s<4, 1>
            hipLaunchKernelGGL((k_matfree_mf64k<4>), grid, block, 0, st, J, sys, nch, chunk, Wm);
        else if (J.lmax == 4) MPA_MF64(4);
        else hipLaunchKernelGGL((k_matfree_mf64k<5>), grid, block, 0, st, J, sys, nch, chunk, Wm);
#undef MPA_MF64
        hipLaunchKernelGGL(k_matfree_sum, dim3(sum_blocks), dim3(kThreads), 0, st, J, sys, nch);
        MPA_LAUNCH_CHECK();
        return MPA_OK;
    }
    if (fac_path(J.N, J.lmax)) {
        const int ne = fac_entries_of(J.lmax);
        if (ne != T->nentries) {
            mpa_set_error("matfree: contraction table has %d entries, the kernel %d", T->nentries, ne);
            return MPA_ERR_ARG;
        }
        double2* W = zbuf + total;
        const int nch = fac_chunks(J.N);
        const int chunk = (J.N + nch - 1) / nch;
        const int groups = (J.N + kThreads - 1) / kThreads;
        const size_t wtot = (size_t)J.S * J.N * ne;
        ProfScope prof(4, st);
        hipLaunchKernelGGL(k_matfree_w, dim3((unsigned)std::min<size_t>(4096, (wtot + kThreads - 1) / kThreads)),
                           dim3(kThreads), 0, st, J, sys, ne, T->wmeta, T->wsrc, T->wcoef, W);
        const dim3 grid((unsigned)(groups * nch), (unsigned)J.S), block(kThreads);
        if (J.lmax == 1) hipLaunchKernelGGL(k_matfree_fac<1>, grid, block, 0, st, J, sys, nch, chunk, W);
        else if (J.lmax == 2) hipLaunchKernelGGL(k_matfree_fac<2>, grid, block, 0, st, J, sys, nch, chunk, W);
        else hipLaunchKernelGGL(k_matfree_fac<3>, grid, block, 0, st, J, sys, nch, chunk, W);
        hipLaunchKernelGGL(k_matfree_sum, dim3(sum_blocks), dim3(kThreads), 0, st, J, sys, nch);
        MPA_LAUNCH_CHECK();
        return MPA_OK;
    }
    const int nchunks = chunks_for(J.N);
    const int chunk = (J.N + nchunks - 1) / nchunks;
    const long long nblocks = (long long)J.N * nchunks;
    const TabDev tab{T->cmeta, T->ccoef, T->clidx};
    const int SG = sg_for(J.S);
    const size_t lds = lds_bytes(J.lmax, SG, T->nterms);
    const bool tl = J.lmax <= kTabLdsLmax;
    const dim3 grid((unsigned)nblocks), block(kThreads);
    ProfScope prof(4, st);
    if (SG == 1 && tl) hipLaunchKernelGGL((k_matfree<1, true>), grid, block, lds, st, J, sys, nchunks, chunk, tab, T->nterms);
    else if (SG == 1) hipLaunchKernelGGL((k_matfree<1, false>), grid, block, lds, st, J, sys, nchunks, chunk, tab, T->nterms);
    else if (SG == 2 && tl) hipLaunchKernelGGL((k_matfree<2, true>), grid, block, lds, st, J, sys, nchunks, chunk, tab, T->nterms);
    else if (SG == 2) hipLaunchKernelGGL((k_matfree<2, false>), grid, block, lds, st, J, sys, nchunks, chunk, tab, T->nterms);
    else if (tl) hipLaunchKernelGGL((k_matfree<kSMax, true>), grid, block, lds, st, J, sys, nchunks, chunk, tab, T->nterms);
    else hipLaunchKernelGGL((k_matfree<kSMax, false>), grid, block, lds, st, J, sys, nchunks, chunk, tab, T->nterms);
    hipLaunchKernelGGL(k_matfree_sum, dim3(sum_blocks), dim3(kThreads), 0, st, J, sys, nchunks);
    MPA_LAUNCH_CHECK();
    return MPA_OK;
}

int matfree_launch_many(const std::vector<MfJob>& jobs, const SysDev* sys, hipStream_t st) {
    std::vector<const MfJob*> shared;
    for (const MfJob& j : jobs) {
        if (j.c64 && (j.lmax == 5 || (mf_shared_for(j.S, j.lmax) && mpa_option(MPA_OPT_MATFREE_C64_KERNEL) == 0)) &&
            matfree_c64_supported(j.N, j.lmax, j.tmat != nullptr)) {
            shared.push_back(&j);
            continue;
        }
        const int rc = matfree_launch(j, sys, st);
        if (rc) return rc;
    }
    return shared.empty() ? MPA_OK : matfree_launch_c64_shared(shared.data(), (int)shared.size(), sys, st);
}

}  // namespace mpa

using namespace mpa;

extern "C" size_t mpa_matfree_workspace_bytes(int N, int lmax, int nrhs) {
    if (N < 1 || nrhs < 1 || lmax < 1 || lmax > MPA_LMAX_MAX) return 0;
    return matfree_workspace_bytes(N, lmax, nrhs);
}

extern "C" int mpa_matfree_matvec_c64(const mpa_cluster_operator* op, int nrhs, int add_identity, const void* x,
                                      void* y, void* stream) {
    if (!op || op->N < 1 || nrhs < 1 || !op->positions || !op->mie || op->tmatrix || !op->workspace || !x || !y ||
        !(op->k > 0) || op->lmax < 1 || op->lmax > MPA_LMAX_MAX) {
        mpa_set_error("matfree_matvec_c64: bad argument (spheres only)");
        return MPA_ERR_ARG;
    }
    MfJob J{op->N, op->lmax, nrhs, add_identity ? 1 : 0, op->positions, (const double2*)op->mie, op->k,
            (const double2*)x, (double2*)y, (double2*)op->workspace, -1, 0, nullptr, nullptr, 1, 0};
    return matfree_launch(J, nullptr, (hipStream_t)stream);
}

extern "C" int mpa_matfree_matvec_batched(int count, const mpa_cluster_operator* const* ops, const int* nrhs, int dtype,
                                          int add_identity, const void* const* x, void* const* y, void* stream) {
    if (count < 1 || !ops || !nrhs || !x || !y || (dtype != MPA_C128 && dtype != MPA_C64)) {
        mpa_set_error("matfree_matvec_batched: bad argument");
        return MPA_ERR_ARG;
    }
    std::vector<MfJob> jobs;
    for (int q = 0; q < count; q++) {
        const mpa_cluster_operator* op = ops[q];
        if (!op || op->N < 1 || nrhs[q] < 1 || !op->positions || !op->mie || op->tmatrix || !op->workspace || !x[q] ||
            !y[q] || !(op->k > 0) || op->lmax < 1 || op->lmax > MPA_LMAX_MAX ||
            (dtype == MPA_C64 && !matfree_c64_supported(op->N, op->lmax, false))) {
            mpa_set_error("matfree_matvec_batched: bad operator %d (spheres only; complex64: lmax <= %d)", q,
                          kC64MaxLmax);
            return MPA_ERR_ARG;
        }
        if (!mpa_get_table(op->lmax)) return MPA_ERR_ARG;
        jobs.push_back(MfJob{op->N, op->lmax, nrhs[q], add_identity ? 1 : 0, op->positions, (const double2*)op->mie,
                             op->k, (const double2*)x[q], (double2*)y[q], (double2*)op->workspace, -1, 0, nullptr,
                             nullptr, dtype == MPA_C64 ? 1 : 0, 0});
    }
    return matfree_launch_many(jobs, nullptr, (hipStream_t)stream);
}

extern "C" int mpa_matfree_matvec(const mpa_cluster_operator* op, int nrhs, int add_identity, const void* x, void* y,
                                  void* stream) {
    if (!op || op->N < 1 || nrhs < 1 || !op->positions || !op->mie == !op->tmatrix || !op->workspace || !x || !y ||
        !(op->k > 0) ||
        op->lmax < 1 || op->lmax > MPA_LMAX_MAX) {
        mpa_set_error("matfree_matvec: bad argument");
        return MPA_ERR_ARG;
    }
    if ((long long)op->N * chunks_for(op->N) > 0x7fffffffLL || (long long)2 * lmax_to_rmax(op->lmax) * op->N > 0x7fffffffLL) {
        mpa_set_error("matfree_matvec: N=%d too large for one launch", op->N);
        return MPA_ERR_ARG;
    }
    MfJob J{op->N, op->lmax, nrhs, add_identity ? 1 : 0, op->positions, (const double2*)op->mie, op->k,
            (const double2*)x, (double2*)y, (double2*)op->workspace, -1, 0, (const double2*)op->tmatrix, nullptr};
    return matfree_launch(J, nullptr, (hipStream_t)stream);
}

// Variant of the matrix-core contraction kernels in use for dtype; set_to >=
// 0 selects one; *count_out (may be NULL) receives the number of variants.
extern "C" int mpa_matfree_variant(int dtype, int set_to, int* count_out) {
    if (dtype != MPA_C128 && dtype != MPA_C64) return -1;
    const int idx = dtype == MPA_C64 ? 0 : 1;
    if (count_out) *count_out = kMfVariants[idx];
    if (set_to >= kMfVariants[idx]) return -1;
    if (set_to >= 0) g_mf_variant[idx] = set_to;
    return g_mf_variant[idx];
}

// Autotuner for the matrix-core contraction kernels: times each variant on
// the caller's operator and vectors (y is overwritten) and keeps the fastest
// for the element type of x (c64 or c128).
extern "C" int mpa_autotune_matfree(const mpa_cluster_operator* op, int nrhs, int dtype, const void* x, void* y,
                                    int reps, int* variant_out, double* ms_out, void* stream) {
    if (!op || nrhs < 1 || !x || !y || reps < 1 || (dtype != MPA_C128 && dtype != MPA_C64)) {
        mpa_set_error("autotune_matfree: bad argument");
        return MPA_ERR_ARG;
    }
    hipStream_t st = (hipStream_t)stream;
    hipEvent_t e0, e1;
    MPA_HIP_CHECK(hipEventCreate(&e0));
    MPA_HIP_CHECK(hipEventCreate(&e1));
    const int idx = dtype == MPA_C64 ? 0 : 1;
    const int saved = g_mf_variant[idx];
    int best = 0, rc = MPA_OK;
    double best_ms = 1e30;
    for (int v = 0; v < kMfVariants[idx] && rc == MPA_OK; v++) {
        g_mf_variant[idx] = v;
        auto run = [&]() {
            return dtype == MPA_C64 ? mpa_matfree_matvec_c64(op, nrhs, 1, x, y, stream)
                                    : mpa_matfree_matvec(op, nrhs, 1, x, y, stream);
        };
        rc = run();
        if (rc) break;
        (void)hipEventRecord(e0, st);
        for (int r = 0; r < reps && rc == MPA_OK; r++) rc = run();
        (void)hipEventRecord(e1, st);
        if (hipEventSynchronize(e1) != hipSuccess) rc = MPA_ERR_HIP;
        float ms = 0;
        if (rc == MPA_OK && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = MPA_ERR_HIP;
        if (rc == MPA_OK && ms / reps < best_ms) {
            best_ms = ms / reps;
            best = v;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    g_mf_variant[idx] = rc == MPA_OK ? best : saved;
    if (rc) return rc;
    if (variant_out) *variant_out = best;
    if (ms_out) *ms_out = best_ms;
    return MPA_OK;
}
