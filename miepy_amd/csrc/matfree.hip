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

#include "matfree.hpp"
#include "translation.hpp"

namespace mpa {
namespace {

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
    const double2* zpre = J.tmat ? J.part + (size_t)nchunks * J.S * n : nullptr;  // z = Tm x
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

size_t matfree_workspace_bytes(int N, int lmax, int S) {
    return sizeof(double2) * (size_t)2 * lmax_to_rmax(lmax) * N * S * (chunks_for(N) + 1);
}

}  // namespace mpa

int mpa_internal_tmatrix_apply(int N, int lmax, int nrhs, const double* tmatrix, const void* x, void* y,
                               hipStream_t st);

namespace mpa {

int matfree_launch(const MfJob& J, const SysDev* sys, hipStream_t st) {
    const TransTable* T = mpa_get_table(J.lmax);
    if (!T) return MPA_ERR_ARG;
    const int nchunks = chunks_for(J.N);
    const int chunk = (J.N + nchunks - 1) / nchunks;
    const long long nblocks = (long long)J.N * nchunks;
    const size_t total = (size_t)J.S * 2 * lmax_to_rmax(J.lmax) * J.N;
    const unsigned sum_blocks = (unsigned)std::min<size_t>(2048, (total + kThreads - 1) / kThreads);
    if (J.tmat) {
        const size_t n = (size_t)2 * lmax_to_rmax(J.lmax) * J.N;
        double2* z = J.part + (size_t)nchunks * J.S * n;
        const int rc = mpa_internal_tmatrix_apply(J.N, J.lmax, J.S, (const double*)J.tmat, J.x, z, st);
        if (rc) return rc;
    }
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

}  // namespace mpa

using namespace mpa;

extern "C" size_t mpa_matfree_workspace_bytes(int N, int lmax, int nrhs) {
    if (N < 1 || nrhs < 1 || lmax < 1 || lmax > MPA_LMAX_MAX) return 0;
    return matfree_workspace_bytes(N, lmax, nrhs);
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
            (const double2*)x, (double2*)y, (double2*)op->workspace, -1, 0, (const double2*)op->tmatrix};
    return matfree_launch(J, nullptr, (hipStream_t)stream);
}
