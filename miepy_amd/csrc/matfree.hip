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
// particles and walks the chunk in batches of JB pairs.  Per batch, phase 1
// gives one thread per (pair, Legendre order mu) the products
// HL[p][mu] = h_p(k d) P_p^mu(cos theta) (same recurrences and operation
// order as the assembly) and e^{i mu phi}, and forms z_j for the batch, all
// in LDS.  Phase 2 gives each thread one (pair, row mode r); it walks the
// column modes s, forms X_A[r,s] and X_B[r,s] from the table's q-sums and
// accumulates both polarizations of y_i for up to four right-hand sides in
// registers.  At the end the JB partial rows meet in LDS in a fixed order and
// the chunk's partial row is written; a second kernel adds the chunks in a
// fixed order plus the identity term, so the result is deterministic.

#include "matfree.hpp"

namespace mpa {
namespace {

constexpr int kJB = 16;   // most pairs per batch
constexpr int kSMax = 4;  // right-hand sides per pass

__host__ __device__ inline int pairs_per_batch(int rmax) {
    const int jb = kThreads / rmax;
    return jb > kJB ? kJB : jb;
}
__host__ __device__ inline int leg_count(int lmax) {
    const int pmax = 2 * lmax + 1;
    return (pmax + 1) * (pmax + 2) / 2;
}

// LDS bytes of one workgroup: HL, e^{i mu phi}, z / reduction buffer.
inline size_t lds_bytes(int lmax, int S) {
    const int rmax = lmax_to_rmax(lmax);
    const int JB = pairs_per_batch(rmax);
    const int sg = S < kSMax ? S : kSMax;
    return sizeof(double2) * ((size_t)JB * leg_count(lmax) + (size_t)JB * (2 * lmax + 1) + (size_t)sg * 2 * JB * rmax);
}

// Chunks of column particles per block row: enough workgroups to fill the
// chip for small clusters, one chunk per row for large ones.
inline int chunks_for(int N) {
    int c = (2048 + N - 1) / N;
    if (c > 64) c = 64;
    if (c > N) c = N;
    return c < 1 ? 1 : c;
}

// Thread (pair, mu): HL[l][mu] = h_l(k r) P_l^mu(cos theta) for l = mu..pmax
// and e^{i mu phi}.  Spherical Hankel h_l = j_l + i y_l by upward recurrence
// (cpp/src/special.cpp:90-112); P_l^mu without Condon-Shortley phase
// (cpp/src/special.cpp:180-204).
__device__ void pair_hl(double2* HL, double2* eph, int mu, int lmax, double dx, double dy, double dz, double k) {
    const int pmax = 2 * lmax + 1;
    const double rad = sqrt(dx * dx + dy * dy + dz * dz);
    const double x = dz / rad;
    const double s = sqrt(fmax(0.0, (1.0 - x) * (1.0 + x)));
    double pmm = 1.0;
    for (int i = 1; i <= mu; i++) pmm *= (2 * i - 1) * s;
    const double z = k * rad;
    double sz, cz;
    sincos(z, &sz, &cz);
    double ja = sz / z, ya = -cz / z;        // h_{l-2} after the first two steps
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
        HL[leg_index(l, mu)] = cscale(make_double2(jl, yl), P);
    }
    if (mu <= 2 * lmax) {
        double sn, cs;
        sincos(mu * atan2(dy, dx), &sn, &cs);
        eph[mu] = make_double2(cs, sn);
    }
}

__device__ inline bool job_done(const SysDev& d) {
    for (int s = 0; s < d.S; s++)
        if (!d.scal[s].done) return false;
    return true;
}

// Partial rows: part[chunk][rhs][row].
__global__ __launch_bounds__(kThreads) void k_matfree(MfJob J, const SysDev* __restrict__ sys, int nchunks, int chunk,
                                                      const int4* __restrict__ tmeta,
                                                      const double2* __restrict__ tcoef) {
    extern __shared__ double2 lds[];
    if (J.skip >= 0 && job_done(sys[J.skip])) return;
    const int lmax = J.lmax;
    const int rmax = lmax_to_rmax(lmax);
    const int bs = 2 * rmax;
    const int n = bs * J.N;
    const int pmax = 2 * lmax + 1;
    const int Lc = leg_count(lmax);
    const int nmu = 2 * lmax + 1;
    const int qa_slots = lmax + 1;
    const int coef_slots = 2 * lmax + 1;
    const int JB = pairs_per_batch(rmax);
    const int zs = JB * rmax;  // z / reduction stride per (rhs, polarization)
    double2* HL = lds;
    double2* eph = HL + JB * Lc;
    double2* zr = eph + JB * nmu;

    const int i = blockIdx.x / nchunks;
    const int c = blockIdx.x % nchunks;
    const int jbeg = c * chunk, jend = min(J.N, jbeg + chunk);
    const int t = threadIdx.x;
    const int jme = t / rmax, r = t - jme * rmax;
    const bool worker = jme < JB;
    const double xi = J.pos[3 * i], yi = J.pos[3 * i + 1], zi = J.pos[3 * i + 2];

    for (int s0 = 0; s0 < J.S; s0 += kSMax) {
        const int sg = min(kSMax, J.S - s0);
        double2 acc[kSMax][2];
#pragma unroll
        for (int q = 0; q < kSMax; q++) acc[q][0] = acc[q][1] = make_double2(0, 0);
        for (int j0 = jbeg; j0 < jend; j0 += JB) {
            const int jc = min(JB, jend - j0);
            __syncthreads();
            for (int u = t; u < jc * (pmax + 1); u += kThreads) {
                const int jj = u / (pmax + 1), mu = u - jj * (pmax + 1);
                const int j = j0 + jj;
                if (j != i)
                    pair_hl(HL + jj * Lc, eph + jj * nmu, mu, lmax, xi - J.pos[3 * j], yi - J.pos[3 * j + 1],
                            zi - J.pos[3 * j + 2], J.k);
            }
            for (int u = t; u < jc * sg * bs; u += kThreads) {
                const int jj = u / (sg * bs), rest = u - jj * (sg * bs);
                const int q = rest / bs, e = rest - q * bs;
                const int b = e >= rmax ? 1 : 0, s = e - b * rmax;
                const int v = (int)sqrtf((float)(s + 1) + 0.5f);
                const int j = j0 + jj;
                zr[(q * 2 + b) * zs + jj * rmax + s] =
                    cmul(J.mie[((size_t)j * 2 + b) * lmax + v - 1], J.x[(size_t)(s0 + q) * n + (size_t)j * bs + e]);
            }
            __syncthreads();
            if (worker && jme < jc && j0 + jme != i) {
                const double2* hl = HL + jme * Lc;
                const double2* ep = eph + jme * nmu;
                const double2* zj = zr + jme * rmax;
                for (int s = 0; s < rmax; s++) {
                    const int tt = r * rmax + s;
                    const int4 meta = tmeta[tt];
                    const int amu = meta.z < 0 ? -meta.z : meta.z;
                    const double2* cf = tcoef + (size_t)tt * coef_slots;
                    double2 xa = make_double2(0, 0), xb = make_double2(0, 0);
                    for (int q = 0; q < meta.x; q++) cfma(xa, cf[q], hl[leg_index(meta.w - 2 * q, amu)]);
                    for (int q = 1; q <= meta.y; q++)
                        cfma(xb, cf[qa_slots + q - 1], hl[leg_index(meta.w - 2 * q + 1, amu)]);
                    double2 e = ep[amu];
                    if (meta.z < 0) e.y = -e.y;
                    xa = cmul(e, xa);
                    xb = cmul(e, xb);
#pragma unroll
                    for (int q = 0; q < kSMax; q++) {
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
            for (int q = 0; q < kSMax; q++)
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
    return sizeof(double2) * (size_t)2 * lmax_to_rmax(lmax) * N * S * chunks_for(N);
}

int matfree_launch(const MfJob& J, const SysDev* sys, hipStream_t st) {
    const TransTable* T = mpa_get_table(J.lmax);
    if (!T) return MPA_ERR_ARG;
    const int nchunks = chunks_for(J.N);
    const int chunk = (J.N + nchunks - 1) / nchunks;
    const long long nblocks = (long long)J.N * nchunks;
    const size_t total = (size_t)J.S * 2 * lmax_to_rmax(J.lmax) * J.N;
    const unsigned sum_blocks = (unsigned)std::min<size_t>(2048, (total + kThreads - 1) / kThreads);
    ProfScope prof(4, st);
    hipLaunchKernelGGL(k_matfree, dim3((unsigned)nblocks), dim3(kThreads), lds_bytes(J.lmax, J.S), st, J, sys, nchunks,
                       chunk, T->meta, T->coef);
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
    if (!op || op->N < 1 || nrhs < 1 || !op->positions || !op->mie || !op->workspace || !x || !y || !(op->k > 0) ||
        op->lmax < 1 || op->lmax > MPA_LMAX_MAX) {
        mpa_set_error("matfree_matvec: bad argument");
        return MPA_ERR_ARG;
    }
    if ((long long)op->N * chunks_for(op->N) > 0x7fffffffLL || (long long)2 * lmax_to_rmax(op->lmax) * op->N > 0x7fffffffLL) {
        mpa_set_error("matfree_matvec: N=%d too large for one launch", op->N);
        return MPA_ERR_ARG;
    }
    MfJob J{op->N, op->lmax, nrhs, add_identity ? 1 : 0, op->positions, (const double2*)op->mie, op->k,
            (const double2*)x, (double2*)y, (double2*)op->workspace, -1, 0};
    return matfree_launch(J, nullptr, (hipStream_t)stream);
}
