// Aggregate interaction matrix assembly for sphere clusters on gfx950.
//
// Replaces cpp/src/interactions.cpp:370-415 (sphere_aggregate_tmatrix) and the
// per-pair translation of cpp/src/vsh_translation.cpp:161-237.
//
// Output: T[n, n] row-major, n = 2 rmax N, entry
//   T[(i,a,r(n,m)), (j,b,r(v,u))] = X_{(a+b)%2}^{ij}[nm, vu] * mie[j, b, v-1]
// for i != j and 0 on the diagonal blocks (plus 1 on the diagonal when
// add_identity is set, giving the system matrix I + T of interactions.cpp:351-353).
//
// Design: one workgroup owns block row i and a tile of J consecutive column
// blocks j0..j0+J-1.  Phase 1 forms, per pair, the products
// HL[p][mu] = h_p(k d) P_p^mu(cos theta) and e^{i mu phi} in LDS
// (translation.hpp), with the compact coefficient table staged in LDS next
// to them (lmax <= 3; larger tables are read through the caches).  Phase 2
// gives each lane one (pair, column mode s) and walks the row modes r: the
// lane forms X_A[r,s] and X_B[r,s] once and writes the four entries they
// make, (a,b) = (0,0) X_A mie_b0, (0,1) X_B mie_b1, (1,0) X_B mie_b0,
// (1,1) X_A mie_b1.  For a fixed r the lanes of a wave cover contiguous runs
// of the row, so every entry of the matrix is written exactly once,
// coalesced, with no memset and no mirrored writes.  Both (i,j) and (j,i)
// blocks are produced by their own row owners; the translation work that
// the reference shares through its symmetry is recomputed (the kernel is
// bound by HBM write bandwidth, not by these FP64 flops).

#include <algorithm>
#include <vector>

#include "translation.hpp"

using namespace mpa;

namespace {

constexpr int kThreads = 256;
constexpr int kJMax = 32;
constexpr int kTabLdsLmax = 3;  // largest lmax whose table is staged in LDS

template <typename OutT>
__device__ inline void store(OutT* p, double2 v);
template <>
__device__ inline void store<double2>(double2* p, double2 v) {
    *p = v;
}
template <>
__device__ inline void store<float2>(float2* p, double2 v) {
    *p = make_float2((float)v.x, (float)v.y);
}

template <typename OutT, bool DUAL>
__device__ inline void put(OutT* out, float2* out32, size_t off, double2 v) {
    store<OutT>(out + off, v);
    if (DUAL) store<float2>(out32 + off, v);
}

// LDS bytes: HL, e^{i mu phi}, Mie coefficients of the J column particles,
// then (TAB_LDS) the compact table.
size_t lds_bytes(int lmax, int J, int nterms, bool tab_lds) {
    const size_t b = sizeof(double2) * ((size_t)J * hl_count(lmax) + (size_t)J * (2 * lmax + 1) + (size_t)J * 2 * lmax);
    return b + (tab_lds ? tab_lds_bytes(lmax_to_rmax(lmax), nterms) : 0);
}

// OutT: element type of `out`; DUAL: also write the complex64 rounding of
// every entry to `out32` in the same pass (the mixed-precision solve needs
// both matrices; the FP64 arithmetic is done once); TAB_LDS: stage the
// coefficient table in LDS.
// The work of workgroup `blk` of one cluster's assembly grid (N x ntiles).
template <typename OutT, bool DUAL, bool TAB_LDS>
__device__ __forceinline__ void assemble_body(int blk, double2* lds, int N, int lmax, int J, int ntiles,
                                              const double* __restrict__ pos, const double2* __restrict__ mie,
                                              double k, TabDev gtab, int nterms, int add_identity,
                                              OutT* __restrict__ out, float2* __restrict__ out32) {
    const int rmax = lmax_to_rmax(lmax);
    const int bs = 2 * rmax;
    const size_t n = (size_t)bs * N;
    const int pmax = 2 * lmax + 1;
    const int Lc = hl_count(lmax);
    const int nmu = 2 * lmax + 1;
    double2* HL = lds;
    double2* eph = HL + J * Lc;
    double2* mieJ = eph + J * nmu;

    const int i = blk / ntiles;
    const int j0 = (blk % ntiles) * J;
    const int jcount = min(J, N - j0);

    TabDev tab = gtab;
    if (TAB_LDS) tab = stage_table(gtab, rmax, nterms, (char*)(mieJ + J * 2 * lmax));
    // phase 1: per-pair special functions
    for (int t = threadIdx.x; t < jcount * (pmax + 1); t += kThreads) {
        const int jj = t / (pmax + 1), mu = t - jj * (pmax + 1);
        const int j = j0 + jj;
        if (j != i)
            pair_hl(HL + jj * Lc, eph + jj * nmu, mu, lmax, pos[3 * i] - pos[3 * j], pos[3 * i + 1] - pos[3 * j + 1],
                    pos[3 * i + 2] - pos[3 * j + 2], k);
    }
    for (int t = threadIdx.x; t < jcount * 2 * lmax; t += kThreads)
        mieJ[t] = mie ? mie[(size_t)j0 * 2 * lmax + t] : make_double2(1.0, 0.0);  // null: the bare X blocks
    __syncthreads();

    // phase 2: lane (pair jj, column mode s) walks the row modes r
    const size_t row0 = (size_t)i * bs;
    const size_t col0 = (size_t)j0 * bs;
    for (int c = threadIdx.x; c < jcount * rmax; c += kThreads) {
        const int jj = c / rmax;
        const int s = c - jj * rmax;
        const int j = j0 + jj;
        const size_t base = row0 * n + col0 + (size_t)jj * bs + s;  // row (0, 0), column (b = 0, s)
        if (j == i) {
            for (int r = 0; r < rmax; r++) {
                const double2 d = make_double2((add_identity && r == s) ? 1.0 : 0.0, 0.0);
                const double2 z = make_double2(0.0, 0.0);
                put<OutT, DUAL>(out, out32, base + (size_t)r * n, d);
                put<OutT, DUAL>(out, out32, base + (size_t)r * n + rmax, z);
                put<OutT, DUAL>(out, out32, base + (size_t)(rmax + r) * n, z);
                put<OutT, DUAL>(out, out32, base + (size_t)(rmax + r) * n + rmax, d);
            }
            continue;
        }
        const int v = (int)sqrtf((float)(s + 1) + 0.5f);
        const double2 m0 = mieJ[(jj * 2 + 0) * lmax + v - 1];
        const double2 m1 = mieJ[(jj * 2 + 1) * lmax + v - 1];
        const double2* hl = HL + jj * Lc;
        const double2* ep = eph + jj * nmu;
        for (int r = 0; r < rmax; r++) {
            double2 xa, xb;
            tuple_xab(tab, r * rmax + s, hl, ep, xa, xb);
            put<OutT, DUAL>(out, out32, base + (size_t)r * n, cmul(xa, m0));
            put<OutT, DUAL>(out, out32, base + (size_t)r * n + rmax, cmul(xb, m1));
            put<OutT, DUAL>(out, out32, base + (size_t)(rmax + r) * n, cmul(xb, m0));
            put<OutT, DUAL>(out, out32, base + (size_t)(rmax + r) * n + rmax, cmul(xa, m1));
        }
    }
}

template <typename OutT, bool DUAL, bool TAB_LDS>
__global__ __launch_bounds__(kThreads) void assemble_kernel(int N, int lmax, int J, int ntiles,
                                                             const double* __restrict__ pos,
                                                             const double2* __restrict__ mie, double k, TabDev gtab,
                                                             int nterms, int add_identity, OutT* __restrict__ out,
                                                             float2* __restrict__ out32) {
    extern __shared__ double2 lds[];
    assemble_body<OutT, DUAL, TAB_LDS>(blockIdx.x, lds, N, lmax, J, ntiles, pos, mie, k, gtab, nterms, add_identity,
                                       out, out32);
}

// ---- many clusters per launch: descriptors by value in the kernel arguments
// (kAsmBatchMax per launch), workgroups dealt to the clusters in order; each
// cluster keeps the tile variant of its own launch, and every entry is
// formed by the same arithmetic whatever the tiling, so a batched matrix
// equals the single-cluster one bit for bit.
constexpr int kAsmBatchMax = 32;

struct AsmJob {
    int N, lmax, J, ntiles, nterms, add_identity, tab_lds, pad;
    const double* pos;
    const double2* mie;
    double k;
    TabDev tab;
    void* out;
    float2* out32;
};

struct AsmBatch {
    int count;
    int first[kAsmBatchMax + 1];  // prefix of each cluster's workgroups
    AsmJob job[kAsmBatchMax];
};

template <typename OutT, bool DUAL>
__global__ __launch_bounds__(kThreads) void assemble_kernel_b(AsmBatch B) {
    extern __shared__ double2 lds[];
    int q = 0;
    while (q + 1 < B.count && (int)blockIdx.x >= B.first[q + 1]) q++;
    const AsmJob& A = B.job[q];
    const int blk = blockIdx.x - B.first[q];
    if (A.tab_lds)
        assemble_body<OutT, DUAL, true>(blk, lds, A.N, A.lmax, A.J, A.ntiles, A.pos, A.mie, A.k, A.tab, A.nterms,
                                        A.add_identity, (OutT*)A.out, A.out32);
    else
        assemble_body<OutT, DUAL, false>(blk, lds, A.N, A.lmax, A.J, A.ntiles, A.pos, A.mie, A.k, A.tab, A.nterms,
                                         A.add_identity, (OutT*)A.out, A.out32);
}

// Column blocks per tile: the J <= 32 that keeps the most lanes busy in
// phase 2 (J * rmax lanes over 256); ties go to the larger J.
int pick_tile(int lmax) {
    const int rmax = lmax_to_rmax(lmax);
    int best = 1;
    double best_eff = -1;
    for (int J = 1; J <= kJMax; J++) {
        const int W = J * rmax;
        const int iters = (W + kThreads - 1) / kThreads;
        const double eff = (double)W / (iters * kThreads);
        if (eff >= best_eff - 1e-12) {
            best = J;
            best_eff = eff;
        }
    }
    return best;
}

// Tile variants per lmax (chosen by mpa_autotune_assembly): column blocks
// per tile J and whether the coefficient table is staged in LDS.
struct AsmVariant {
    int J;
    bool tab_lds;
};
constexpr size_t kLdsLimit = 65536;
int g_asm_variant[MPA_LMAX_MAX + 1] = {};

std::vector<AsmVariant> asm_variants(int lmax, int nterms) {
    const int J0 = pick_tile(lmax);
    std::vector<AsmVariant> v;
    for (int J : {J0, (J0 + 1) / 2})
        for (bool tl : {lmax <= kTabLdsLmax, !(lmax <= kTabLdsLmax)}) {
            if (lds_bytes(lmax, J, nterms, tl) > kLdsLimit) continue;
            bool dup = false;
            for (const AsmVariant& w : v) dup |= (w.J == J && w.tab_lds == tl);
            if (!dup) v.push_back({J, tl});
        }
    return v;
}

template <typename OutT, bool DUAL>
int launch_assembly(int N, int lmax, const double* positions, const double* mie, double k, int add_identity,
                    OutT* out, float2* out32, hipStream_t s, const char* who) {
    const TransTable* T = mpa_get_table(lmax);
    if (!T) return MPA_ERR_ARG;
    const std::vector<AsmVariant> vars = asm_variants(lmax, T->nterms);
    const AsmVariant var = vars[std::min<size_t>(g_asm_variant[lmax], vars.size() - 1)];
    const int J = var.J;
    const int ntiles = (N + J - 1) / J;
    const long long nblocks = (long long)N * ntiles;
    if (nblocks > 0x7fffffffLL) {
        mpa_set_error("%s: N=%d too large for one launch", who, N);
        return MPA_ERR_ARG;
    }
    const TabDev tab{T->cmeta, T->ccoef, T->clidx};
    const bool tab_lds = var.tab_lds;
    const size_t lds = lds_bytes(lmax, J, T->nterms, tab_lds);
    if (tab_lds)
        hipLaunchKernelGGL((assemble_kernel<OutT, DUAL, true>), dim3((unsigned)nblocks), dim3(kThreads), lds, s, N,
                           lmax, J, ntiles, positions, (const double2*)mie, k, tab, T->nterms, add_identity, out,
                           out32);
    else
        hipLaunchKernelGGL((assemble_kernel<OutT, DUAL, false>), dim3((unsigned)nblocks), dim3(kThreads), lds, s, N,
                           lmax, J, ntiles, positions, (const double2*)mie, k, tab, T->nterms, add_identity, out,
                           out32);
    MPA_LAUNCH_CHECK();
    return MPA_OK;
}

// ---- particles with general T-matrices --------------------------------------
//
// T_ij = X^{ij} Tm_j (cpp/src/vsh_translation.cpp:53-159: the pair's bare
// translation blocks times the column particle's T-matrix; the reference
// forms the product entry by entry for every pair inside an n x n scratch
// matrix).  Here the bare X blocks are assembled in place (unit Mie
// coefficients) and every column block is then multiplied on the right by
// its particle's T-matrix: one workgroup per (column block j, 32 rows)
// stages the old 32 x bs values and (bs <= 48) Tm_j in LDS, forms the
// products in FP64 and writes the block back; the identity is added on the
// diagonal blocks (which are zero before the product).

constexpr int kApplyRows = 32;
constexpr int kApplyTmLdsMax = 48;  // bs up to which Tm_j is staged in LDS (lmax <= 4)

template <typename OutT>
__device__ inline double2 load_c(const OutT* p);
template <>
__device__ inline double2 load_c<double2>(const double2* p) {
    return *p;
}
template <>
__device__ inline double2 load_c<float2>(const float2* p) {
    const float2 v = *p;
    return make_double2(v.x, v.y);
}

template <typename OutT>
__global__ __launch_bounds__(kThreads) void k_apply_tmatrix(int N, int bs, int nrowblocks,
                                                             const double2* __restrict__ tm, int add_identity,
                                                             OutT* __restrict__ out) {
    extern __shared__ double2 lds[];
    const int j = blockIdx.x / nrowblocks;
    const int r0 = (blockIdx.x % nrowblocks) * kApplyRows;
    const size_t n = (size_t)bs * N;
    const int nr = (int)min((size_t)kApplyRows, n - r0);
    double2* old = lds;                         // [kApplyRows][bs]
    double2* tml = old + kApplyRows * bs;       // [bs][bs] when staged
    const double2* tmj = tm + (size_t)j * bs * bs;
    const bool staged = bs <= kApplyTmLdsMax;
    for (int u = threadIdx.x; u < nr * bs; u += kThreads) {
        const int rr = u / bs, c = u - rr * bs;
        old[u] = load_c<OutT>(out + (size_t)(r0 + rr) * n + (size_t)j * bs + c);
    }
    if (staged)
        for (int u = threadIdx.x; u < bs * bs; u += kThreads) tml[u] = tmj[u];
    __syncthreads();
    const double2* T = staged ? tml : tmj;
    for (int u = threadIdx.x; u < nr * bs; u += kThreads) {
        const int rr = u / bs, c = u - rr * bs;
        double2 acc = make_double2(0.0, 0.0);
        for (int q = 0; q < bs; q++) cfma(acc, old[rr * bs + q], T[q * bs + c]);
        const size_t row = (size_t)r0 + rr;
        if (add_identity && row == (size_t)j * bs + c) acc.x += 1.0;
        store<OutT>(out + row * n + (size_t)j * bs + c, acc);
    }
}

template <typename OutT>
int launch_apply_tmatrix(int N, int lmax, const double* tmatrix, int add_identity, OutT* out, hipStream_t s) {
    const int bs = 2 * lmax_to_rmax(lmax);
    const long long n = (long long)bs * N;
    const int nrb = (int)((n + kApplyRows - 1) / kApplyRows);
    const long long nblocks = (long long)N * nrb;
    if (nblocks > 0x7fffffffLL) {
        mpa_set_error("particle_aggregate_tmatrix: N=%d too large for one launch", N);
        return MPA_ERR_ARG;
    }
    const size_t lds = sizeof(double2) * ((size_t)kApplyRows * bs + (bs <= kApplyTmLdsMax ? (size_t)bs * bs : 0));
    hipLaunchKernelGGL(k_apply_tmatrix<OutT>, dim3((unsigned)nblocks), dim3(kThreads), lds, s, N, bs, nrb,
                       (const double2*)tmatrix, add_identity, out);
    MPA_LAUNCH_CHECK();
    return MPA_OK;
}

}  // namespace

extern "C" int mpa_particle_aggregate_tmatrix(int N, int lmax, const double* positions, const double* tmatrix,
                                              double k, int out_dtype, int add_identity, void* out, void* stream) {
    if (N < 1 || !positions || !tmatrix || !out || !(k > 0)) {
        mpa_set_error("particle_aggregate_tmatrix: bad argument (N=%d, k=%g)", N, k);
        return MPA_ERR_ARG;
    }
    if (out_dtype != MPA_C128 && out_dtype != MPA_C64) {
        mpa_set_error("particle_aggregate_tmatrix: unknown dtype %d", out_dtype);
        return MPA_ERR_ARG;
    }
    hipStream_t s = (hipStream_t)stream;
    ProfScope prof(out_dtype == MPA_C128 ? 2 : 3, s);
    int rc;
    if (out_dtype == MPA_C128) {
        rc = launch_assembly<double2, false>(N, lmax, positions, nullptr, k, 0, (double2*)out, nullptr, s,
                                             "particle_aggregate_tmatrix");
        if (rc == MPA_OK) rc = launch_apply_tmatrix<double2>(N, lmax, tmatrix, add_identity, (double2*)out, s);
    } else {
        rc = launch_assembly<float2, false>(N, lmax, positions, nullptr, k, 0, (float2*)out, nullptr, s,
                                            "particle_aggregate_tmatrix");
        if (rc == MPA_OK) rc = launch_apply_tmatrix<float2>(N, lmax, tmatrix, add_identity, (float2*)out, s);
    }
    return rc;
}

extern "C" int mpa_sphere_aggregate_tmatrix(int N, int lmax, const double* positions, const double* mie, double k,
                                            int out_dtype, int add_identity, void* out, void* stream) {
    if (N < 1 || !positions || !mie || !out || !(k > 0)) {
        mpa_set_error("sphere_aggregate_tmatrix: bad argument (N=%d, k=%g)", N, k);
        return MPA_ERR_ARG;
    }
    if (out_dtype != MPA_C128 && out_dtype != MPA_C64) {
        mpa_set_error("sphere_aggregate_tmatrix: unknown dtype %d", out_dtype);
        return MPA_ERR_ARG;
    }
    hipStream_t s = (hipStream_t)stream;
    ProfScope prof(out_dtype == MPA_C128 ? 2 : 3, s);
    if (out_dtype == MPA_C128)
        return launch_assembly<double2, false>(N, lmax, positions, mie, k, add_identity, (double2*)out, nullptr, s,
                                               "sphere_aggregate_tmatrix");
    return launch_assembly<float2, false>(N, lmax, positions, mie, k, add_identity, (float2*)out, nullptr, s,
                                          "sphere_aggregate_tmatrix");
}

extern "C" int mpa_assembly_tile(int lmax) {
    const TransTable* T = mpa_get_table(lmax);
    if (!T) return -1;
    const std::vector<AsmVariant> vars = asm_variants(lmax, T->nterms);
    return vars[std::min<size_t>(g_asm_variant[lmax], vars.size() - 1)].J;
}

// Variant in use for lmax (set_to >= 0 selects one); -1 for a bad lmax or
// variant.  *count_out (may be NULL) receives the number of variants.
extern "C" int mpa_assembly_variant(int lmax, int set_to, int* count_out) {
    const TransTable* T = mpa_get_table(lmax);
    if (!T) return -1;
    const int nv = (int)asm_variants(lmax, T->nterms).size();
    if (count_out) *count_out = nv;
    if (set_to >= nv) return -1;
    if (set_to >= 0) g_asm_variant[lmax] = set_to;
    return g_asm_variant[lmax];
}

// Autotuner: time every tile variant for this lmax on the caller's cluster
// (out is overwritten) and keep the fastest for the process.
extern "C" int mpa_autotune_assembly(int N, int lmax, const double* positions, const double* mie, double k,
                                     int out_dtype, void* out, int reps, int* variant_out, double* ms_out,
                                     void* stream) {
    if (N < 1 || !positions || !mie || !out || !(k > 0) || reps < 1 || (out_dtype != MPA_C128 && out_dtype != MPA_C64)) {
        mpa_set_error("autotune_assembly: bad argument");
        return MPA_ERR_ARG;
    }
    const TransTable* T = mpa_get_table(lmax);
    if (!T) return MPA_ERR_ARG;
    const int nv = (int)asm_variants(lmax, T->nterms).size();
    hipStream_t st = (hipStream_t)stream;
    hipEvent_t e0, e1;
    MPA_HIP_CHECK(hipEventCreate(&e0));
    MPA_HIP_CHECK(hipEventCreate(&e1));
    const int saved = g_asm_variant[lmax];
    int best = 0, rc = MPA_OK;
    double best_ms = 1e30;
    for (int v = 0; v < nv && rc == MPA_OK; v++) {
        g_asm_variant[lmax] = v;
        rc = mpa_sphere_aggregate_tmatrix(N, lmax, positions, mie, k, out_dtype, 0, out, stream);  // warm-up
        if (rc) break;
        (void)hipEventRecord(e0, st);
        for (int r = 0; r < reps && rc == MPA_OK; r++)
            rc = mpa_sphere_aggregate_tmatrix(N, lmax, positions, mie, k, out_dtype, 0, out, stream);
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
    g_asm_variant[lmax] = rc == MPA_OK ? best : saved;
    if (rc) return rc;
    if (variant_out) *variant_out = best;
    if (ms_out) *ms_out = best_ms;
    return MPA_OK;
}

extern "C" int mpa_sphere_aggregate_tmatrix_dual(int N, int lmax, const double* positions, const double* mie, double k,
                                                 int add_identity, void* out64, void* out32, void* stream) {
    if (N < 1 || !positions || !mie || !out64 || !out32 || !(k > 0)) {
        mpa_set_error("sphere_aggregate_tmatrix_dual: bad argument (N=%d, k=%g)", N, k);
        return MPA_ERR_ARG;
    }
    hipStream_t s = (hipStream_t)stream;
    ProfScope prof(2, s);
    return launch_assembly<double2, true>(N, lmax, positions, mie, k, add_identity, (double2*)out64, (float2*)out32,
                                          s, "sphere_aggregate_tmatrix_dual");
}

// ---- many clusters --------------------------------------------------------------

namespace {

template <typename OutT, bool DUAL>
int launch_assembly_batched(const std::vector<const mpa_asm_cluster*>& cs, hipStream_t s) {
    for (size_t b = 0; b < cs.size(); b += kAsmBatchMax) {
        const int cnt = (int)std::min<size_t>(kAsmBatchMax, cs.size() - b);
        AsmBatch B{};
        B.count = cnt;
        size_t lds = 0;
        for (int q = 0; q < cnt; q++) {
            const mpa_asm_cluster& C = *cs[b + q];
            const TransTable* T = mpa_get_table(C.lmax);
            if (!T) return MPA_ERR_ARG;
            const std::vector<AsmVariant> vars = asm_variants(C.lmax, T->nterms);
            const AsmVariant var = vars[std::min<size_t>(g_asm_variant[C.lmax], vars.size() - 1)];
            AsmJob& A = B.job[q];
            A.N = C.N;
            A.lmax = C.lmax;
            A.J = var.J;
            A.ntiles = (C.N + var.J - 1) / var.J;
            A.nterms = T->nterms;
            A.add_identity = C.add_identity;
            A.tab_lds = var.tab_lds ? 1 : 0;
            A.pos = C.positions;
            A.mie = (const double2*)C.mie;
            A.k = C.k;
            A.tab = TabDev{T->cmeta, T->ccoef, T->clidx};
            A.out = C.out;
            A.out32 = (float2*)C.out32;
            const long long wg = (long long)B.first[q] + (long long)C.N * A.ntiles;
            if (wg > 0x7fffffffLL) {
                mpa_set_error("sphere_aggregate_tmatrix_batched: batch too large for one launch");
                return MPA_ERR_ARG;
            }
            B.first[q + 1] = (int)wg;
            lds = std::max(lds, lds_bytes(C.lmax, var.J, T->nterms, var.tab_lds));
        }
        hipLaunchKernelGGL((assemble_kernel_b<OutT, DUAL>), dim3((unsigned)B.first[cnt]), dim3(kThreads), lds, s, B);
        MPA_LAUNCH_CHECK();
    }
    return MPA_OK;
}

}  // namespace

extern "C" int mpa_sphere_aggregate_tmatrix_batched(int count, const mpa_asm_cluster* clusters, void* stream) {
    if (count < 0 || (count > 0 && !clusters)) {
        mpa_set_error("sphere_aggregate_tmatrix_batched: bad argument");
        return MPA_ERR_ARG;
    }
    std::vector<const mpa_asm_cluster*> c128, c64, dual;
    for (int q = 0; q < count; q++) {
        const mpa_asm_cluster& C = clusters[q];
        if (C.N < 1 || C.lmax < 1 || C.lmax > MPA_LMAX_MAX || !C.positions || !C.mie || !C.out || !(C.k > 0) ||
            (C.dtype != MPA_C128 && C.dtype != MPA_C64) || (C.out32 && C.dtype != MPA_C128)) {
            mpa_set_error("sphere_aggregate_tmatrix_batched: bad cluster %d", q);
            return MPA_ERR_ARG;
        }
        (C.out32 ? dual : C.dtype == MPA_C128 ? c128 : c64).push_back(&C);
    }
    hipStream_t s = (hipStream_t)stream;
    ProfScope prof(2, s);
    int rc = MPA_OK;
    if (!dual.empty()) rc = launch_assembly_batched<double2, true>(dual, s);
    if (rc == MPA_OK && !c128.empty()) rc = launch_assembly_batched<double2, false>(c128, s);
    if (rc == MPA_OK && !c64.empty()) rc = launch_assembly_batched<float2, false>(c64, s);
    return rc;
}
