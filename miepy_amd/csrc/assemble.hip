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
// blocks j0..j0+J-1.  Phase 1 computes, per pair, h_p(k d) (FP64 upward
// recurrence, special.cpp:90-112), P_p^mu(cos theta) without Condon-Shortley
// phase (special.cpp:180-204) for mu >= 0, and e^{i mu phi} into LDS.  Phase 2
// gives each lane a fixed output column and walks the 2 rmax rows of the
// block row, so every row store of the wave is one contiguous run of
// J * 2 rmax entries: the matrix is written exactly once, fully coalesced,
// with no memset and no scattered mirror writes.  Both (i,j) and (j,i) blocks
// are produced by their own row owners; the translation work that the
// reference shares through its symmetry is recomputed (the kernel is bound by
// HBM write bandwidth, not by these FP64 flops).

#include "common.hpp"

namespace {

constexpr int kThreads = 256;
constexpr int kJMax = 16;
constexpr int kLegMax = (MPA_PMAX_MAX + 1) * (MPA_PMAX_MAX + 2) / 2;
constexpr int kMuMax = 2 * MPA_LMAX_MAX + 1;  // mu = 0 .. 2 lmax

struct PairLDS {
    double2 hp[kJMax][MPA_PMAX_MAX + 2];
    double leg[kJMax][kLegMax];
    double2 eph[kJMax][kMuMax];
    int valid[kJMax];
};

// Per-pair special functions into LDS; one thread per (pair, mu column).
__device__ void pair_setup(PairLDS& L, int jj, int mu, int lmax, double dx, double dy, double dz, double k) {
    const int pmax = 2 * lmax + 1;
    const double rad = sqrt(dx * dx + dy * dy + dz * dz);
    const double x = dz / rad;
    const double s = sqrt(fmax(0.0, (1.0 - x) * (1.0 + x)));
    // Legendre column mu: P_mu^mu = (2mu-1)!! s^mu; P_{mu+1}^mu = x (2mu+1) P_mu^mu;
    // (l-mu) P_l^mu = (2l-1) x P_{l-1}^mu - (l+mu-1) P_{l-2}^mu
    double pmm = 1.0;
    for (int i = 1; i <= mu; i++) pmm *= (2 * i - 1) * s;
    L.leg[jj][leg_index(mu, mu)] = pmm;
    if (mu + 1 <= pmax) {
        double p1 = x * (2 * mu + 1) * pmm;
        L.leg[jj][leg_index(mu + 1, mu)] = p1;
        double p2 = pmm;
        for (int l = mu + 2; l <= pmax; l++) {
            double pl = ((2 * l - 1) * x * p1 - (l + mu - 1) * p2) / (l - mu);
            L.leg[jj][leg_index(l, mu)] = pl;
            p2 = p1;
            p1 = pl;
        }
    }
    if (mu <= 2 * lmax) {
        const double phi = atan2(dy, dx);
        double sn, cs;
        sincos(mu * phi, &sn, &cs);
        L.eph[jj][mu] = make_double2(cs, sn);
    }
    if (mu == 0) {
        const double z = k * rad;
        double sz, cz;
        sincos(z, &sz, &cz);
        double j0 = sz / z, y0 = -cz / z;
        double j1 = j0 / z + y0, y1 = y0 / z - j0;
        L.hp[jj][0] = make_double2(j0, y0);
        L.hp[jj][1] = make_double2(j1, y1);
        for (int i = 2; i <= pmax; i++) {
            double j2 = (2 * i - 1) / z * j1 - j0;
            double y2 = (2 * i - 1) / z * y1 - y0;
            L.hp[jj][i] = make_double2(j2, y2);
            j0 = j1;
            j1 = j2;
            y0 = y1;
            y1 = y2;
        }
    }
}

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

// OutT: element type of `out`; DUAL: also write the complex64 rounding of
// every entry to `out32` in the same pass (the mixed-precision solve needs
// both matrices; the FP64 arithmetic is done once).
template <typename OutT, bool DUAL>
__global__ __launch_bounds__(kThreads) void assemble_kernel(int N, int lmax, int J, int ntiles,
                                                             const double* __restrict__ pos,
                                                             const double2* __restrict__ mie, double k,
                                                             const int4* __restrict__ tmeta,
                                                             const double2* __restrict__ tcoef, int add_identity,
                                                             OutT* __restrict__ out, float2* __restrict__ out32) {
    __shared__ PairLDS L;
    __shared__ double2 mieJ[kJMax][2][MPA_LMAX_MAX];

    const int rmax = lmax_to_rmax(lmax);
    const int bs = 2 * rmax;
    const size_t n = (size_t)bs * N;
    const int pmax = 2 * lmax + 1;
    const int qa_slots = lmax + 1;
    const int coef_slots = 2 * lmax + 1;

    const int i = blockIdx.x / ntiles;
    const int j0 = (blockIdx.x % ntiles) * J;
    const int jcount = min(J, N - j0);

    // phase 1: per-pair special functions
    const int ncol = pmax + 1;
    for (int t = threadIdx.x; t < jcount * ncol; t += kThreads) {
        const int jj = t / ncol, mu = t % ncol;
        const int j = j0 + jj;
        if (j != i) {
            const double dx = pos[3 * i] - pos[3 * j];
            const double dy = pos[3 * i + 1] - pos[3 * j + 1];
            const double dz = pos[3 * i + 2] - pos[3 * j + 2];
            pair_setup(L, jj, mu, lmax, dx, dy, dz, k);
        }
    }
    for (int t = threadIdx.x; t < jcount * 2 * lmax; t += kThreads) {
        const int jj = t / (2 * lmax), rest = t % (2 * lmax);
        mieJ[jj][rest / lmax][rest % lmax] = mie[(size_t)(j0 + jj) * 2 * lmax + rest];
    }
    __syncthreads();

    // phase 2: each lane owns one output column of the tile, walks the rows
    const int W = jcount * bs;
    const size_t row0 = (size_t)i * bs;
    const size_t col0 = (size_t)j0 * bs;
    for (int c = threadIdx.x; c < W; c += kThreads) {
        const int jj = c / bs;
        const int cb = c - jj * bs;
        const int b = cb >= rmax ? 1 : 0;
        const int s = cb - b * rmax;
        const int j = j0 + jj;
        OutT* col = out + row0 * n + col0 + c;
        float2* col32 = DUAL ? out32 + row0 * n + col0 + c : nullptr;
        if (j == i) {
            for (int l = 0; l < bs; l++) {
                double2 v = make_double2((add_identity && l == cb) ? 1.0 : 0.0, 0.0);
                store<OutT>(col + (size_t)l * n, v);
                if (DUAL) store<float2>(col32 + (size_t)l * n, v);
            }
            continue;
        }
        const int v = (int)sqrtf((float)(s + 1) + 0.5f);
        const double2 mj = mieJ[jj][b][v - 1];
        for (int l = 0; l < bs; l++) {
            const int a = l >= rmax ? 1 : 0;
            const int r = l - a * rmax;
            const int t = r * rmax + s;
            const int4 meta = tmeta[t];
            const int amu = meta.z < 0 ? -meta.z : meta.z;
            const double2* cf = tcoef + (size_t)t * coef_slots;
            double2 acc = make_double2(0.0, 0.0);
            if (((a + b) & 1) == 0) {
                for (int q = 0; q < meta.x; q++) {
                    const int p = meta.w - 2 * q;
                    const double2 hl = cscale(L.hp[jj][p], L.leg[jj][leg_index(p, amu)]);
                    cfma(acc, cf[q], hl);
                }
            } else {
                for (int q = 1; q <= meta.y; q++) {
                    const int p = meta.w - 2 * q + 1;
                    const double2 hl = cscale(L.hp[jj][p], L.leg[jj][leg_index(p, amu)]);
                    cfma(acc, cf[qa_slots + q - 1], hl);
                }
            }
            double2 e = L.eph[jj][amu];
            if (meta.z < 0) e.y = -e.y;
            const double2 val = cmul(cmul(e, acc), mj);
            store<OutT>(col + (size_t)l * n, val);
            if (DUAL) store<float2>(col32 + (size_t)l * n, val);
        }
    }
}

// Column blocks per tile: the J <= 16 that keeps the most lanes busy in
// phase 2 (W = J * 2 rmax columns over 256 lanes); ties go to the larger J.
int pick_tile(int lmax) {
    const int bs = 2 * lmax_to_rmax(lmax);
    int best = 1;
    double best_eff = -1;
    for (int J = 1; J <= kJMax; J++) {
        const int W = J * bs;
        const int iters = (W + kThreads - 1) / kThreads;
        const double eff = (double)W / (iters * kThreads);
        if (eff >= best_eff - 1e-12) {
            best = J;
            best_eff = eff;
        }
    }
    return best;
}

}  // namespace

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
    const TransTable* T = mpa_get_table(lmax);
    if (!T) return MPA_ERR_ARG;
    const int J = pick_tile(lmax);
    const int ntiles = (N + J - 1) / J;
    const long long nblocks = (long long)N * ntiles;
    if (nblocks > 0x7fffffffLL) {
        mpa_set_error("sphere_aggregate_tmatrix: N=%d too large for one launch", N);
        return MPA_ERR_ARG;
    }
    hipStream_t s = (hipStream_t)stream;
    ProfScope prof(out_dtype == MPA_C128 ? 2 : 3, s);
    if (out_dtype == MPA_C128)
        hipLaunchKernelGGL((assemble_kernel<double2, false>), dim3((unsigned)nblocks), dim3(kThreads), 0, s, N, lmax, J,
                           ntiles, positions, (const double2*)mie, k, T->meta, T->coef, add_identity, (double2*)out,
                           (float2*)nullptr);
    else
        hipLaunchKernelGGL((assemble_kernel<float2, false>), dim3((unsigned)nblocks), dim3(kThreads), 0, s, N, lmax, J,
                           ntiles, positions, (const double2*)mie, k, T->meta, T->coef, add_identity, (float2*)out,
                           (float2*)nullptr);
    MPA_LAUNCH_CHECK();
    return MPA_OK;
}

extern "C" int mpa_assembly_tile(int lmax) { return pick_tile(lmax); }

extern "C" int mpa_sphere_aggregate_tmatrix_dual(int N, int lmax, const double* positions, const double* mie, double k,
                                                 int add_identity, void* out64, void* out32, void* stream) {
    if (N < 1 || !positions || !mie || !out64 || !out32 || !(k > 0)) {
        mpa_set_error("sphere_aggregate_tmatrix_dual: bad argument (N=%d, k=%g)", N, k);
        return MPA_ERR_ARG;
    }
    const TransTable* T = mpa_get_table(lmax);
    if (!T) return MPA_ERR_ARG;
    const int J = pick_tile(lmax);
    const int ntiles = (N + J - 1) / J;
    const long long nblocks = (long long)N * ntiles;
    if (nblocks > 0x7fffffffLL) {
        mpa_set_error("sphere_aggregate_tmatrix_dual: N=%d too large for one launch", N);
        return MPA_ERR_ARG;
    }
    hipStream_t s = (hipStream_t)stream;
    ProfScope prof(2, s);
    hipLaunchKernelGGL((assemble_kernel<double2, true>), dim3((unsigned)nblocks), dim3(kThreads), 0, s, N, lmax, J,
                       ntiles, positions, (const double2*)mie, k, T->meta, T->coef, add_identity, (double2*)out64,
                       (float2*)out32);
    MPA_LAUNCH_CHECK();
    return MPA_OK;
}
