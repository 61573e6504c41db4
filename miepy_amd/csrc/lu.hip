// Dense LU with partial pivoting: the `exact` method of solve_linear_system.
//
// Replaces interaction_matrix.partialPivLu().solve(p_src)
// (cpp/src/interactions.cpp:356-357, and :338-339 on the preconditioned path).
// Right-looking blocked LU with panels of kNb = 32 columns:
//
//   for each panel [k0, k0 + 32):
//     for each column j of the panel:
//       pivot p = the first row i >= j of largest |A[i][j]| (partial
//       pivoting as Eigen's PartialPivLU; the modulus is compared squared);
//       swap the whole rows j and p; A[i][j] /= A[j][j] for i > j; rank-1
//       update of the panel columns right of j (fused with the search for
//       the next column's pivot);
//     U12 = L11^-1 A12 (unit lower triangular, L11 in LDS, one thread per
//     column);
//     A22 -= L21 U12 on the FP64 matrix cores (v_mfma_f64_16x16x4f64),
//     64 x 64 complex tiles, K = 32 complex = 64 real positions.
//
// L (unit diagonal) and U overwrite A; the row swaps, in order, become one
// permutation at the end of the factorization.  The solve gathers b through
// that permutation and runs blocked forward and back substitution (a
// one-workgroup solve of each 32 x 32 diagonal block, then a block-column
// update of the rows below / above).  Everything is FP64.

#include <algorithm>
#include <climits>
#include <cstdio>
#include <vector>

#include "common.hpp"



namespace {

constexpr int kNb = 32;         // panel width = K of the trailing update
constexpr int kT = 256;         // threads per workgroup
constexpr int kMaxParts = 240;  // workgroups of the pivot searches (partial maxima)
constexpr int kTile = 64;       // trailing-update tile (complex rows and columns)

size_t al256(size_t b) { return (b + 255) & ~size_t(255); }

struct LuWs {
    int* piv;     // [n] pivot row of each column
    int* perm;    // [n] row permutation: (P b)[k] = b[perm[k]]
    double* pval; // [kMaxParts] partial maxima of |A[i][j]|^2
    int* pidx;    // [kMaxParts] their rows
    int* info;    // 0, or 1 + the first column without a nonzero pivot
};

size_t ws_bytes(int n) {
    return 2 * al256(sizeof(int) * (size_t)n) + al256(sizeof(double) * kMaxParts) + al256(sizeof(int) * kMaxParts) +
           256;
}

LuWs ws_view(void* w, int n) {
    char* p = (char*)w;
    LuWs v;
    v.piv = (int*)p; p += al256(sizeof(int) * (size_t)n);
    v.perm = (int*)p; p += al256(sizeof(int) * (size_t)n);
    v.pval = (double*)p; p += al256(sizeof(double) * kMaxParts);
    v.pidx = (int*)p; p += al256(sizeof(int) * kMaxParts);
    v.info = (int*)p;
    return v;
}

// (value, row) with the larger value winning and the lower row on ties: the
// first maximum, as a sequential scan finds it.
__device__ inline void better(double& v, int& i, double v2, int i2) {
    if (v2 > v || (v2 == v && i2 < i)) {
        v = v2;
        i = i2;
    }
}

// Block-wide reduction of (v, i) to thread 0 (all threads must call).
__device__ inline void block_best(double& v, int& i) {
    __shared__ double sv[kT / 64];
    __shared__ int si[kT / 64];
    for (int off = 32; off > 0; off >>= 1) {
        const double v2 = __shfl_xor(v, off);
        const int i2 = __shfl_xor(i, off);
        better(v, i, v2, i2);
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) {
        sv[w] = v;
        si[w] = i;
    }
    __syncthreads();
    if (threadIdx.x == 0)
        for (int k = 1; k < kT / 64; k++) better(v, i, sv[k], si[k]);
    __syncthreads();
}

__global__ __launch_bounds__(kT) void k_lu_add_identity(double2* __restrict__ A, int n) {
    const int i = blockIdx.x * kT + threadIdx.x;
    if (i < n) A[(size_t)i * n + i].x += 1.0;
}

// Partial maxima of |A[i][j]|^2 over rows i >= j (the first column of a panel).
__global__ __launch_bounds__(kT) void k_lu_pivot_search(const double2* __restrict__ A, int n, int j,
                                                        double* __restrict__ pval, int* __restrict__ pidx) {
    double v = -1.0;
    int idx = INT_MAX;
    for (int i = j + blockIdx.x * kT + threadIdx.x; i < n; i += gridDim.x * kT) {
        const double2 a = A[(size_t)i * n + j];
        better(v, idx, a.x * a.x + a.y * a.y, i);
    }
    block_best(v, idx);
    if (threadIdx.x == 0) {
        pval[blockIdx.x] = v;
        pidx[blockIdx.x] = idx;
    }
}

// Reduce the partial maxima to column j's pivot row p (every workgroup
// redundantly), record it, and swap the whole rows j and p.  A column
// without a nonzero (finite) pivot records info and keeps p = j.
__global__ __launch_bounds__(kT) void k_lu_swap(double2* __restrict__ A, int n, int j, const double* __restrict__ pval,
                                                const int* __restrict__ pidx, int nparts, int* __restrict__ piv,
                                                int* __restrict__ info) {
    __shared__ int sp;
    double v = -1.0;
    int idx = INT_MAX;
    if ((int)threadIdx.x < nparts) {
        v = pval[threadIdx.x];
        idx = pidx[threadIdx.x];
    }
    block_best(v, idx);
    if (threadIdx.x == 0) {
        const bool ok = v > 0.0 && v <= 1.7976931348623157e308 && idx < n;
        sp = ok ? idx : j;
        if (blockIdx.x == 0) {
            piv[j] = sp;
            if (!ok && *info == 0) *info = j + 1;
        }
    }
    __syncthreads();
    const int p = sp;
    if (p == j) return;
    double2* rj = A + (size_t)j * n;
    double2* rp = A + (size_t)p * n;
    for (int c = blockIdx.x * kT + threadIdx.x; c < n; c += gridDim.x * kT) {
        const double2 t = rj[c];
        rj[c] = rp[c];
        rp[c] = t;
    }
}

// Column j of the panel [.., k_end): A[i][j] /= A[j][j] for i > j, then
// A[i][c] -= A[i][j] A[j][c] for the panel columns j < c < k_end, and the
// partial maxima of the updated column j + 1 (when it is in the panel).
__global__ __launch_bounds__(kT) void k_lu_panel(double2* __restrict__ A, int n, int j, int k_end,
                                                 double* __restrict__ pval, int* __restrict__ pidx) {
    __shared__ double2 urow[kNb];
    __shared__ double2 dpiv;
    const int w = k_end - j - 1;  // 0 .. kNb - 1
    if ((int)threadIdx.x < w) urow[threadIdx.x] = A[(size_t)j * n + j + 1 + threadIdx.x];
    if (threadIdx.x == 0) dpiv = A[(size_t)j * n + j];
    __syncthreads();
    const double2 d = dpiv;
    double v = -1.0;
    int idx = INT_MAX;
    for (int i = j + 1 + blockIdx.x * kT + threadIdx.x; i < n; i += gridDim.x * kT) {
        double2* row = A + (size_t)i * n;
        const double2 l = cdiv(row[j], d);
        row[j] = l;
        for (int c = 0; c < w; c++) row[j + 1 + c] = csub(row[j + 1 + c], cmul(l, urow[c]));
        if (w > 0) {
            const double2 a = row[j + 1];
            better(v, idx, a.x * a.x + a.y * a.y, i);
        }
    }
    if (w > 0) {
        block_best(v, idx);
        if (threadIdx.x == 0) {
            pval[blockIdx.x] = v;
            pidx[blockIdx.x] = idx;
        }
    }
}

// U12 = L11^-1 A12: rows [k0, k0 + kNb), columns [k0 + kNb, n).
__global__ __launch_bounds__(kT) void k_lu_trsm(double2* __restrict__ A, int n, int k0) {
    __shared__ double2 L[kNb][kNb + 1];
    for (int u = threadIdx.x; u < kNb * kNb; u += kT) {
        const int r = u / kNb, q = u - r * kNb;
        L[r][q] = A[(size_t)(k0 + r) * n + k0 + q];
    }
    __syncthreads();
    const int c = k0 + kNb + blockIdx.x * kT + threadIdx.x;
    if (c >= n) return;
    double2 x[kNb];
#pragma unroll
    for (int r = 0; r < kNb; r++) x[r] = A[(size_t)(k0 + r) * n + c];
#pragma unroll
    for (int r = 1; r < kNb; r++) {
        double2 s = x[r];
#pragma unroll
        for (int q = 0; q < r; q++) s = csub(s, cmul(L[r][q], x[q]));
        x[r] = s;
    }
#pragma unroll
    for (int r = 1; r < kNb; r++) A[(size_t)(k0 + r) * n + c] = x[r];
}

typedef double d4v __attribute__((ext_vector_type(4)));

// B operand of the complex product as a real one: K position 2q + odd holds
// Re (odd = 0) or Im (odd = 1) of L[.][q]; real output column 2c + part is
// Re (part 0) or Im (part 1) of the product, so B'[2q][2c] = Re U, B'[2q+1][2c]
// = -Im U, B'[2q][2c+1] = Im U, B'[2q+1][2c+1] = Re U.
__device__ inline double bprime(double2 u, int odd, int part) {
    return odd ? (part ? u.x : -u.y) : (part ? u.y : u.x);
}

// A22 -= L21 U12 for rows and columns >= k0 + kNb: one workgroup per 64 x 64
// complex tile, each wave 16 rows x 64 complex columns (8 MFMA tiles of 16
// real columns), K = 64 real positions in 16 steps.  The MFMA operand and
// result layouts: A[row = lane & 15][k = lane >> 4], B[k = lane >> 4][col =
// lane & 15], D[row = 4 q + (lane >> 4)][col = lane & 15] for q = 0..3.
__global__ __launch_bounds__(kT) void k_lu_gemm(double2* __restrict__ A, int n, int k0) {
    __shared__ double2 Ls[kTile][kNb + 1];
    __shared__ double2 Us[kNb][kTile + 1];
    const int r0 = k0 + kNb;
    const int tr = r0 + blockIdx.y * kTile, tc = r0 + blockIdx.x * kTile;
    for (int u = threadIdx.x; u < kTile * kNb; u += kT) {
        const int r = u / kNb, q = u - r * kNb;
        const int i = tr + r;
        Ls[r][q] = i < n ? A[(size_t)i * n + k0 + q] : make_double2(0.0, 0.0);
    }
    for (int u = threadIdx.x; u < kNb * kTile; u += kT) {
        const int q = u / kTile, cc = u - q * kTile;
        const int c = tc + cc;
        Us[q][cc] = c < n ? A[(size_t)(k0 + q) * n + c] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int g = lane >> 4, col = lane & 15, part = col & 1;
    double* Ad = (double*)A;
    d4v acc[8];
#pragma unroll
    for (int t = 0; t < 8; t++)
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int i = tr + w * 16 + 4 * q + g, c = tc + t * 8 + (col >> 1);
            acc[t][q] = (i < n && c < n) ? Ad[2 * ((size_t)i * n + c) + part] : 0.0;
        }
    const int odd = g & 1;
#pragma unroll
    for (int s = 0; s < 2 * kNb / 4; s++) {
        const int q = 2 * s + (g >> 1);
        const double2 lv = Ls[w * 16 + col][q];
        const double a = odd ? -lv.y : -lv.x;
#pragma unroll
        for (int t = 0; t < 8; t++)
            acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bprime(Us[q][t * 8 + (col >> 1)], odd, part), acc[t], 0,
                                                          0, 0);
    }
#pragma unroll
    for (int t = 0; t < 8; t++)
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int i = tr + w * 16 + 4 * q + g, c = tc + t * 8 + (col >> 1);
            if (i < n && c < n) Ad[2 * ((size_t)i * n + c) + part] = acc[t][q];
        }
}

// x[s][k] = b[s][perm[k]]
__global__ __launch_bounds__(kT) void k_lu_permute(const double2* __restrict__ b, double2* __restrict__ x,
                                                   const int* __restrict__ perm, int n) {
    const int k = blockIdx.x * kT + threadIdx.x;
    const size_t o = (size_t)blockIdx.y * n;
    if (k < n) x[o + k] = b[o + perm[k]];
}

// Diagonal block [b0, b0 + nb) of L (unit lower, lower = 1) or U (upper,
// with its diagonal): one wave per right-hand side (grid.x), lane q holds
// unknown q; each row is a wave-wide sum of the solved unknowns.
__global__ __launch_bounds__(64) void k_lu_diag_solve(const double2* __restrict__ LU, int n, int b0, int nb,
                                                      double2* __restrict__ x, int lower) {
    __shared__ double2 B[kNb][kNb + 1];
    for (int u = threadIdx.x; u < nb * nb; u += 64) {
        const int r = u / nb, q = u - r * nb;
        B[r][q] = LU[(size_t)(b0 + r) * n + b0 + q];
    }
    __syncthreads();
    const int q = threadIdx.x;
    double2* xs = x + (size_t)blockIdx.x * n + b0;
    double2 y = q < nb ? xs[q] : make_double2(0.0, 0.0);
    for (int k = 0; k < nb; k++) {
        const int r = lower ? k : nb - 1 - k;
        const bool term = q < nb && (lower ? q < r : q > r);
        double2 t = term ? cmul(B[r][q], y) : make_double2(0.0, 0.0);
        for (int off = 32; off > 0; off >>= 1) {
            t.x += __shfl_xor(t.x, off);
            t.y += __shfl_xor(t.y, off);
        }
        if (q == r) y = lower ? csub(y, t) : cdiv(csub(y, t), B[r][r]);
    }
    if (q < nb) xs[q] = y;
}

// x[s][i] -= sum_q LU[i][b0 + q] x[s][b0 + q] for the rows i in [i0, i1)
// (below the block for L, above it for U); grid.y = right-hand sides.
__global__ __launch_bounds__(kT) void k_lu_block_update(const double2* __restrict__ LU, int n, int b0, int nb, int i0,
                                                        int i1, double2* __restrict__ x) {
    __shared__ double2 xb[kNb];
    double2* xs = x + (size_t)blockIdx.y * n;
    if ((int)threadIdx.x < nb) xb[threadIdx.x] = xs[b0 + threadIdx.x];
    __syncthreads();
    const int i = i0 + blockIdx.x * kT + threadIdx.x;
    if (i >= i1) return;
    const double2* row = LU + (size_t)i * n + b0;
    double2 t = xs[i];
    for (int q = 0; q < nb; q++) t = csub(t, cmul(row[q], xb[q]));
    xs[i] = t;
}

int parts_for(int rows) {
    int g = (rows + kT - 1) / kT;
    return g < 1 ? 1 : (g > kMaxParts ? kMaxParts : g);
}

int lu_factor(int n, double2* A, int add_identity, void* workspace, int* info_out, hipStream_t st) {
    const LuWs W = ws_view(workspace, n);
    MPA_HIP_CHECK(hipMemsetAsync(W.info, 0, sizeof(int), st));
    if (add_identity) hipLaunchKernelGGL(k_lu_add_identity, dim3((n + kT - 1) / kT), dim3(kT), 0, st, A, n);
    const int swap_grid = std::min(64, (n + kT - 1) / kT);
    for (int k0 = 0; k0 < n; k0 += kNb) {
        const int k_end = std::min(n, k0 + kNb);
        int nparts = parts_for(n - k0);
        hipLaunchKernelGGL(k_lu_pivot_search, dim3(nparts), dim3(kT), 0, st, (const double2*)A, n, k0, W.pval, W.pidx);
        for (int j = k0; j < k_end; j++) {
            hipLaunchKernelGGL(k_lu_swap, dim3(swap_grid), dim3(kT), 0, st, A, n, j, (const double*)W.pval,
                               (const int*)W.pidx, nparts, W.piv, W.info);
            const int g = parts_for(n - j - 1);
            hipLaunchKernelGGL(k_lu_panel, dim3(g), dim3(kT), 0, st, A, n, j, k_end, W.pval, W.pidx);
            nparts = g;  // the partial maxima of column j + 1 (when j + 1 < k_end)
        }
        if (k_end < n) {
            const int m = n - k_end;
            hipLaunchKernelGGL(k_lu_trsm, dim3((m + kT - 1) / kT), dim3(kT), 0, st, A, n, k0);
            const int tiles = (m + kTile - 1) / kTile;
            hipLaunchKernelGGL(k_lu_gemm, dim3(tiles, tiles), dim3(kT), 0, st, A, n, k0);
        }
        MPA_LAUNCH_CHECK();
    }
    // the swaps, in order, as one permutation of the right-hand side's rows
    std::vector<int> piv(n), perm(n);
    int info = 0;
    MPA_HIP_CHECK(hipMemcpyAsync(piv.data(), W.piv, sizeof(int) * n, hipMemcpyDeviceToHost, st));
    MPA_HIP_CHECK(hipMemcpyAsync(&info, W.info, sizeof(int), hipMemcpyDeviceToHost, st));
    MPA_HIP_CHECK(hipStreamSynchronize(st));
    for (int k = 0; k < n; k++) perm[k] = k;
    for (int j = 0; j < n; j++) std::swap(perm[j], perm[piv[j]]);
    MPA_HIP_CHECK(hipMemcpyAsync(W.perm, perm.data(), sizeof(int) * n, hipMemcpyHostToDevice, st));
    MPA_HIP_CHECK(hipStreamSynchronize(st));
    if (info_out) *info_out = info;
    if (info) {
        char msg[96];
        std::snprintf(msg, sizeof msg, "lu_factor: singular matrix (no nonzero pivot in column %d)", info - 1);
        mpa_set_error(msg);
        return MPA_BREAKDOWN;
    }
    return MPA_OK;
}

int lu_solve(int n, int S, const double2* LU, void* workspace, const double2* b, double2* x, hipStream_t st) {
    const LuWs W = ws_view(workspace, n);
    hipLaunchKernelGGL(k_lu_permute, dim3((n + kT - 1) / kT, S), dim3(kT), 0, st, b, x, (const int*)W.perm, n);
    for (int b0 = 0; b0 < n; b0 += kNb) {  // L y = P b
        const int nb = std::min(kNb, n - b0);
        hipLaunchKernelGGL(k_lu_diag_solve, dim3(S), dim3(64), 0, st, LU, n, b0, nb, x, 1);
        const int i0 = b0 + nb;
        if (i0 < n)
            hipLaunchKernelGGL(k_lu_block_update, dim3((n - i0 + kT - 1) / kT, S), dim3(kT), 0, st, LU, n, b0, nb,
                               i0, n, x);
    }
    for (int b0 = ((n - 1) / kNb) * kNb; b0 >= 0; b0 -= kNb) {  // U x = y
        const int nb = std::min(kNb, n - b0);
        hipLaunchKernelGGL(k_lu_diag_solve, dim3(S), dim3(64), 0, st, LU, n, b0, nb, x, 0);
        if (b0 > 0)
            hipLaunchKernelGGL(k_lu_block_update, dim3((b0 + kT - 1) / kT, S), dim3(kT), 0, st, LU, n, b0, nb, 0, b0,
                               x);
    }
    MPA_LAUNCH_CHECK();
    return MPA_OK;
}

}  // namespace

extern "C" {

size_t mpa_lu_workspace_bytes(int n) { return n < 1 ? 0 : ws_bytes(n); }

int mpa_lu_factor(int n, void* A, int add_identity, void* workspace, int* info_out, void* stream) {
    if (n < 1 || !A || !workspace) {
        mpa_set_error("lu_factor: bad argument");
        return MPA_ERR_ARG;
    }
    return lu_factor(n, (double2*)A, add_identity, workspace, info_out, (hipStream_t)stream);
}

int mpa_lu_solve(int n, int nrhs, const void* LU, void* workspace, const void* b, void* x, void* stream) {
    if (n < 1 || nrhs < 1 || nrhs > 65535 || !LU || !workspace || !b || !x || b == x) {
        mpa_set_error("lu_solve: bad argument (b and x must be distinct buffers)");
        return MPA_ERR_ARG;
    }
    return lu_solve(n, nrhs, (const double2*)LU, workspace, (const double2*)b, (double2*)x, (hipStream_t)stream);
}

}  // extern "C"
