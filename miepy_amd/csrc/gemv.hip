// Dense complex matvec / multi-RHS product: kernels, tile variants, the
// autotuner and the C ABI.  Replaces the OpenMP row-block GEMV of
// cpp/src/interactions.cpp:62-95 (and the products inside Eigen's BiCGSTAB).
//
// The system matrix is streamed from HBM exactly once per pass with 16-byte
// nontemporal loads (the matrix is not reused within a call, so it is kept
// out of the caches that hold x).  Vector-ALU tiles: each wave of a
// workgroup owns R consecutive rows of one system and runs along the full
// rows with U chunks per lane in flight; the x chunk loaded for a column is
// reused for the R rows and for up to SG right-hand sides held in registers.
// Jobs whose system has converged for all right-hand sides return before
// touching the matrix.

#include <vector>

#include "gemv.hpp"

namespace mpa {
namespace {

typedef double dvec2 __attribute__((ext_vector_type(2)));
typedef float fvec4 __attribute__((ext_vector_type(4)));

template <typename T> struct Chunk;
template <> struct Chunk<double> {
    using v = dvec2;
    static constexpr int C = 1;  // complex entries per 16-byte chunk
};
template <> struct Chunk<float> {
    using v = fvec4;
    static constexpr int C = 2;
};

__device__ inline void chunk_fma(double& ar, double& ai, dvec2 a, dvec2 x) {
    ar = fma(a.x, x.x, ar);
    ar = fma(-a.y, x.y, ar);
    ai = fma(a.x, x.y, ai);
    ai = fma(a.y, x.x, ai);
}
__device__ inline void chunk_fma(float& ar, float& ai, fvec4 a, fvec4 x) {
    ar = fmaf(a.x, x.x, ar);
    ar = fmaf(-a.y, x.y, ar);
    ai = fmaf(a.x, x.y, ai);
    ai = fmaf(a.y, x.x, ai);
    ar = fmaf(a.z, x.z, ar);
    ar = fmaf(-a.w, x.w, ar);
    ai = fmaf(a.z, x.w, ai);
    ai = fmaf(a.w, x.z, ai);
}

__device__ inline bool all_done(const SysDev& d) {
    for (int s = 0; s < d.S; s++)
        if (!d.scal[s].done) return false;
    return true;
}

// One workgroup's share of a matvec job: its four waves own R rows each,
// starting at row 4 R rb; SG right-hand sides per pass, U chunks in flight.
template <typename T, int R, int SG, int U>
__device__ inline void gemv_rows(const MvDev& J, int rb) {
    using V = typename Chunk<T>::v;
    constexpr int C = Chunk<T>::C;
    const int n = J.n;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int row0 = (rb * 4 + w) * R;
    if (row0 >= n) return;
    const int nv = n / C;  // 16-byte chunks per row (n is even)
    const T* A = (const T*)J.A;
    const V* Ar[R];
#pragma unroll
    for (int r = 0; r < R; r++) Ar[r] = (const V*)(A + (size_t)2 * n * min(row0 + r, n - 1));
    for (int s0 = 0; s0 < J.S; s0 += SG) {
        T ar[R][SG], ai[R][SG];
#pragma unroll
        for (int r = 0; r < R; r++)
#pragma unroll
            for (int s = 0; s < SG; s++) ar[r][s] = ai[r][s] = 0;
        const V* xs[SG];
#pragma unroll
        for (int s = 0; s < SG; s++) xs[s] = (const V*)((const T*)J.x + (size_t)2 * n * min(s0 + s, J.S - 1));
        int c = lane;
        for (; c + 64 * (U - 1) < nv; c += 64 * U) {
            V xv[SG][U];
#pragma unroll
            for (int s = 0; s < SG; s++)
#pragma unroll
                for (int u = 0; u < U; u++) xv[s][u] = xs[s][c + 64 * u];
            V av[R][U];
#pragma unroll
            for (int r = 0; r < R; r++)
#pragma unroll
                for (int u = 0; u < U; u++) av[r][u] = __builtin_nontemporal_load(Ar[r] + c + 64 * u);
#pragma unroll
            for (int r = 0; r < R; r++)
#pragma unroll
                for (int s = 0; s < SG; s++)
#pragma unroll
                    for (int u = 0; u < U; u++) chunk_fma(ar[r][s], ai[r][s], av[r][u], xv[s][u]);
        }
        for (; c < nv; c += 64) {
            V xv[SG];
#pragma unroll
            for (int s = 0; s < SG; s++) xv[s] = xs[s][c];
#pragma unroll
            for (int r = 0; r < R; r++) {
                const V a = __builtin_nontemporal_load(Ar[r] + c);
#pragma unroll
                for (int s = 0; s < SG; s++) chunk_fma(ar[r][s], ai[r][s], a, xv[s]);
            }
        }
#pragma unroll
        for (int r = 0; r < R; r++)
#pragma unroll
            for (int s = 0; s < SG; s++) {
                T re = ar[r][s], im = ai[r][s];
                for (int off = 32; off > 0; off >>= 1) {
                    re += __shfl_xor(re, off);
                    im += __shfl_xor(im, off);
                }
                const int row = row0 + r;
                if (lane == 0 && row < n && s0 + s < J.S) {
                    const size_t o = (size_t)(s0 + s) * n + row;
                    cplx_t<T> yv = cmake<T>(re, im);
                    if (J.add_identity) yv = cadd(yv, ((const cplx_t<T>*)J.x)[o]);
                    ((cplx_t<T>*)J.y)[o] = yv;
                }
            }
    }
}

// ---- multi-RHS products on the matrix cores ------------------------------------
//
// Y = (I +) A X for up to kGemmS right-hand sides per pass as a real GEMM:
// a complex row of A is the real row (Re a_0, Im a_0, Re a_1, ...), and
// right-hand side s contributes two real columns (one per output part),
//   X'[2j][(s,re)] = Re x_j,  X'[2j][(s,im)] = Im x_j,
//   X'[2j+1][(s,re)] = -Im x_j, X'[2j+1][(s,im)] = Re x_j,
// so that Y_re = sum Re a Re x - Im a Im x and Y_im = sum Re a Im x + Im a Re x.
// A is streamed once per pass with 16-byte nontemporal loads straight into
// the MFMA operand registers (each lane reads a contiguous run of its row);
// the X' block of a K step is staged in LDS once per workgroup and shared by
// its four waves, whose lanes form their B operands from it.
//   c64:  v_mfma_f32_32x32x2f32, FP32 accumulate; a wave owns 32 rows x
//         32 real columns (16 RHS); K step 32 complex columns.
//   c128: v_mfma_f64_16x16x4f64; a wave owns 16 rows x 2 tiles of 16 real
//         columns (16 RHS); K step 32 complex columns.
// Lane layouts (checked on the device): 32x32x2 A lane l -> (row l%32,
// k l/32), B -> (col l%32, k l/32), D[i] -> (row 8(i/4) + 4(l/32) + i%4,
// col l%32); f64 16x16x4 A lane l -> (row l%16, k l/16), B -> (col l%16,
// k l/16), D[i] -> (row 4 i + l/16, col l%16).

typedef float f16v __attribute__((ext_vector_type(16)));
typedef double d4v __attribute__((ext_vector_type(4)));

constexpr int kGemmS = 16;   // right-hand sides per pass
constexpr int kGemmKB = 32;  // complex columns per K step
template <typename T> struct GemmRows;
template <> struct GemmRows<float> { static constexpr int value = 128; };   // 4 waves x 32
template <> struct GemmRows<double> { static constexpr int value = 64; };   // 4 waves x 16

// element of X' for K parity `odd` and output part `part` from x = (re, im)
template <typename R>
__device__ inline R xprime(R re, R im, int odd, int part) {
    return odd ? (part ? re : -im) : (part ? im : re);
}

// A tile of a K step in LDS: rows x (32 complex + 1 pad chunk); the odd
// 16-byte pad spreads the rows of one MFMA operand read over the banks.
constexpr int kPadC64 = 2;  // floats of padding per row (c64: 64 + 4 floats -> stride 68 dwords)

__device__ inline void gemm_rows_c64(const MvDev& J, int rb) {
    constexpr int RW = 128, RS = 2 * kGemmKB + 2 * kPadC64;  // rows per workgroup, row stride (floats)
    __shared__ float As[RW * RS];
    __shared__ float2 xs[kGemmS][kGemmKB + 1];
    const int n = J.n;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int h = lane >> 5, col = lane & 31, s_of = col >> 1, part = col & 1;
    const float2* A = (const float2*)J.A;
    const float2* X = (const float2*)J.x;
    const int nkb = (n + kGemmKB - 1) / kGemmKB;
    // global -> register staging of one K step: thread loads 8 chunks of 16 B
    // (chunk id = it * 256 + tid: 16 chunks per row, a wave covers 4 full rows)
    for (int s0 = 0; s0 < J.S; s0 += kGemmS) {
        auto load_a = [&](int kb, fvec4 (&ra)[8]) {
#pragma unroll
            for (int it = 0; it < 8; it++) {
                const int id = it * kThreads + tid;
                const int r = id >> 4, c16 = id & 15;
                const int row = min(rb * RW + r, n - 1);
                const int j = kb * kGemmKB + 2 * c16;  // first complex column of the chunk (even; n even)
                fvec4 z = {0.f, 0.f, 0.f, 0.f};
                ra[it] = j < n ? __builtin_nontemporal_load((const fvec4*)(A + (size_t)row * n + j)) : z;
            }
        };
        auto load_x = [&](int kb, float2 (&rx)[2]) {
#pragma unroll
            for (int it = 0; it < 2; it++) {
                const int u = it * kThreads + tid;
                const int s = u / kGemmKB, jj = u % kGemmKB, j = kb * kGemmKB + jj;
                rx[it] = (s0 + s < J.S && j < n) ? X[(size_t)(s0 + s) * n + j] : make_float2(0.f, 0.f);
            }
        };
        auto store_lds = [&](const fvec4 (&ra)[8], const float2 (&rx)[2]) {
#pragma unroll
            for (int it = 0; it < 8; it++) {
                const int id = it * kThreads + tid;
                *(fvec4*)(As + (id >> 4) * RS + (id & 15) * 4) = ra[it];
            }
#pragma unroll
            for (int it = 0; it < 2; it++) {
                const int u = it * kThreads + tid;
                xs[u / kGemmKB][u % kGemmKB] = rx[it];
            }
        };
        f16v acc;
#pragma unroll
        for (int i = 0; i < 16; i++) acc[i] = 0.f;
        fvec4 ra[8];
        float2 rx[2];
        load_a(0, ra);
        load_x(0, rx);
        for (int kb = 0; kb < nkb; kb++) {
            __syncthreads();  // the previous step's reads of As / xs are done
            store_lds(ra, rx);
            __syncthreads();
            if (kb + 1 < nkb) {
                load_a(kb + 1, ra);
                load_x(kb + 1, rx);
            }
            const float* arow = As + (w * 32 + (lane & 31)) * RS + h * 32;
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const fvec4 av = *(const fvec4*)(arow + 4 * q);
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    const int t = 4 * q + e;
                    const float2 xv = xs[s_of][16 * h + (t >> 1)];
                    acc = __builtin_amdgcn_mfma_f32_32x32x2f32(av[e], xprime<float>(xv.x, xv.y, t & 1, part), acc, 0, 0,
                                                               0);
                }
            }
        }
        const int s = s0 + s_of;
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const int r = rb * RW + w * 32 + 8 * (i >> 2) + 4 * h + (i & 3);
            if (r < n && s < J.S && s_of < kGemmS) {
                const size_t o = 2 * ((size_t)s * n + r) + part;
                float v = acc[i];
                if (J.add_identity) v += ((const float*)J.x)[o];
                ((float*)J.y)[o] = v;
            }
        }
    }
}

// Eight right-hand sides per pass (16 real columns, one 16x16x4 tile wide):
// v_mfma_f32_16x16x4f32, a wave owns 32 rows as two 16-row tiles sharing
// each B operand; half the matrix-core work per byte of A of the 16-RHS
// form, whose upper 16 columns would be idle at S <= 8.  Each lane reads a
// float4 of its row (real K 16q + 4g .. +3, g = lane / 16) and feeds its
// four components to four MFMAs, so MFMA (q, e) covers real K 16q + 4g + e
// over g; the B operands follow the same order.  D[i] of 16x16x4f32 ->
// (row 4 (l/16) + i, col l%16).
constexpr int kGemmS8 = 8;

__device__ inline void gemm_rows_c64_s8(const MvDev& J, int rb) {
    constexpr int RW = 128, RS = 2 * kGemmKB + 2 * kPadC64;  // rows per workgroup, row stride (floats)
    __shared__ float As[RW * RS];
    __shared__ fvec4 xs4[kGemmS8][kGemmKB / 2 + 1];  // two complex per float4; one float4 of padding
    const int n = J.n;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int g = lane >> 4, col = lane & 15, s_of = col >> 1, part = col & 1;
    const float2* A = (const float2*)J.A;
    const float2* X = (const float2*)J.x;
    const int nkb = (n + kGemmKB - 1) / kGemmKB;
    for (int s0 = 0; s0 < J.S; s0 += kGemmS8) {
        // A: 128 rows x 16 chunks of 16 B, 8 per thread (a wave covers 4 full rows per it)
        auto load_a = [&](int kb, fvec4 (&ra)[8]) {
#pragma unroll
            for (int it = 0; it < 8; it++) {
                const int id = it * kThreads + tid;
                const int r = id >> 4, c16 = id & 15;
                const int row = min(rb * RW + r, n - 1);
                const int j = kb * kGemmKB + 2 * c16;
                fvec4 z = {0.f, 0.f, 0.f, 0.f};
                ra[it] = j < n ? __builtin_nontemporal_load((const fvec4*)(A + (size_t)row * n + j)) : z;
            }
        };
        // X: 8 right-hand sides x 32 complex, one per thread
        auto load_x = [&](int kb, float2& rx) {
            const int s = tid / kGemmKB, jj = tid % kGemmKB, j = kb * kGemmKB + jj;
            rx = (s0 + s < J.S && j < n) ? X[(size_t)(s0 + s) * n + j] : make_float2(0.f, 0.f);
        };
        auto store_lds = [&](const fvec4 (&ra)[8], const float2& rx) {
#pragma unroll
            for (int it = 0; it < 8; it++) {
                const int id = it * kThreads + tid;
                *(fvec4*)(As + (id >> 4) * RS + (id & 15) * 4) = ra[it];
            }
            ((float2*)&xs4[tid / kGemmKB][0])[tid % kGemmKB] = rx;
        };
        fvec4 acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
        fvec4 ra[8];
        float2 rx;
        load_a(0, ra);
        load_x(0, rx);
        for (int kb = 0; kb < nkb; kb++) {
            __syncthreads();  // the previous step's reads of As / xs4 are done
            store_lds(ra, rx);
            __syncthreads();
            if (kb + 1 < nkb) {
                load_a(kb + 1, ra);
                load_x(kb + 1, rx);
            }
            const float* arow0 = As + (w * 32 + col) * RS + 4 * g;
            const float* arow1 = arow0 + 16 * RS;
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const fvec4 a0 = *(const fvec4*)(arow0 + 16 * q);
                const fvec4 a1 = *(const fvec4*)(arow1 + 16 * q);
                const fvec4 xv = xs4[s_of][4 * q + g];  // complex columns 8q + 2g and 8q + 2g + 1
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    const float xr = (e >> 1) ? xv.z : xv.x, xi = (e >> 1) ? xv.w : xv.y;
                    const float b = xprime<float>(xr, xi, e & 1, part);
                    acc0 = __builtin_amdgcn_mfma_f32_16x16x4f32(a0[e], b, acc0, 0, 0, 0);
                    acc1 = __builtin_amdgcn_mfma_f32_16x16x4f32(a1[e], b, acc1, 0, 0, 0);
                }
            }
        }
        const int s = s0 + s_of;
#pragma unroll
        for (int t = 0; t < 2; t++)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int r = rb * RW + w * 32 + 16 * t + 4 * g + i;
                if (r < n && s < J.S) {
                    const size_t o = 2 * ((size_t)s * n + r) + part;
                    float v = t ? acc1[i] : acc0[i];
                    if (J.add_identity) v += ((const float*)J.x)[o];
                    ((float*)J.y)[o] = v;
                }
            }
    }
}

constexpr int kPadC128 = 2;  // doubles of padding per row (c128: 64 + 2 doubles -> stride 132 dwords)

// NT tiles of 8 right-hand sides per pass (NT = 2: 16 RHS; NT = 1: 8 RHS at
// half the matrix-core work)
template <int NT>
__device__ inline void gemm_rows_c128(const MvDev& J, int rb) {
    constexpr int RW = 64, RS = 2 * kGemmKB + kPadC128;  // rows per workgroup, row stride (doubles)
    __shared__ double As[RW * RS];
    __shared__ double2 xs[kGemmS][kGemmKB + 1];
    const int n = J.n;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int g = lane >> 4, col = lane & 15, s_of = col >> 1, part = col & 1;
    const double2* A = (const double2*)J.A;
    const double2* X = (const double2*)J.x;
    const int nkb = (n + kGemmKB - 1) / kGemmKB;
    // 64 rows x 32 complex = 2048 chunks of 16 B (one complex each): thread loads 8
    // (chunk id = it * 256 + tid: 32 chunks per row, a wave covers 2 full rows)
    for (int s0 = 0; s0 < J.S; s0 += 8 * NT) {
        auto load_a = [&](int kb, dvec2 (&ra)[8]) {
#pragma unroll
            for (int it = 0; it < 8; it++) {
                const int id = it * kThreads + tid;
                const int r = id >> 5, c = id & 31;
                const int row = min(rb * RW + r, n - 1);
                const int j = kb * kGemmKB + c;
                dvec2 z = {0.0, 0.0};
                ra[it] = j < n ? __builtin_nontemporal_load((const dvec2*)(A + (size_t)row * n + j)) : z;
            }
        };
        auto load_x = [&](int kb, double2 (&rx)[2]) {
#pragma unroll
            for (int it = 0; it < 2; it++) {
                const int u = it * kThreads + tid;
                const int s = u / kGemmKB, jj = u % kGemmKB, j = kb * kGemmKB + jj;
                rx[it] = (s0 + s < J.S && j < n) ? X[(size_t)(s0 + s) * n + j] : make_double2(0.0, 0.0);
            }
        };
        auto store_lds = [&](const dvec2 (&ra)[8], const double2 (&rx)[2]) {
#pragma unroll
            for (int it = 0; it < 8; it++) {
                const int id = it * kThreads + tid;
                *(dvec2*)(As + (id >> 5) * RS + (id & 31) * 2) = ra[it];
            }
#pragma unroll
            for (int it = 0; it < 2; it++) {
                const int u = it * kThreads + tid;
                xs[u / kGemmKB][u % kGemmKB] = rx[it];
            }
        };
        d4v acc0, acc1;
#pragma unroll
        for (int i = 0; i < 4; i++) acc0[i] = acc1[i] = 0.0;
        dvec2 ra[8];
        double2 rx[2];
        load_a(0, ra);
        load_x(0, rx);
        for (int kb = 0; kb < nkb; kb++) {
            __syncthreads();
            store_lds(ra, rx);
            __syncthreads();
            if (kb + 1 < nkb) {
                load_a(kb + 1, ra);
                load_x(kb + 1, rx);
            }
            const double* arow = As + (w * 16 + (lane & 15)) * RS + g * 16;
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const dvec2 av = *(const dvec2*)(arow + 2 * q);
#pragma unroll
                for (int e = 0; e < 2; e++) {
                    const int t = 2 * q + e;
                    const double2 x0 = xs[s_of][8 * g + (t >> 1)];
                    acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(av[e], xprime<double>(x0.x, x0.y, t & 1, part), acc0, 0,
                                                                0, 0);
                    if (NT == 2) {
                        const double2 x1 = xs[8 + s_of][8 * g + (t >> 1)];
                        acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(av[e], xprime<double>(x1.x, x1.y, t & 1, part),
                                                                    acc1, 0, 0, 0);
                    }
                }
            }
        }
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int r = rb * RW + w * 16 + 4 * i + g;
#pragma unroll
            for (int c = 0; c < NT; c++) {
                const int s = s0 + 8 * c + s_of;
                if (r < n && s < J.S) {
                    const size_t o = 2 * ((size_t)s * n + r) + part;
                    double v = c ? acc1[i] : acc0[i];
                    if (J.add_identity) v += ((const double*)J.x)[o];
                    ((double*)J.y)[o] = v;
                }
            }
        }
    }
}

// SP right-hand sides per pass: 16 or 8
template <typename T, int SP>
__device__ inline void gemm_rows(const MvDev& J, int rb);
template <>
__device__ inline void gemm_rows<float, 16>(const MvDev& J, int rb) {
    gemm_rows_c64(J, rb);
}
template <>
__device__ inline void gemm_rows<float, 8>(const MvDev& J, int rb) {
    gemm_rows_c64_s8(J, rb);
}
template <>
__device__ inline void gemm_rows<double, 16>(const MvDev& J, int rb) {
    gemm_rows_c128<2>(J, rb);
}
template <>
__device__ inline void gemm_rows<double, 8>(const MvDev& J, int rb) {
    gemm_rows_c128<1>(J, rb);
}

template <typename T, int SP>
__global__ __launch_bounds__(kThreads) void k_gemm(const MvDev* __restrict__ jobs, const SysDev* __restrict__ sys,
                                                   const int2* __restrict__ work) {
    const int2 wk = work[blockIdx.x];
    const MvDev J = jobs[wk.x];
    if (J.skip >= 0 && all_done(sys[J.skip])) return;
    gemm_rows<T, SP>(J, wk.y);
}

template <typename T, int SP>
__global__ __launch_bounds__(kThreads) void k_gemm1(MvDev J) {
    gemm_rows<T, SP>(J, blockIdx.x);
}

// batched: one workgroup per (job, row block) of the work list
template <typename T, int RB, int SG, int U>
__global__ __launch_bounds__(kThreads) void k_gemv(const MvDev* __restrict__ jobs, const SysDev* __restrict__ sys,
                                                   const int2* __restrict__ work) {
    const int2 wk = work[blockIdx.x];
    const MvDev J = jobs[wk.x];
    if (J.skip >= 0 && all_done(sys[J.skip])) return;
    gemv_rows<T, RB, SG, U>(J, wk.y);
}

// one job passed by value: workgroup b owns row block b
template <typename T, int RB, int SG, int U>
__global__ __launch_bounds__(kThreads) void k_gemv1(MvDev J) {
    gemv_rows<T, RB, SG, U>(J, blockIdx.x);
}


// ---- tile variants ------------------------------------------------------------------

struct Variant {
    int R, SG, U;  // vector-ALU tile; R = 0 marks the matrix-core GEMM
    const char* name;
};
constexpr int kMaxVariants = 5;
const Variant kVariants[kTileClasses][kMaxVariants] = {
    {{8, 1, 2, "r8u2"}, {4, 1, 2, "r4u2"}, {8, 1, 4, "r8u4"}, {16, 1, 1, "r16u1"}, {4, 1, 4, "r4u4"}},
    {{4, 2, 2, "r4u2"}, {2, 2, 2, "r2u2"}, {4, 2, 4, "r4u4"}, {8, 2, 1, "r8u1"}, {0, 0, 0, nullptr}},
    {{2, 4, 2, "r2u2"}, {1, 4, 2, "r1u2"}, {2, 4, 4, "r2u4"}, {0, 16, 0, "mfma"}, {0, 0, 0, nullptr}},
    {{0, 16, 0, "mfma"}, {1, 8, 2, "r1s8u2"}, {2, 8, 1, "r2s8u1"}, {0, 8, 0, "mfma8"}, {0, 0, 0, nullptr}},
};
int g_variant[2][kTileClasses] = {{0, 0, 0, 0}, {0, 0, 0, 0}};  // [es == 8][class]

inline int nvariants(int c) {
    int k = 0;
    while (k < kMaxVariants && kVariants[c][k].name) k++;
    return k;
}
inline const Variant& variant_of(int c, int es) { return kVariants[c][g_variant[es == 8][c]]; }

template <typename T>
int launch_variant(int c, int v, const MvDev* jobs, const SysDev* sys, const int2* work, int nwork, const MvDev* one,
                   hipStream_t st) {
    const dim3 grid(nwork), block(kThreads);
#define MPA_VAR(R_, SG_, U_)                                                                          \
    do {                                                                                              \
        if (one) hipLaunchKernelGGL((k_gemv1<T, R_, SG_, U_>), grid, block, 0, st, *one);             \
        else hipLaunchKernelGGL((k_gemv<T, R_, SG_, U_>), grid, block, 0, st, jobs, sys, work);      \
    } while (0)
    const Variant& V = kVariants[c][v];
    if (V.R == 0 && V.SG == 8) {
        if (one) hipLaunchKernelGGL((k_gemm1<T, 8>), grid, block, 0, st, *one);
        else hipLaunchKernelGGL((k_gemm<T, 8>), grid, block, 0, st, jobs, sys, work);
    } else if (V.R == 0) {
        if (one) hipLaunchKernelGGL((k_gemm1<T, 16>), grid, block, 0, st, *one);
        else hipLaunchKernelGGL((k_gemm<T, 16>), grid, block, 0, st, jobs, sys, work);
    } else if (c == 0) {
        switch (v) {
            case 0: MPA_VAR(8, 1, 2); break;
            case 1: MPA_VAR(4, 1, 2); break;
            case 2: MPA_VAR(8, 1, 4); break;
            case 3: MPA_VAR(16, 1, 1); break;
            default: MPA_VAR(4, 1, 4); break;
        }
    } else if (c == 1) {
        switch (v) {
            case 0: MPA_VAR(4, 2, 2); break;
            case 1: MPA_VAR(2, 2, 2); break;
            case 2: MPA_VAR(4, 2, 4); break;
            default: MPA_VAR(8, 2, 1); break;
        }
    } else if (c == 2) {
        switch (v) {
            case 0: MPA_VAR(2, 4, 2); break;
            case 1: MPA_VAR(1, 4, 2); break;
            default: MPA_VAR(2, 4, 4); break;
        }
    } else {
        switch (v) {
            case 1: MPA_VAR(1, 8, 2); break;
            default: MPA_VAR(2, 8, 1); break;
        }
    }
#undef MPA_VAR
    MPA_LAUNCH_CHECK();
    return MPA_OK;
}

int rows_of_variant(int c, int v, int es) {
    const Variant& V = kVariants[c][v];
    if (V.R == 0) return es == 8 ? GemmRows<double>::value : GemmRows<float>::value;
    return 4 * V.R;
}

}  // namespace

int gemv_class(int S) { return S == 1 ? 0 : (S == 2 ? 1 : (S <= 4 ? 2 : 3)); }

int gemv_rows_of_class(int c, int es) { return rows_of_variant(c, g_variant[es == 8][c], es); }

int gemv_row_blocks(int n, int S, int es) {
    const int rb = gemv_rows_of_class(gemv_class(S), es);
    return (n + rb - 1) / rb;
}

int gemv_launch(int es, const MvDev* jobs, const SysDev* sys, int2* const* work, const int* nwork, hipStream_t st) {
    ProfScope prof(es == 8 ? 0 : 1, st);
    for (int c = 0; c < kTileClasses; c++) {
        if (!nwork[c]) continue;
        const int v = g_variant[es == 8][c];
        const int rc = es == 8 ? launch_variant<double>(c, v, jobs, sys, work[c], nwork[c], nullptr, st)
                               : launch_variant<float>(c, v, jobs, sys, work[c], nwork[c], nullptr, st);
        if (rc) return rc;
    }
    return MPA_OK;
}

int gemv_launch1(int es, const MvDev& J, hipStream_t st) {
    ProfScope prof(es == 8 ? 0 : 1, st);
    const int c = gemv_class(J.S);
    const int v = g_variant[es == 8][c];
    const int nb = (J.n + rows_of_variant(c, v, es) - 1) / rows_of_variant(c, v, es);
    return es == 8 ? launch_variant<double>(c, v, nullptr, nullptr, nullptr, nb, &J, st)
                   : launch_variant<float>(c, v, nullptr, nullptr, nullptr, nb, &J, st);
}

}  // namespace mpa

using namespace mpa;

extern "C" int mpa_matvec(int n, int nrhs, int dtype, const void* A, int add_identity, const void* x, void* y,
                          void* stream) {
    if (n < 1 || nrhs < 1 || !A || !x || !y || (n & 1)) {
        mpa_set_error("matvec: bad argument (n must be even and positive)");
        return MPA_ERR_ARG;
    }
    if (dtype != MPA_C128 && dtype != MPA_C64) {
        mpa_set_error("matvec: unknown dtype %d", dtype);
        return MPA_ERR_ARG;
    }
    MvDev J{A, x, y, n, nrhs, add_identity, -1};
    return gemv_launch1(dtype == MPA_C128 ? 8 : 4, J, (hipStream_t)stream);
}

// Autotuner: time every tile variant of the class of `nrhs` on the caller's
// operands (y is overwritten), keep the fastest for (dtype, class).
extern "C" int mpa_autotune_matvec(int n, int nrhs, int dtype, const void* A, const void* x, void* y, int reps,
                                   int* variant_out, double* ms_out, void* stream) {
    if (n < 2 || (n & 1) || nrhs < 1 || !A || !x || !y || reps < 1 || (dtype != MPA_C128 && dtype != MPA_C64)) {
        mpa_set_error("autotune_matvec: bad argument");
        return MPA_ERR_ARG;
    }
    const int es = dtype == MPA_C128 ? 8 : 4;
    const int c = gemv_class(nrhs);
    hipStream_t st = (hipStream_t)stream;
    MvDev J{A, x, y, n, nrhs, 0, -1};
    hipEvent_t e0, e1;
    MPA_HIP_CHECK(hipEventCreate(&e0));
    MPA_HIP_CHECK(hipEventCreate(&e1));
    int best = 0;
    double best_ms = 1e30;
    const int saved = g_variant[es == 8][c];
    int rc = MPA_OK;
    for (int v = 0; v < nvariants(c) && rc == MPA_OK; v++) {
        g_variant[es == 8][c] = v;
        rc = gemv_launch1(es, J, st);  // warm-up
        if (rc) break;
        (void)hipEventRecord(e0, st);
        for (int r = 0; r < reps && rc == MPA_OK; r++) rc = gemv_launch1(es, J, st);
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
    g_variant[es == 8][c] = rc == MPA_OK ? best : saved;
    if (rc) return rc;
    if (variant_out) *variant_out = best;
    if (ms_out) *ms_out = best_ms;
    return MPA_OK;
}

// Batched autotuner: every variant of the class of nrhs[0] timed on the
// batched launch over all systems of that class (the launch shape of the
// solvers: one workgroup per (system, row block)), the fastest kept.
extern "C" int mpa_autotune_matvec_batched(int count, const int* n, const int* nrhs, int dtype, const void* const* A,
                                           const void* const* x, void* const* y, int reps, int* variant_out,
                                           double* ms_out, void* stream) {
    if (count < 1 || !n || !nrhs || !A || !x || !y || reps < 1 || (dtype != MPA_C128 && dtype != MPA_C64)) {
        mpa_set_error("autotune_matvec_batched: bad argument");
        return MPA_ERR_ARG;
    }
    const int es = dtype == MPA_C128 ? 8 : 4;
    const int c = gemv_class(nrhs[0] < 1 ? 1 : nrhs[0]);
    std::vector<MvDev> jobs;
    for (int i = 0; i < count; i++) {
        if (n[i] < 2 || (n[i] & 1) || nrhs[i] < 1 || !A[i] || !x[i] || !y[i]) {
            mpa_set_error("autotune_matvec_batched: bad system %d", i);
            return MPA_ERR_ARG;
        }
        if (gemv_class(nrhs[i]) == c) jobs.push_back(MvDev{A[i], x[i], y[i], n[i], nrhs[i], 0, -1});
    }
    const int nv = nvariants(c);
    size_t max_work = 0;
    for (int v = 0; v < nv; v++) {
        size_t w = 0;
        for (const MvDev& J : jobs) w += (J.n + rows_of_variant(c, v, es) - 1) / rows_of_variant(c, v, es);
        max_work = w > max_work ? w : max_work;
    }
    hipStream_t st = (hipStream_t)stream;
    MvDev* d_jobs = nullptr;
    int2* d_work = nullptr;
    MPA_HIP_CHECK(hipMalloc(&d_jobs, sizeof(MvDev) * jobs.size()));
    if (hipMalloc(&d_work, sizeof(int2) * max_work) != hipSuccess) {
        (void)hipFree(d_jobs);
        mpa_set_error("autotune_matvec_batched: out of device memory");
        return MPA_ERR_OOM;
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = MPA_OK;
    if (hipMemcpyAsync(d_jobs, jobs.data(), sizeof(MvDev) * jobs.size(), hipMemcpyHostToDevice, st) != hipSuccess ||
        hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
        rc = MPA_ERR_HIP;
    const int saved = g_variant[es == 8][c];
    int best = saved;
    double best_ms = 1e30;
    std::vector<int2> work;
    for (int v = 0; v < nv && rc == MPA_OK; v++) {
        const int rows = rows_of_variant(c, v, es);
        work.clear();
        for (int j = 0; j < (int)jobs.size(); j++)
            for (int b = 0; b < (jobs[j].n + rows - 1) / rows; b++) work.push_back(make_int2(j, b));
        order_work_longest_first(work, [&](int j) { return jobs[j].n; });  // as the solvers order it
        if (hipMemcpyAsync(d_work, work.data(), sizeof(int2) * work.size(), hipMemcpyHostToDevice, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess) {
            rc = MPA_ERR_HIP;
            break;
        }
        g_variant[es == 8][c] = v;
        int2* wl[kTileClasses] = {nullptr, nullptr, nullptr, nullptr};
        int nw[kTileClasses] = {0, 0, 0, 0};
        wl[c] = d_work;
        nw[c] = (int)work.size();
        rc = gemv_launch(es, d_jobs, nullptr, wl, nw, st);  // warm-up
        if (rc) break;
        (void)hipEventRecord(e0, st);
        for (int r = 0; r < reps && rc == MPA_OK; r++) rc = gemv_launch(es, d_jobs, nullptr, wl, nw, st);
        (void)hipEventRecord(e1, st);
        if (hipEventSynchronize(e1) != hipSuccess) rc = MPA_ERR_HIP;
        float ms = 0;
        if (rc == MPA_OK && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = MPA_ERR_HIP;
        if (rc == MPA_OK && ms / reps < best_ms) {
            best_ms = ms / reps;
            best = v;
        }
    }
    (void)hipStreamSynchronize(st);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(d_work);
    (void)hipFree(d_jobs);
    g_variant[es == 8][c] = rc == MPA_OK ? best : saved;
    if (rc) return rc;
    if (variant_out) *variant_out = best;
    if (ms_out) *ms_out = best_ms;
    return MPA_OK;
}

extern "C" int mpa_matvec_variant(int dtype, int nrhs, int set_to) {
    if (dtype != MPA_C128 && dtype != MPA_C64) return -1;
    const int c = gemv_class(nrhs < 1 ? 1 : nrhs);
    int& v = g_variant[dtype == MPA_C128][c];
    if (set_to >= 0 && set_to < nvariants(c)) v = set_to;
    return v;
}

extern "C" const char* mpa_matvec_variant_name(int dtype, int nrhs, int variant) {
    const int c = gemv_class(nrhs < 1 ? 1 : nrhs);
    if (variant < 0 || variant >= nvariants(c) || (dtype != MPA_C128 && dtype != MPA_C64)) return "";
    return kVariants[c][variant].name;
}

// Plain 16-byte copy of `bytes` (a multiple of 16): the reference point for
// scaling hardware byte counters (FETCH_SIZE / WRITE_SIZE) to HBM bytes.
__global__ __launch_bounds__(kThreads) void k_copy16(size_t nchunks, const dvec2* __restrict__ src,
                                                     dvec2* __restrict__ dst) {
    for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < nchunks; i += (size_t)gridDim.x * kThreads)
        __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}

extern "C" int mpa_stream_copy(const void* src, void* dst, size_t bytes, void* stream) {
    if (!src || !dst || bytes % 16) {
        mpa_set_error("stream_copy: bad argument");
        return MPA_ERR_ARG;
    }
    hipLaunchKernelGGL(k_copy16, dim3(8192), dim3(kThreads), 0, (hipStream_t)stream, bytes / 16, (const dvec2*)src,
                       (dvec2*)dst);
    MPA_LAUNCH_CHECK();
    return MPA_OK;
}
