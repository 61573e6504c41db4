// Shared device/host helpers for the MI355X cluster-solve library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <mutex>

#include "../../include/miepy_amd.h"

// ---- complex arithmetic on HIP vector types ---------------------------------
// c128 = double2 (re, im), c64 = float2 (re, im); the interleaved layout is
// numpy's complex128 / complex64.

template <typename T> struct cplx;
template <> struct cplx<double> { using type = double2; };
template <> struct cplx<float> { using type = float2; };
template <typename T> using cplx_t = typename cplx<T>::type;

template <typename T> __host__ __device__ inline cplx_t<T> cmake(T re, T im) {
    cplx_t<T> z;
    z.x = re;
    z.y = im;
    return z;
}
__host__ __device__ inline double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ inline float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// conj(a) * b
__host__ __device__ inline double2 cmulc(double2 a, double2 b) {
    return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__host__ __device__ inline float2 cmulc(float2 a, float2 b) {
    return make_float2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__host__ __device__ inline double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__host__ __device__ inline float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__host__ __device__ inline double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__host__ __device__ inline float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__host__ __device__ inline double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
__host__ __device__ inline float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__host__ __device__ inline double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
__host__ __device__ inline float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
// acc += a * b
__host__ __device__ inline void cfma(double2& acc, double2 a, double2 b) {
    acc.x = fma(a.x, b.x, acc.x);
    acc.x = fma(-a.y, b.y, acc.x);
    acc.y = fma(a.x, b.y, acc.y);
    acc.y = fma(a.y, b.x, acc.y);
}
__host__ __device__ inline void cfma(float2& acc, float2 a, float2 b) {
    acc.x = fmaf(a.x, b.x, acc.x);
    acc.x = fmaf(-a.y, b.y, acc.x);
    acc.y = fmaf(a.x, b.y, acc.y);
    acc.y = fmaf(a.y, b.x, acc.y);
}
__host__ __device__ inline double2 cdiv(double2 a, double2 b) {
    double d = b.x * b.x + b.y * b.y;
    return make_double2((a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d);
}

// ---- mode-index conventions (src/miepy/vsh/mode_indices.py:4-16) -----------
__host__ __device__ inline int lmax_to_rmax(int lmax) { return lmax * (lmax + 2); }
// r(n, m) = n(n+2) - n + m - 1
__host__ __device__ inline int mode_r(int n, int m) { return n * (n + 2) - n + m - 1; }
// the degree n of mode index r: n^2 <= r + 1 < (n + 1)^2
__host__ __device__ inline int mode_n(int r) {
    int n = (int)sqrt((double)(r + 1));
    while (n * n > r + 1) n--;
    while ((n + 1) * (n + 1) <= r + 1) n++;
    return n;
}

// Largest lmax the kernels are compiled for (tables, per-pair scratch).
#define MPA_LMAX_MAX 6
#define MPA_PMAX_MAX (2 * MPA_LMAX_MAX + 1)

// ---- translation coefficient table (built on the host, per lmax) ----------
// One record per ordered mode tuple t = r(n,m) * rmax + r(v,u):
//   meta = {number of A terms (q = 0..qA), number of B terms (q = 1..qB),
//           mu = u - m (the Legendre order / azimuthal index), n + v}
//   coef[0 .. QA_SLOTS)            A terms  f * A(q) * legendre sign factor
//   coef[QA_SLOTS .. COEF_SLOTS)   B terms -f * B(q) * legendre sign factor
// so that, for a pair at distance k*d and direction (theta, phi),
//   X_A = e^{i mu phi} sum_q coefA[q] P_{p}^{|mu|}(cos theta) h_p(kd),
//   X_B = e^{i mu phi} sum_q coefB[q] P_{p+1}^{|mu|}(cos theta) h_{p+1}(kd),
// with p = n + v - 2q (the q-th A term) or p = n + v - 2(q+1) (B terms).
struct TransTable {
    int lmax = 0;
    int rmax = 0;
    int qa_slots = 0;    // lmax + 1
    int coef_slots = 0;  // 2 lmax + 1
    int4* meta = nullptr;      // device, [rmax*rmax]
    double2* coef = nullptr;   // device, [rmax*rmax*coef_slots]
    // compact form (only the nonzero terms; translation.hpp TabDev):
    int nterms = 0;
    int2* cmeta = nullptr;            // device, [rmax*rmax]
    double2* ccoef = nullptr;         // device, [nterms]
    unsigned short* clidx = nullptr;  // device, [nterms]
    // contraction form (factorized matrix-free operator, matfree.hip):
    int nentries = 0;
    int2* wmeta = nullptr;     // device, [nentries] {first term, count | kind << 16}
    int* wsrc = nullptr;       // device, column mode s of each term
    double2* wcoef = nullptr;  // device, coefficient of each term
    int nh = 0;                // HLE values per pair, (2 lmax + 1)^2
    int2* hmap = nullptr;      // device, [rmax][nh]: the A and B entry of (r, h), or -1
    // the same terms flattened per (r, h): hlist[r * nh + h] = (first, count)
    // into hterm_src (sidx | kind << 8) and hterm_coef, A terms before B terms
    int nhterms = 0;
    int2* hlist = nullptr;
    int* hterm_src = nullptr;
    double2* hterm_coef = nullptr;
};

// Returns the (cached, per device) table for lmax; nullptr on failure.
const TransTable* mpa_get_table(int lmax);

// ---- optional launch timing (profile.cpp) -----------------------------------
// kinds: 0 matvec c128, 1 matvec c64, 2 assembly c128, 3 assembly c64,
//        4 matrix-free matvec c128, 5 matrix-free matvec c64
class ProfScope {
   public:
    ProfScope(int kind, hipStream_t stream);
    ~ProfScope();

   private:
    int kind_;
    hipStream_t stream_;
    void* start_;
};

// ---- runtime settings and resources (runtime.cpp) ------------------------------
int mpa_option(int key);
bool mpa_profile_active();
// Non-blocking stream of the current device on which iterations are captured.
hipStream_t mpa_capture_stream();
// A second non-blocking stream of the current device (default priority) with
// a fork and a join event, for work that runs beside the caller's stream.
struct SideStream {
    hipStream_t stream = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
};
// Exclusive use of the device's side stream and its events for the lifetime
// of the lease (the events are re-recorded by every user: the record and the
// waits on it must not interleave with another caller's); get() is null if
// they could not be created.
class SideLease {
   public:
    SideLease();
    const SideStream* get() const { return side_; }

   private:
    std::lock_guard<std::mutex> lock_;
    SideStream* side_;
};
// Exclusive use of the process's pinned host staging buffer (>= bytes) for
// the lifetime of the lease; ptr() is null if the allocation failed.
class PinnedLease {
   public:
    explicit PinnedLease(size_t bytes);
    void* ptr() const { return ptr_; }

   private:
    std::lock_guard<std::mutex> lock_;
    void* ptr_;
};

// ---- error plumbing ------------------------------------------------------------
void mpa_set_error(const char* fmt, ...);

#define MPA_HIP_CHECK(expr)                                                         \
    do {                                                                            \
        hipError_t _e = (expr);                                                     \
        if (_e != hipSuccess) {                                                     \
            mpa_set_error("%s:%d: %s: %s", __FILE__, __LINE__, #expr, hipGetErrorString(_e)); \
            return MPA_ERR_HIP;                                                     \
        }                                                                           \
    } while (0)

#define MPA_LAUNCH_CHECK()                                                          \
    do {                                                                            \
        hipError_t _e = hipGetLastError();                                          \
        if (_e != hipSuccess) {                                                     \
            mpa_set_error("%s:%d: kernel launch: %s", __FILE__, __LINE__, hipGetErrorString(_e)); \
            return MPA_ERR_HIP;                                                     \
        }                                                                           \
    } while (0)

// Legendre table layout: P_p^mu (mu >= 0) at index p(p+1)/2 + mu.
__host__ __device__ inline int leg_index(int p, int mu) { return p * (p + 1) / 2 + mu; }
