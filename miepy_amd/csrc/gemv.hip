// C ABI of the dense complex matvec (see gemv.hpp for the kernel design).
// Replaces the OpenMP row-block GEMV of cpp/src/interactions.cpp:62-95.

#include "gemv.hpp"

using namespace mpa;

extern "C" int mpa_matvec(int n, int nrhs, int dtype, const void* A, int add_identity, const void* x, void* y,
                          void* stream) {
    if (n < 1 || nrhs < 1 || !A || !x || !y || (n & 1)) {
        mpa_set_error("matvec: bad argument (n must be even and positive)");
        return MPA_ERR_ARG;
    }
    MvDev J{A, x, y, n, nrhs, add_identity, -1};
    if (dtype == MPA_C128) return gemv_launch1<double>(J, (hipStream_t)stream);
    if (dtype == MPA_C64) return gemv_launch1<float>(J, (hipStream_t)stream);
    mpa_set_error("matvec: unknown dtype %d", dtype);
    return MPA_ERR_ARG;
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
