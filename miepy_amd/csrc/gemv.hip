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
