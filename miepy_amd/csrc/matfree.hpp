// Matrix-free sphere-cluster operator shared by the standalone matvec and the
// batched Krylov engine.
#pragma once

#include <vector>

#include "krylov.hpp"

namespace mpa {

// One y = (I +) T x job, T applied from positions and Mie coefficients.
struct MfJob {
    int N, lmax, S, add_identity;
    const double* pos;   // [N, 3]
    const double2* mie;  // [N, 2, lmax]
    double k;
    const double2* x;    // [S, n]
    double2* y;          // [S, n]
    double2* part;       // [chunks, S, n], then [S, n] for z = Tm x (general T-matrices)
    int skip;            // SysDev index whose all-done flag skips the job, or -1
    int pad;
    const double2* tmat;  // [N, 2 rmax, 2 rmax] per-particle T-matrices, or null (spheres: mie)
    const double2* zpre;  // z = Tm x when tmat is set (filled by matfree_launch), else null
    int c64;              // 1: x, y are complex64 and the product runs on the FP32 matrix cores
    int pad2;
};

// Whether the complex64 matrix-core operator covers (N, lmax, T-matrices).
bool matfree_c64_supported(int N, int lmax, bool tmatrix);

// Workspace bytes of a job (the per-chunk partial rows and the z buffer).
size_t matfree_workspace_bytes(int N, int lmax, int S);

// Launches the two kernels of one job; `sys` may be null when skip < 0.
int matfree_launch(const MfJob& job, const SysDev* sys, hipStream_t st);

// Launches every job: complex64 sphere jobs share launches (grouped by lmax)
// when the shared kernel is selected, the others run one by one.
int matfree_launch_many(const std::vector<MfJob>& jobs, const SysDev* sys, hipStream_t st);

}  // namespace mpa
