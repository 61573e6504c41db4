/* C ABI of the MI355X cluster-solve library (libmiepy_amd.so, gfx950).
 *
 * Every pointer argument named like a matrix or vector is a DEVICE pointer
 * (HBM) unless documented as host; `stream` is a hipStream_t passed as void*
 * (NULL = default stream).  Complex arrays are interleaved (re, im): c128 =
 * two doubles, c64 = two floats, i.e. numpy complex128 / complex64.  Matrices
 * are row-major [n, n] with n = 2 * rmax * N, rmax = lmax (lmax + 2), and the
 * flat index particle * 2 rmax + pol * rmax + r(n, m), r(n, m) = n(n+2)-n+m-1
 * (the reference's ComplexMatrix layout, cpp/src/vec.hpp:14,
 * src/miepy/vsh/mode_indices.py:16).  Several right-hand sides are stored
 * RHS-major: [nrhs, n].
 *
 * System sizes n are even (n = 2 rmax N for every cluster system); the
 * 16-byte vector loads of the matvec rely on it.
 *
 * Every function returns an int status (MPA_OK = 0).  On failure
 * mpa_last_error() describes the cause.  Nothing here allocates device
 * memory except the per-lmax translation tables, built once per process and
 * device on first use (a few hundred KB at most).
 */
#ifndef MIEPY_AMD_H
#define MIEPY_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes (the reference returns silently at maxiter and never checks
 * for breakdown or non-finite values, cpp/src/interactions.cpp:138-139,365) */
#define MPA_OK 0
#define MPA_ERR_ARG 1
#define MPA_ERR_HIP 2
#define MPA_NOT_CONVERGED 3
#define MPA_BREAKDOWN 4
#define MPA_NONFINITE 5
#define MPA_ERR_OOM 6

/* element types */
#define MPA_C128 0
#define MPA_C64 1

/* runtime options (mpa_set_option / mpa_get_option) */
#define MPA_OPT_GRAPHS 0 /* 1 (default): the solver loops replay their iterations from a captured
                            HIP graph (one launch per iteration instead of one per kernel); 0: eager
                            launches.  Results are identical either way.  Graphs are not used while
                            launch timing (mpa_profile_enable) is on. */
#define MPA_OPT_MATFREE_FP64_KERNEL 1 /* kernel of the FP64 matrix-free product: 0 (default) the FP64
                                         matrix-core kernels (lmax <= 5), 1 the LDS kernel (always used
                                         for lmax > 5), 2 the factorized register-resident kernel
                                         (lmax <= 3, N >= 128).  Results agree to rounding. */
#define MPA_OPT_MATFREE_C64_KERNEL 2 /* kernel of the complex64 matrix-free product: 0 (default) FP16
                                        matrix cores on hi/lo split operands (three products, FP32
                                        accumulation, per-pair and per-column scaling; relative error
                                        per term about 2^-21), 1 FP32 matrix cores. */
#define MPA_OPT_OVERLAP 3 /* 1 (default): in a product phase of the batched solvers that has both
                             stored-matrix systems and matrix-free systems, the batched GEMM over the
                             stored matrices runs on a second stream of the device beside the
                             matrix-free products (HBM-bound and matrix-core-bound work side by
                             side), joined before the phase ends; 0: one after the other on the
                             caller's stream.  Results are identical either way.  Off while launch
                             timing is on (mpa_bicgstab_noalloc solves one stored system, so it never
                             forks). */
#define MPA_OPT_COUNT 4

const char* mpa_version(void);
const char* mpa_last_error(void);

/* Aggregate T-matrix of a sphere cluster.
 * Replaces interactions.sphere_aggregate_tmatrix(positions, mie, k)
 * (cpp/src/interactions_pyb.cpp:26-31 -> cpp/src/interactions.cpp:370-415).
 *   positions  [N, 3] f64 (device)
 *   mie        [N, 2, lmax] c128 (device): mie(j, b*lmax + v-1), as the
 *              reference's mie argument
 *   k          medium wavenumber
 *   out_dtype  MPA_C128 or MPA_C64 (FP64 arithmetic, rounded on store)
 *   add_identity  0: T exactly as the reference; 1: I + T (the system matrix
 *              of cpp/src/interactions.cpp:351-353)
 *   out        [n, n] of out_dtype (device); every entry is written. */
int mpa_sphere_aggregate_tmatrix(int N, int lmax, const double* positions, const double* mie, double k, int out_dtype,
                                 int add_identity, void* out, void* stream);

/* The same matrix written twice in one pass: out64 [n, n] c128 and its
 * complex64 rounding out32 [n, n] (device), for the mixed-precision solve,
 * which needs both; the FP64 arithmetic is done once. */
int mpa_sphere_aggregate_tmatrix_dual(int N, int lmax, const double* positions, const double* mie, double k,
                                      int add_identity, void* out64, void* out32, void* stream);

/* Aggregate T-matrix of a cluster of particles with general T-matrices.
 * Replaces interactions.particle_aggregate_tmatrix(positions, tmatrix, k)
 * (cpp/src/interactions_pyb.cpp:12-24 -> cpp/src/interactions.cpp:417-462,
 * cpp/src/vsh_translation.cpp:53-159): block (i, j) = X^{ij} Tm_j.
 *   tmatrix  [N, 2 rmax, 2 rmax] c128 (device), row-major per particle
 * Other arguments as mpa_sphere_aggregate_tmatrix. */
int mpa_particle_aggregate_tmatrix(int N, int lmax, const double* positions, const double* tmatrix, double k,
                                   int out_dtype, int add_identity, void* out, void* stream);

/* y_j = Tm_j x_j for every particle j and right-hand side: x, y [nrhs, n]
 * c128 (device, distinct), tmatrix as above.  The einsum
 * "naibj,nbj->nai" of src/miepy/cluster.py:650,657 (p_scat = T p_inc). */
int mpa_tmatrix_apply(int N, int lmax, int nrhs, const double* tmatrix, const void* x, void* y, void* stream);

/* y = (add_identity ? x : 0) + A x for nrhs vectors ([nrhs, n]).
 * Replaces the OpenMP row-block GEMV inside Eigen's BiCGSTAB
 * (cpp/src/interactions.cpp:62-95). */
int mpa_matvec(int n, int nrhs, int dtype, const void* A, int add_identity, const void* x, void* y, void* stream);

/* Autotuner for the matvec tiles (no counterpart in the reference).  Times
 * every tile variant of the RHS-count class of `nrhs` (vector-ALU row tiles
 * of several shapes; the matrix-core GEMM for 3+ RHS) on the caller's
 * operands, `reps` launches each (y is overwritten), and makes the fastest
 * the variant used for (dtype, class) by every later matvec and solve in
 * the process.  Outputs (host, may be NULL): the chosen variant and its
 * average launch time in ms. */
int mpa_autotune_matvec(int n, int nrhs, int dtype, const void* A, const void* x, void* y, int reps,
                        int* variant_out, double* ms_out, void* stream);
/* The same over a batch (no counterpart in the reference): every variant of
 * the class of nrhs[0] timed on one batched launch over the `count` systems
 * of that class (n[i], nrhs[i], device pointers A[i], x[i], y[i]; systems of
 * other classes are left out), as the batched solvers launch it.  Allocates
 * its job and work lists for the duration of the call. */
int mpa_autotune_matvec_batched(int count, const int* n, const int* nrhs, int dtype, const void* const* A,
                                const void* const* x, void* const* y, int reps, int* variant_out, double* ms_out,
                                void* stream);
/* Variant in use for (dtype, class of nrhs); set_to >= 0 selects one. */
int mpa_matvec_variant(int dtype, int nrhs, int set_to);
const char* mpa_matvec_variant_name(int dtype, int nrhs, int variant);

/* Bytes of device workspace mpa_bicgstab needs. */
size_t mpa_bicgstab_workspace_bytes(int n, int nrhs, int dtype);

/* BiCGSTAB for (add_identity ? I + A : A) x = b, nrhs systems sharing A.
 * Replaces the bicgstab branch of solve_linear_system
 * (cpp/src/interactions.cpp:347-368; Eigen::BiCGSTAB, identity
 * preconditioner): zero initial guess, stop at ||r|| <= tol ||b|| or maxiter.
 * The reference hard-codes tol = 1e-5, maxiter = 1000.
 *   dtype       MPA_C128 (FP64) or MPA_C64 (FP32 vectors, FP64 scalars)
 *   b, x        [nrhs, n] of dtype (device); x is overwritten
 *   workspace   mpa_bicgstab_workspace_bytes(n, nrhs, dtype) bytes (device)
 *   iters_out, resid_out, status_out  host arrays [nrhs] or NULL:
 *               iterations, ||r||/||b||, per-RHS status
 *   sync_every  iterations enqueued between host convergence checks (>= 1)
 * Returns MPA_OK, or the first non-OK per-RHS status (MPA_NOT_CONVERGED,
 * MPA_NONFINITE). */
int mpa_bicgstab(int n, int nrhs, int dtype, const void* A, int add_identity, const void* b, void* x, double tol,
                 int maxiter, void* workspace, int* iters_out, double* resid_out, int* status_out, int sync_every,
                 void* stream);

/* mpa_bicgstab without a workspace argument: no counterpart in the reference;
 * for callers that cannot allocate device memory (code beside HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The same
 * solve and results as mpa_bicgstab (sync_every = 1), with the workspace in a
 * static 64 MiB device array of this library (one per GPU, held until the
 * process ends): allocates nothing, launches eagerly (no graph), reads the
 * per-RHS scalars into pageable host memory.  Returns MPA_ERR_ARG when
 * mpa_bicgstab_workspace_bytes(n, nrhs, dtype) exceeds 64 MiB (96 bytes
 * per unknown and RHS for MPA_C128: n * nrhs up to about 6.9e5).  Calls on one GPU share the array: calls
 * that overlap on one GPU race. */
int mpa_bicgstab_noalloc(int n, int nrhs, int dtype, const void* A, int add_identity, const void* b, void* x,
                         double tol, int maxiter, int* iters_out, double* resid_out, int* status_out, void* stream);

/* Solver flags (mpa_bicgstab_ex, mpa_system.flags). */
#define MPA_SOLVE_REF_LOOP 1 /* the reference's hand-written loop (cpp/src/interactions.cpp:98-209,
                                230-288): x0 = b, stop at ||r|| < tol (absolute) or after maxiter + 1
                                iterations, no restart on rho breakdown */

/* Bytes of device workspace mpa_bicgstab_ex needs; preconditioned != 0 when
 * minv is passed (two more vectors per RHS). */
size_t mpa_bicgstab_ex_workspace_bytes(int n, int nrhs, int dtype, int preconditioned);

/* mpa_bicgstab with the options of the reference's other BiCGSTAB entry points:
 *   flags       0: Eigen's BiCGSTAB, as mpa_bicgstab; MPA_SOLVE_REF_LOOP: the
 *               reference's own loop, bound as interactions.bicgstab /
 *               bicgstab_profiled (cpp/src/interactions_pyb.cpp:39-55)
 *   minv        NULL, or [n, block_size] of dtype (device): the stacked inverses
 *               of the diagonal blocks, applied as a right preconditioner
 *               (y = M^-1 p, z = M^-1 s, x += alpha y + w z): with flags 0 the
 *               BlockDiagonalPreconditioner solve of
 *               solve_linear_system_preconditioned (interactions.cpp:291-345,
 *               interactions_pyb.cpp:71-87 with M_inv_blocks), with
 *               MPA_SOLVE_REF_LOOP bicgstab_preconditioned_profiled
 *               (interactions.cpp:230-288, interactions_pyb.cpp:57-69)
 *   block_size  rows per diagonal block (divides n); ignored without minv
 *   workspace   mpa_bicgstab_ex_workspace_bytes(n, nrhs, dtype, minv != NULL)
 * resid_out is ||r||/||b|| as for mpa_bicgstab (the reference's loop returns
 * ||r||); MPA_NOT_CONVERGED marks an iteration cap the reference's loop
 * reaches silently. */
int mpa_bicgstab_ex(int n, int nrhs, int dtype, const void* A, int add_identity, const void* b, void* x, double tol,
                    int maxiter, int flags, const void* minv, int block_size, void* workspace, int* iters_out,
                    double* resid_out, int* status_out, int sync_every, void* stream);

/* ---- per-wavelength inputs ------------------------------------------------------ */

/* Mie coefficients of N spheres: mie_scat[i] = (a_n, b_n), mie_int[i] = (c_n,
 * d_n), each [N, 2, lmax] c128 (device).  Replaces the per-particle, per-order
 * calls of mie_sphere_scattering_coefficients / _interior_coefficients
 * (src/miepy/mie_single/mie_sphere.py:16-76, driven by
 * src/miepy/sphere_cluster.py:99-125).  Per particle (device arrays [N]):
 * radius, k (the medium wavenumber) f64; eps, mu, eps_b, mu_b c128;
 * conducting int (nonzero: perfect conductor) or NULL.  1 <= lmax <= 20. */
int mpa_mie_coefficients(int N, int lmax, const double* radius, const void* eps, const void* mu, const double* k,
                         const void* eps_b, const void* mu_b, const int* conducting, void* mie_scat, void* mie_int,
                         void* stream);

/* Plane-wave source coefficients p_src [nsrc, N, 2, rmax] c128 (device) at
 * the particle positions [N, 3] (device): source s is the sum of the waves
 * w in [wave_begin[s], wave_begin[s+1]), p_src[s][i] += e^{i (k r_i . khat_w
 * + phase_w)} table_w; table [waves, 2, rmax] c128 holds each wave's
 * amplitude |E_mn| e^{-i m phi} (tau pol0 - i pi pol1, pi pol0 - i tau pol1)
 * (plane_wave.structure, src/miepy/sources/plane_waves.py:126-156).
 * wave_begin [nsrc + 1] int, khat [waves, 3] and phase [waves] f64 (device). */
int mpa_plane_wave_structure(int N, int lmax, int nsrc, const double* positions, double k, const int* wave_begin,
                             const double* khat, const double* phase, const void* table, void* p_src, void* stream);

/* ---- cluster coefficients ------------------------------------------------------ */

/* Device workspace of mpa_cluster_coefficients (per-chunk partial sums). */
size_t mpa_cluster_coefficients_workspace_bytes(int N, int lmax_out, int nrhs);

/* The cluster's scattered-field coefficients about one origin: every
 * particle's p_scat re-expanded with the incident-mode translation of
 * origin - r_i, summed over particles (a particle exactly at the origin adds
 * its coefficients).  Replaces miepy.cluster_coefficients
 * (src/miepy/vsh/cluster_coefficients.py:14-53), which loops over the scalar
 * vsh_translation (cpp/src/vsh_translation.cpp:428-470) per particle and mode.
 *   positions  [N, 3] f64 (device); origin (ox, oy, oz)
 *   p_scat     [nrhs, N, 2, rmax(lmax_in)] c128 (device)
 *   p_cluster  [nrhs, 2, rmax(lmax_out)] c128 (device), overwritten
 *   1 <= lmax_in, lmax_out <= 6; k the medium wavenumber. */
int mpa_cluster_coefficients(int N, int lmax_in, int lmax_out, int nrhs, const double* positions, double ox,
                             double oy, double oz, const void* p_scat, double k, void* p_cluster, void* workspace,
                             void* stream);

/* The source matrix of a sphere cluster's T-matrix (the S of
 * tmatrix_sphere_cluster, src/miepy/tmatrix/common.py:199-247, built there by
 * loops over the scalar vsh_translation): origin mode (b, n, m), n <= lmax_out,
 * expanded about each particle in modes (a, v, u), v <= lmax_p, with the
 * incident-mode translation of r_i - origin (a particle within 1e-15 of the
 * origin takes the identity).  Stored transposed, one origin mode per row:
 *   out [2 rmax(lmax_out), N, 2, rmax(lmax_p)] c128 (device), overwritten;
 * each row is one right-hand side of (I + T) X = S.  N <= 65535. */
int mpa_cluster_source(int N, int lmax_p, int lmax_out, const double* positions, double ox, double oy, double oz,
                       double k, void* out, void* stream);

/* The cluster T-matrix from the solved cluster system: T = R diag(mie) X,
 * R the re-expansion of every particle's scattered modes about the origin
 * (src/miepy/tmatrix/common.py:250-295; steps 5-6 of tmatrix_sphere_cluster,
 * :352-364).
 *   X        [2 rmax(lmax_out), N, 2, rmax(lmax_p)] c128 (device): the
 *            solutions of (I + T) X = S, one origin mode per row
 *   mie      [N, 2, lmax_p] c128 (device): a_n, b_n
 *   tmatrix  [2, rmax(lmax_out), 2, rmax(lmax_out)] c128 (device),
 *            row-major, overwritten
 *   workspace: mpa_cluster_coefficients_workspace_bytes(N, lmax_out,
 *              2 rmax(lmax_out)) bytes. */
int mpa_cluster_tmatrix(int N, int lmax_p, int lmax_out, const double* positions, double ox, double oy, double oz,
                        const void* mie, const void* X, double k, void* tmatrix, void* workspace, void* stream);

/* Translation coefficients (A, B) of `count` tuples (m, n, u, v) (mnuv
 * [count][4] int, device; 1 <= n, v <= lmax, |m| <= n, |u| <= v) for nvec
 * translation vectors in spherical coordinates rad, theta, phi [nvec] f64
 * (device, rad > 0), wavenumber k and VSH mode 0 outgoing, 1 ingoing,
 * 2 incident, 3 interior: the scalar vsh_translation and its numpy form
 * (cpp/src/vsh_translation.cpp:366-470, bound in
 * cpp/src/vsh_translation_pyb.cpp:14-31).  out [nvec][count][2] c128
 * (device). */
int mpa_vsh_translation(int count, const int* mnuv, int lmax, int nvec, const double* rad, const double* theta,
                        const double* phi, double k, int mode, void* out, void* stream);

/* Cluster cross-sections per multipole (cpp/src/flux.cpp:32-57,
 * src/miepy/flux.py:35-56): p_cluster and p_src (the source's coefficients at
 * the origin) [nrhs, 2, rmax] c128 -> C [nrhs, 3 (scattering, absorption,
 * extinction), 2, lmax] f64 (all device). */
int mpa_cluster_cross_sections(int lmax, int nrhs, const void* p_cluster, const void* p_src, double k, double* C,
                               void* stream);

/* ---- dense LU: the exact method ---------------------------------------------- */

/* Device workspace of one factorization (pivots, row permutation); the same
 * workspace goes to mpa_lu_factor and mpa_lu_solve. */
size_t mpa_lu_workspace_bytes(int n);

/* In-place LU with partial pivoting of (add_identity ? I + A : A), A [n, n]
 * c128 row-major (device): L (unit diagonal) and U overwrite A, the row
 * permutation goes to the workspace.  Replaces the partialPivLu() of the
 * exact branch of solve_linear_system (cpp/src/interactions.cpp:356-357 and
 * :338-339).  Pivot: the first row of largest modulus in the column.
 * *info_out (may be NULL): 0, or 1 + the first column without a nonzero
 * pivot, returned as MPA_BREAKDOWN (Eigen divides by the zero pivot
 * silently).  Synchronizes the stream. */
int mpa_lu_factor(int n, void* A, int add_identity, void* workspace, int* info_out, void* stream);

/* x = A^-1 b from mpa_lu_factor's factors and workspace; b, x [nrhs, n] c128
 * (device, distinct buffers). */
int mpa_lu_solve(int n, int nrhs, const void* LU, void* workspace, const void* b, void* x, void* stream);

/* ---- matrix-free operator ----------------------------------------------------
 * The aggregate interaction operator of a sphere cluster applied without
 * storing it: the translation blocks of interactions.cpp:370-415 /
 * vsh_translation.cpp:161-237 are recomputed inside every product.  Memory
 * is O(n) (positions, Mie coefficients, a workspace of partial rows), so
 * clusters whose n x n matrix does not fit in HBM can be solved; no
 * counterpart exists in the reference, which always assembles.  Host
 * struct, device pointers. */
typedef struct {
    int N, lmax;
    const double* positions; /* [N, 3] */
    const double* mie;       /* [N, 2, lmax] c128 (mie_scat) for spheres, or NULL */
    double k;                /* wavenumber in the medium */
    void* workspace;         /* mpa_matfree_workspace_bytes(N, lmax, nrhs) bytes for the largest nrhs used */
    const double* tmatrix;   /* [N, 2 rmax, 2 rmax] c128 per-particle T-matrices, or NULL; exactly one of
                                mie / tmatrix is set (the operator of particle_aggregate_tmatrix) */
} mpa_cluster_operator;

size_t mpa_matfree_workspace_bytes(int N, int lmax, int nrhs);

/* y = (add_identity ? x : 0) + T x, x and y [nrhs, n] c128 (device), FP64.
 * Same operator as mpa_sphere_aggregate_tmatrix followed by mpa_matvec. */
int mpa_matfree_matvec(const mpa_cluster_operator* op, int nrhs, int add_identity, const void* x, void* y,
                       void* stream);

/* The same product in complex64 on the matrix cores (the separation and
 * argument reduction in FP64, special functions and sums in FP32, products
 * on the FP16 matrix cores as hi/lo splits of scaled operands), x and y
 * [nrhs, n] c64: the operator of mixed-precision inner solves without a
 * stored matrix.  Spheres with lmax <= 4. */
int mpa_matfree_matvec_c64(const mpa_cluster_operator* op, int nrhs, int add_identity, const void* x, void* y,
                           void* stream);

/* Batched variant of mpa_matfree_matvec / mpa_matfree_matvec_c64 over a
 * table of sphere-cluster operators (the batched form the reference's
 * sphere_aggregate_tmatrix + GEMV would take over many clusters,
 * cpp/src/interactions.cpp:370-415 and :347-368): y[q] = (I +) T_q x[q] for
 * every q, nrhs[q] right-hand sides each, all in dtype.  With the complex64
 * kernel whose sources share each pair's special functions selected
 * (mpa_matfree_variant(MPA_C64, 6)) the complex64 products of all operators
 * of one lmax share launches (16 per launch); otherwise one launch sequence
 * per operator.  Results equal the one-operator calls bit for bit. */
int mpa_matfree_matvec_batched(int count, const mpa_cluster_operator* const* ops, const int* nrhs, int dtype,
                               int add_identity, const void* const* x, void* const* y, void* stream);

/* BiCGSTAB (as mpa_bicgstab, FP64) on (add_identity ? I + T : T) x = b with
 * the matrix-free operator; workspace: mpa_bicgstab_workspace_bytes(n,
 * nrhs, MPA_C128) bytes, n = 2 lmax (lmax + 2) N; op->workspace sized for
 * nrhs.  Same role as the bicgstab branch of solve_linear_system
 * (cpp/src/interactions.cpp:347-368) for a cluster whose matrix is not
 * stored. */
int mpa_bicgstab_matfree(const mpa_cluster_operator* op, int nrhs, int add_identity, const void* b, void* x,
                         double tol, int maxiter, void* workspace, int* iters_out, double* resid_out, int* status_out,
                         int sync_every, void* stream);

/* Autotuner for the matrix-core kernels of the matrix-free operator (no
 * counterpart in the reference): times their variants (one or two
 * accumulator chains per tile) on the caller's operator and vectors x, y
 * ([nrhs, n] of dtype; y is overwritten) and keeps the fastest for dtype. */
int mpa_autotune_matfree(const mpa_cluster_operator* op, int nrhs, int dtype, const void* x, void* y, int reps,
                         int* variant_out, double* ms_out, void* stream);
/* Matrix-core kernel variant of the matrix-free product in use for dtype;
 * set_to >= 0 selects one; *count_out (may be NULL) receives the number of
 * variants; -1 on a bad argument. */
int mpa_matfree_variant(int dtype, int set_to, int* count_out);

/* ---- batched solves: many independent systems, one set of launches -------- */

/* One system (I + T) x = b with nrhs right-hand sides (host struct; all
 * pointers are device pointers).  `state` holds the system's Krylov vectors
 * and scalars: mpa_system_state_bytes(n, nrhs, dtype) bytes.  With `op`
 * set, A must be NULL and the operator is applied matrix-free (dtype
 * MPA_C128 only; op->workspace sized for nrhs). */
typedef struct {
    int n, nrhs, add_identity, flags; /* flags: MPA_SOLVE_* bits */
    const void* A; /* [n, n] of dtype (T, or I + T with add_identity = 0) */
    const void* b; /* [nrhs, n] */
    void* x;       /* [nrhs, n], overwritten */
    void* state;
    double tol;    /* relative: ||r|| <= tol ||b|| (absolute with MPA_SOLVE_REF_LOOP) */
    int maxiter, block_size;
    const mpa_cluster_operator* op; /* host pointer or NULL */
    const void* minv; /* NULL, or the [n, block_size] right preconditioner of mpa_bicgstab_ex (dtype) */
} mpa_system;

size_t mpa_system_state_bytes(int n, int nrhs, int dtype);
/* State bytes of a system with (preconditioned != 0) or without minv. */
size_t mpa_system_state_bytes_ex(int n, int nrhs, int dtype, int preconditioned);
/* Device scratch for the descriptors and work lists of one batch. */
size_t mpa_batch_scratch_bytes(int nsys, const mpa_system* systems, int dtype);

/* BiCGSTAB (as mpa_bicgstab) for nsys systems at once: every kernel launch
 * covers all systems; converged systems cost nothing further.  The outputs
 * are host arrays of sum(nrhs) entries in (system, RHS) order. */
int mpa_bicgstab_batched(int nsys, const mpa_system* systems, int dtype, void* scratch, int* iters_out,
                         double* resid_out, int* status_out, int sync_every, void* stream);

/* Iteration-level control of the same engine, for continuous batching:
 * mpa_systems_start initializes the given systems (x = 0, r = b, scalars);
 * mpa_systems_advance runs `iterations` BiCGSTAB iterations over all given
 * systems (converged ones are no-ops; a pending restart is served first),
 * then synchronizes and reports per RHS, in (system, RHS) order: done flag,
 * iterations, ||r||/||b||, status (host arrays of sum(nrhs), any may be NULL).
 * The set of systems may change between calls; each system's state lives in
 * its own `state` buffer. */
int mpa_systems_start(int nsys, const mpa_system* systems, int dtype, void* scratch, void* stream);
int mpa_systems_advance(int nsys, const mpa_system* systems, int dtype, void* scratch, int iterations, int* done_out,
                        int* iters_out, double* resid_out, int* status_out, void* stream);

/* One system for the mixed-precision batched solve (host struct, device
 * pointers); `state`: mpa_refine_state_bytes(n, nrhs) bytes. */
typedef struct {
    int n, nrhs, add_identity, pad;
    const void* A32; /* [n, n] c64 */
    const void* A64; /* [n, n] c128, the FP64 residual operator */
    const void* b;   /* [nrhs, n] c128 */
    void* x;         /* [nrhs, n] c128 */
    void* state;
    double tol, inner_tol;
    int maxiter, max_refine;
    const mpa_cluster_operator* op; /* host pointer or NULL: with op set, A32 and A64 are NULL, the
                                       corrections use the complex64 matrix-core operator and the
                                       residual the FP64 one (no matrix is stored; spheres, lmax <= 5) */
} mpa_refine_system;

size_t mpa_refine_state_bytes(int n, int nrhs);
size_t mpa_refine_scratch_bytes(int nsys, const mpa_refine_system* systems);
/* mpa_solve_refined for nsys systems at once (one batched complex64 engine
 * for the corrections, one batched c128 matvec for the residuals). */
int mpa_solve_refined_batched(int nsys, const mpa_refine_system* systems, void* scratch, int* outer_out,
                              int* inner_out, double* resid_out, int* status_out, void* stream);

/* Iteration-level mixed-precision refinement (no counterpart in the
 * reference; the serving loop's form of mpa_solve_refined_batched): systems
 * join and leave between calls.  mpa_refine_systems_start sets x = 0 and
 * starts each system's first complex64 correction; each
 * mpa_refine_systems_advance call runs `iterations` BiCGSTAB iterations of
 * the running corrections of the given systems (any subset, in any order of
 * calls), and a system whose correction finished gets its FP64 update,
 * residual, per-RHS decision and next correction in the same call.
 * done_out[k] = 1 once every RHS of the system is decided; outer, inner,
 * resid and status as mpa_solve_refined_batched.  Same arguments and
 * scratch (mpa_refine_scratch_bytes) as mpa_solve_refined_batched; the
 * per-RHS control records live in each system's state. */
int mpa_refine_systems_start(int nsys, const mpa_refine_system* systems, void* scratch, void* stream);
int mpa_refine_systems_advance(int nsys, const mpa_refine_system* systems, void* scratch, int iterations,
                               int* done_out, int* outer_out, int* inner_out, double* resid_out, int* status_out,
                               void* stream);

/* Bytes of device workspace mpa_solve_refined needs. */
size_t mpa_refine_workspace_bytes(int n, int nrhs);

/* Mixed-precision solve of (I + T) x = b: complex64 BiCGSTAB corrections on
 * the FP32 matrix A32, FP64 iterative refinement with the FP64 residual
 * r = b - A64 x, until ||r|| <= tol ||b|| (per RHS) or max_refine
 * corrections or stagnation.  Same system as mpa_bicgstab (the reference's
 * solve_linear_system, cpp/src/interactions.cpp:347-368), answered to FP64
 * accuracy; it has no counterpart in the reference, which solves in FP64 only.
 *   A32  [n, n] c64, A64 [n, n] c128 (device); add_identity* as mpa_bicgstab
 *   b, x [nrhs, n] c128 (device); inner_tol = relative tolerance of each
 *   complex64 correction solve, maxiter its iteration cap
 *   outer_out, inner_out, resid_out, status_out: host arrays [nrhs] or NULL. */
int mpa_solve_refined(int n, int nrhs, const void* A32, int add_identity32, const void* A64, int add_identity64,
                      const void* b, void* x, double tol, double inner_tol, int maxiter, int max_refine,
                      void* workspace, int* outer_out, int* inner_out, double* resid_out, int* status_out,
                      void* stream);

/* Observables of a solved cluster for nsrc sources, all FP64 (device):
 *   p_inc, p_src [nsrc, N, 2, rmax] c128; mie_scat, mie_int [N, 2, lmax] c128
 *   p_scat (out) = p_inc * mie_scat[:, :, n_idx]        (sphere_cluster.py:716)
 *   p_int  (out, may be NULL) = p_inc * mie_int[:, ::-1, n_idx] (:717)
 *   C_part (out) [nsrc, N, 3, 2, lmax] scattering/absorption/extinction per
 *          multipole: flux.particle_cross_sections (cpp/src/flux.cpp:5-30)
 *   C_total (out, may be NULL) [nsrc, 3] cluster sums (sphere_cluster.py:442-454)
 *   F (out) [nsrc, N, 3]: forces.force_all (cpp/src/forces.cpp:7-86,139-154)
 *   T (out, may be NULL) [nsrc, N, 3]: forces.torque_all (:88-137,156-171)
 * eps_b is the (real) medium permittivity; mu_b is unused, as in the reference
 * (cpp/src/forces.cpp:10). */
int mpa_cluster_observables(int N, int lmax, int nsrc, const double* p_inc, const double* p_src,
                            const double* mie_scat, const double* mie_int, double k, double eps_b, double* p_scat,
                            double* p_int, double* C_part, double* C_total, double* F, double* T, void* stream);

/* One cluster of mpa_sphere_aggregate_tmatrix_batched: the arguments of
 * mpa_sphere_aggregate_tmatrix (dtype = its out_dtype), or, with out32 set
 * (dtype MPA_C128), of mpa_sphere_aggregate_tmatrix_dual (out = the c128
 * matrix, out32 = its complex64 rounding). */
typedef struct {
    int N, lmax, dtype, add_identity;
    const double* positions;
    const double* mie;
    double k;
    void* out;
    void* out32;
} mpa_asm_cluster;

/* Sphere-cluster aggregate T-matrices of `count` clusters of any sizes and
 * lmax in a few launches (descriptors in the kernel arguments, 32 per launch
 * and output kind); each matrix equals its own single-cluster call's bit for
 * bit.  The reference assembles each cluster on its own
 * (cpp/src/interactions.cpp:370-415 through sphere_aggregate_tmatrix,
 * cpp/src/interactions_pyb.cpp); this batched form is for many independent
 * clusters solved together. */
int mpa_sphere_aggregate_tmatrix_batched(int count, const mpa_asm_cluster* clusters, void* stream);

/* One cluster of mpa_cluster_observables_batched: the arguments of
 * mpa_cluster_observables for that cluster (same shapes and meanings; p_int,
 * mie_int, C_total and T may be NULL as there). */
typedef struct {
    int N, lmax, nsrc, pad;
    const double* p_inc;
    const double* p_src;
    const double* mie_scat;
    const double* mie_int;
    double k, eps_b;
    double* p_scat;
    double* p_int;
    double* C_part;
    double* C_total;
    double* F;
    double* T;
} mpa_obs_cluster;

/* mpa_cluster_observables for `count` clusters of any sizes and lmax in a few
 * launches (the cluster descriptors travel in the kernel arguments, 32 per
 * launch); each cluster's results equal those of its own
 * mpa_cluster_observables call bit for bit.  The reference reduces each
 * cluster on its own (src/miepy/sphere_cluster.py:442-471,715-717 over
 * cpp/src/flux.cpp:5-30 and cpp/src/forces.cpp:7-171); this batched form is
 * for many independent clusters solved together. */
int mpa_cluster_observables_batched(int count, const mpa_obs_cluster* clusters, void* stream);

/* Force and/or torque on M particles from explicit coefficients.
 * Replaces forces.force_all / torque_all (cpp/src/forces_pyb.cpp:27-43 ->
 * cpp/src/forces.cpp:139-171): p_scat, p_inc [M, 2 rmax] c128 (device);
 * F, T [M, 3] f64 (device, either may be NULL). */
int mpa_force_torque(int M, int lmax, const double* p_scat, const double* p_inc, double k, double eps_b, double* F,
                     double* T, void* stream);

/* Per-multipole cross-sections of M particles.
 * Replaces flux.particle_cross_sections (cpp/src/flux_pyb.cpp:12 ->
 * cpp/src/flux.cpp:5-30), batched: p_scat, p_inc, p_src [M, 2 rmax] c128;
 * C_part [M, 3, 2, lmax] f64 (scattering, absorption, extinction). */
int mpa_particle_cross_sections(int M, int lmax, const double* p_scat, const double* p_inc, const double* p_src,
                                double k, double* C_part, void* stream);

/* Process-wide runtime options (no counterpart in the reference). */
int mpa_set_option(int key, int value);
int mpa_get_option(int key);

/* Launch timing with HIP events recorded on the launching stream (off by
 * default).  kind: 0 matvec c128, 1 matvec c64, 2 assembly c128,
 * 3 assembly c64, 4 matrix-free matvec c128, 5 matrix-free matvec c64.
 * mpa_profile_enable(1) clears the records and starts recording;
 * mpa_profile_read sums the recorded launch groups of one kind (host outputs). */
int mpa_profile_enable(int on);
int mpa_profile_read(int kind, int* count, double* total_ms);

/* Plain device-to-device copy of `bytes` (multiple of 16) with 16-byte
 * loads and stores: the known-size reference kernel used to scale hardware
 * byte counters to HBM bytes in measurements. */
int mpa_stream_copy(const void* src, void* dst, size_t bytes, void* stream);

/* Host copy of the per-lmax translation coefficient table (tests):
 * meta_out [rmax*rmax*4] int, coef_out [rmax*rmax*(2 lmax+1)] c128. */
int mpa_translation_table(int lmax, int* meta_out, double* coef_out);

/* Column blocks per assembly workgroup tile in use for lmax. */
int mpa_assembly_tile(int lmax);

/* Autotuner for the assembly tiles (no counterpart in the reference): times
 * every tile variant for lmax (column blocks per workgroup; coefficient
 * table in LDS or read through the caches) on the caller's cluster, `reps`
 * launches each (`out`, [n, n] of out_dtype, is overwritten), and keeps the
 * fastest for the process.  Outputs as mpa_autotune_matvec. */
int mpa_autotune_assembly(int N, int lmax, const double* positions, const double* mie, double k, int out_dtype,
                          void* out, int reps, int* variant_out, double* ms_out, void* stream);
/* Assembly variant in use for lmax; set_to >= 0 selects one; *count_out
 * (may be NULL) receives the number of variants; -1 on a bad argument. */
int mpa_assembly_variant(int lmax, int set_to, int* count_out);

#ifdef __cplusplus
}
#endif

#endif /* MIEPY_AMD_H */
