/*
 * negf_gpu.h -- C-ABI of the MI355X RGF energy-loop library (libnegf_gpu.so).
 *
 * Plain C types only: int32 integers, double, and complex*16 passed as
 * interleaved (re, im) double pairs (`double _Complex` / Fortran complex(c_double_complex)
 * are layout-compatible).  All arrays are Fortran column-major, 1-based in Fortran
 * and 0-based here.  Every call is synchronous with respect to the host: results
 * are in caller memory when it returns.  One host thread per context at a time.
 *
 * Return status (all entry points): 0 = OK, < 0 = bad argument or HIP error
 * (negf_last_error() gives the message), > 0 = 1-based slice index at which an
 * exactly-zero pivot was met (the reference ignores LAPACK's INFO and lets the
 * resulting Inf/NaN propagate; the library does the same and only reports it).
 *
 * Citations are to the NEGFtrans Fortran sources (src/...:line).
 */
#ifndef NEGF_GPU_H
#define NEGF_GPU_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct negf_ctx negf_ctx;

/* Precision of the bulk block products (flags of negf_ctx_create). */
#define NEGF_PREC_FP64       0   /* complex128 everywhere (parity path)              */
#define NEGF_PREC_MIXED_FP32 1   /* FP32 sweep kept in HBM + FP64 iterative refinement  */
/* V-hat window: faithful copies only |p|<=NGX/2, |q|<=NGY/2 (src/main.f90:242-246). */
#define NEGF_FULL_WINDOW     16  /* extension: use the whole +-NGX, +-NGY window    */
/* Extension, no reference counterpart (the reference adds exp(i q Delta) with the
 * continuum q, src/grid.f90:284-289,323-331): lead terms of a semi-infinite flat lead on
 * the device's own finite-difference lattice, computed on the GPU by decimation (FP64,
 * |coupling| < 1e-14 convergence test, at most 200 doubling steps; a call fails when a
 * channel does not converge).  Trace weights become Im(term) / Delta^2 on the
 * lattice-open channels.  A V = 0 device then transmits exactly 1 per open channel. */
#define NEGF_LEAD_DECIMATION 32
/* Extension, no reference counterpart (the reference's only tuning knob is LAPACK's
 * ilaenv block size, src/math_kernel.f90:17): FP64 negf_transmission_batch picks its
 * kernel shape and form by measurement.  Variants: 0 one workgroup per job with the
 * 32-column two-tile Gauss-Jordan (Np <= 512), 1 two workgroups per job (two-sided) with
 * it, 2 one workgroup per job with the 16-column four-tile form, 3 two-sided with it
 * (above Np = 512 only 2 and 3).  On the first call of a job-count class (ceil(log2
 * jobs)) every variant runs over the call's own jobs; the fastest is kept for the class
 * and its results are returned.  Without the flag a fixed rule picks: two-sided when two
 * workgroups per job fit the GPU at once, two-tile up to Np = 512.  The environment
 * variable NEGF_TRANS_VARIANT (read at context creation) forces one variant. */
#define NEGF_AUTOTUNE        64

/* Library version string and last error message (thread-local). */
const char *negf_version(void);
const char *negf_last_error(void);
/* Number of visible GPUs (0 when none); does not create a context. */
int negf_device_count(void);

/*
 * Create a context for one device structure (one transport system).
 *   N        : plane-wave basis size (size(nx_grid)), N >= 1
 *   N_z      : number of z slices, N_z >= 2
 *   nx_grid, ny_grid : basis (src/grid.f90:410-431), length N
 *   Lx, Ly, Lz, V_L, V_R, eta : atomic units as in t_parameters (src/global.f90:5-7)
 *   vrec     : V_reciprocal_all(-NGX:NGX, -NGY:NGY, 1:N_z) complex128 (src/main.f90:218),
 *              uploaded once and kept resident on the GPU
 * Replaces the per-job setup that src/negf.f90:139-150 repeats for every energy.
 */
int negf_ctx_create(negf_ctx **out, int N, int N_z, const int *nx_grid, const int *ny_grid,
                    double Lx, double Ly, double Lz, double V_L, double V_R, double eta,
                    const double *vrec, int NGX, int NGY, int flags);
int negf_ctx_destroy(negf_ctx *ctx);

/*
 * Batched transmission (replaces the job loop src/main.f90:384-398 over
 * Transmission_Coefficient, src/negf.f90:115-174):
 *   tau[ie] += sum_ik kw[ik] * tau(E[ie], k = (kx[ik], ky[ik]))      (ie < nE, ik < nk)
 * E in Hartree (real; +i*eta is added as at src/negf.f90:149).  All nE*nk jobs run
 * in one persistent launch; when they number at most half the workgroups the GPU
 * holds, each job runs on two workgroups (forward and backward sweeps joined at
 * N_z/2, a second short launch), which agrees with the one-workgroup form to rounding.
 * NEGF_DIAG_ONE_SIDED=1 in the environment at context creation keeps one workgroup per
 * job.  tau_jobs (optional, may be NULL) receives the per-job tau before the k weight,
 * ordered job = ie + nE*ik (the reference's i_job - 1).
 */
int negf_transmission_batch(negf_ctx *ctx, const double *E, int nE, const double *kx,
                            const double *ky, const double *kw, int nk, double *tau,
                            double *tau_jobs);

/*
 * Drop-in for GreensFunction_tri_solver(G_inv_D, G_Function, DFL)
 * (src/math_kernel.f90:44-165): off-diagonal blocks are identities.
 *   G_inv_D : N*N*N_z complex128
 *   mode    : bit mask, 1 = "D" (diagonal), 2 = "F" (first column), 4 = "L" (last column)
 *   diag, first, last : N*N*N_z complex128 each, or NULL when not requested; requested
 *              outputs are fully written, the t_gfunc components not requested are the
 *              caller's to zero (src/math_kernel.f90:75-77).
 */
int negf_rgf_tri_solve(int N, int N_z, const double *G_inv_D, int mode, double *diag,
                       double *first, double *last);

/*
 * Batched equilibrium density (replaces src/main.f90:272-288 over Equilibrium_Density,
 * src/negf.f90:9-113): for each complex node z[n] (without eta; eta is added) and
 * complex weight w[n] (the contour weights of src/negf.f90:86-111, k weight applied here):
 *   density_matrix(:, :, k) += sum_n sum_ik kw[ik] * w[n] * 2*Delta^2 * G_kk(z[n] + i eta)
 * density_matrix : N*N*N_z complex128, accumulated in place (intent(inout)).
 */
int negf_density_batch(negf_ctx *ctx, const double *z, const double *w, int nnode,
                       const double *kx, const double *ky, const double *kw, int nk,
                       double *density_matrix);

/*
 * LDOS for a list of real energies (replaces Local_Density_Of_State,
 * src/negf.f90:176-236, summed over k): per energy ie
 *   ldos_matrix(:, :, k, ie) = - sum_ik kw[ik] / pi * 2*Delta^2 * G_kk(E[ie] + i eta)
 * ldos_matrix : N*N*N_z*nE complex128 (overwritten).
 */
int negf_ldos_batch(negf_ctx *ctx, const double *E, int nE, const double *kx,
                    const double *ky, const double *kw, int nk, double *ldos_matrix);

/*
 * Plane-wave -> real-space diagonal of nmat N x N blocks
 * (Greensfunction_PlanewaveToReal, src/grid.f90:71-135):
 *   out(m_x, m_y, s) = sum_ij M_s(i,j) exp(+2 pi i [(nx_i-nx_j) m_x/N_x + (ny_i-ny_j) m_y/N_y])
 * mats : N*N*nmat complex128 ; out : N_x*N_y*nmat complex128.
 */
int negf_pw_to_real(int N, const int *nx_grid, const int *ny_grid, int nmat,
                    const double *mats, int N_x, int N_y, double *out);

/*
 * Real-space potential -> V_reciprocal_all (LocalPotential_RealToPlanewave,
 * src/grid.f90:12-69, for every slice, plus the window copy of src/main.f90:238-247).
 *   V_real : N_x*N_y*N_z real (Hartree) ; vrec : (2NGX+1)*(2NGY+1)*N_z complex128 (overwritten)
 *   full_window : 0 = copy only |p|<=NGX/2, |q|<=NGY/2 (reference), 1 = whole window
 */
int negf_vhat_from_real(int N_x, int N_y, int N_z, const double *V_real, int NGX, int NGY,
                        int full_window, double *vrec);

/*
 * Batched drop-ins for the two private LAPACK/BLAS wrappers of src/math_kernel.f90:
 *   negf_inverse_batch        inv[b] = A[b]^{-1}   (inverse, :8-34: zgetrf + zgetri)
 *   negf_matrix_product_batch C[b] = A[b] B[b]     (matrix_product, :36-42: zgemm 'N','N')
 * count N x N complex128 matrices, contiguous.  Pivoting follows zgetrf (largest
 * |re|+|im| in the column, first index on ties).
 */
int negf_inverse_batch(int N, int count, const double *A, double *inv);
int negf_matrix_product_batch(int N, int count, const double *A, const double *B, double *C);

/*
 * negf_inverse_batch without any allocation.  It has no counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under graph capture, serving
 * loops, harnesses that forbid hipMalloc).  A, inv (count N x N complex128 each) and
 * status (count ints: 0, or the 1-based column of the first exactly-zero pivot) are
 * DEVICE pointers owned by the caller; the work areas come from a 64 MiB array the
 * library holds on each GPU, so count * (Np^2 + 2 PB Np) * 16 bytes must fit in it
 * (Np = N padded to 32, PB = 32; above Np = 512, padded to 16 and PB = 16) -- larger
 * batches are rejected with an error.  The work is queued on `stream` (a hipStream_t;
 * null = the default stream) and not waited for.  Calls that overlap on one GPU race on
 * the shared array.  Returns 0 or a negative error code.
 */
int negf_inverse_batch_noalloc(int N, int count, const double *A, double *inv, int *status, void *stream);

/*
 * Slice blocks on the GPU, for callers that want them materialised:
 *   negf_hamiltonian : H(:, :, k) of Hamiltonian_construction (src/grid.f90:236-262)
 *   negf_e_minus_h   : D_k of E_minus_H_construction (src/grid.f90:264-333) at complex
 *                      energy (E_re + i E_im) -- eta is NOT added here, as in the reference
 *                      where the caller passes energy + i*ETA
 * out : N*N*N_z complex128 (overwritten).  (The batched solvers build the same blocks
 * on the fly and never store them.)
 */
int negf_hamiltonian(negf_ctx *ctx, double kx, double ky, double *out);
int negf_e_minus_h(negf_ctx *ctx, double E_re, double E_im, double kx, double ky, double *out);

/*
 * Context-free forms with the argument lists of the reference routines:
 *   negf_hamiltonian_direct  Hamiltonian_construction(nx, ny, Lx, Ly, kx, ky, V_reciprocal, H)
 *                            (src/grid.f90:236-262)
 *   negf_e_minus_h_from_h    E_minus_H_construction(H, energy, nx, ny, Lx, Ly, Lz, kx, ky,
 *                            V_L, V_R, E_minus_H) (src/grid.f90:264-333), energy complex
 *   negf_transmission_trace  Transmission_solver on G = first_column(:, :, N_z) already
 *                            rescaled by 2 Delta^2 (src/grid.f90:335-385); Delta = Lz/N_z
 * H, out : N*N*N_z complex128 ; G : N*N complex128.
 */
int negf_hamiltonian_direct(int N, int N_z, const int *nx_grid, const int *ny_grid, double Lx,
                            double Ly, double kx, double ky, const double *vrec, int NGX, int NGY,
                            double *out);
int negf_e_minus_h_from_h(int N, int N_z, const double *H, double E_re, double E_im,
                          const int *nx_grid, const int *ny_grid, double Lx, double Ly, double Lz,
                          double kx, double ky, double V_L, double V_R, double *out);
int negf_transmission_trace(int N, int N_z, const double *G, double E, const int *nx_grid,
                            const int *ny_grid, double Lx, double Ly, double Lz, double kx,
                            double ky, double V_L, double V_R, double *tau);

/*
 * Extension, no reference counterpart (the reference rejects VDS != 0,
 * src/main.f90:129-132): the bias-window part of the non-equilibrium density,
 *   density_matrix(:, :, k) += sum_{n, kp} kw(kp) w(n) [G Gamma_L G^dagger]_kk (E(n) + i eta)
 * with G rescaled by 2 Delta^2 (as Equilibrium_Density rescales, src/negf.f90:73-84) and
 * Gamma_L = diag(sin(Q^L_i Delta) / Delta^2) on slice 1 (left lead, real energy, closed
 * channels 0, the trace's formula at src/grid.f90:355-360).  E : nE real energies
 * (Hartree, eta added inside); w : nE complex weights (re, im interleaved);
 * density_matrix : N*N*N_z complex128, accumulated as negf_density_batch does.
 */
int negf_neq_density_batch(negf_ctx *ctx, const double *E, const double *w, int nE, const double *kx,
                           const double *ky, const double *kw, int nk, double *density_matrix);

/* Wall time (seconds) of the last batched call's device work, from HIP events. */
double negf_last_kernel_seconds(negf_ctx *ctx);

/* Extension, no reference counterpart.  Number of energies of the last
 * negf_transmission_batch call on a NEGF_PREC_MIXED_FP32 context whose refinement did
 * not reach its tolerance within the iteration limit and were recomputed on the FP64
 * path (their results are FP64 results).  0 for FP64 contexts. */
int negf_last_fallbacks(const negf_ctx *ctx);

/* Extension, no reference counterpart.  Variant (see NEGF_AUTOTUNE) of the last FP64
 * negf_transmission_batch call, -1 before any; and the device seconds each variant took
 * in the context's last tuning call (-1 for a variant not run). */
int negf_last_variant(const negf_ctx *ctx);
double negf_tune_seconds(const negf_ctx *ctx, int variant);

/*
 * NonLocalPotential_RealToPlanewave (src/grid.f90:137-234; the reference never calls it):
 *   vrec(p, q, p', q') = (N_x N_y)^-2 sum V(x, y, x', y')
 *                        exp(-2 pi i (p x / N_x + q y / N_y - p' x' / N_x - q' y' / N_y))
 * for p, p' in -Rx..Rx and q, q' in -Ry..Ry.  V_real: N_x x N_y x N_x x N_y complex128,
 * column-major; vrec: (2Rx+1) x (2Ry+1) x (2Rx+1) x (2Ry+1) complex128, column-major,
 * index 0 = frequency -R.  The reference's check 2R <= N is kept; at |f| = N/2 the
 * reference reads past its work array, here the periodic value is returned.
 */
int negf_nonlocal_vhat_from_real(int N_x, int N_y, const double *V_real, int Rx, int Ry, double *vrec);

/*
 * Extension, no reference counterpart: replace V_reciprocal_all of a context (same NGX,
 * NGY, N_z), e.g. one step of a self-consistent potential loop.  Synchronous.
 */
int negf_ctx_set_potential(negf_ctx *ctx, const double *vrec);

/*
 * Extension, no reference counterpart (the reference runs one device per program):
 * uneven batches.  A negf_multi holds FP64 contexts of one GPU with any N and N_z.
 * negf_multi_transmission adds, for every device d and energy e,
 *   tau[d][e] += sum_k kw[d][k] tau(E[d][e], kx[d][k], ky[d][k])
 * exactly as negf_transmission_batch on that context, in one hipGraph: contexts of the
 * same padded size share one launch (jobs handed to workgroups longest-N_z first), and
 * the launches of different sizes run concurrently on forked streams.  The inversion
 * variant of each padded size (matrix-core two-lanes-per-row or VALU one-thread-per-row
 * Gauss-Jordan) is chosen by timing both on the first call; NEGF_MULTI_VARIANT=mfma|valu
 * forces one.  Returns 0, the first zero-pivot slice (> 0), or < 0.  The contexts are
 * not owned: they must outlive the negf_multi.
 */
typedef struct negf_multi negf_multi;
int negf_multi_create(negf_multi **out, int ndev, negf_ctx *const *ctxs);
int negf_multi_transmission(negf_multi *m, const double *const *E, const int *nE, const double *const *kx,
                            const double *const *ky, const double *const *kw, const int *nk,
                            double *const *tau);
/* Launch the graph of the last negf_multi_transmission again (same jobs; the contexts'
 * potentials may have been replaced) and accumulate into tau the same way. */
int negf_multi_replay(negf_multi *m, double *const *tau);
/* LDOS of the uneven batch at selected slices: for every device d, energy e and kept
 * slice q (slices[d][q], 0-based), ldos[d] (N_d x N_d x nsel x nE[d] complex128,
 * overwritten) receives -sum_k kw[d][k] / pi G_kk(E[d][e] + i eta) at that slice -- the
 * same blocks negf_ldos_batch gives for all slices.  One hipGraph as above (the outputs
 * are zeroed by a kernel inside it). */
int negf_multi_ldos(negf_multi *m, const double *const *E, const int *nE, const double *const *kx,
                    const double *const *ky, const double *const *kw, const int *nk, int nsel,
                    const int *const *slices, double *const *ldos);
/* Device time of the last negf_multi_transmission, replay or LDOS call (HIP events), seconds. */
double negf_multi_last_seconds(const negf_multi *m);
/* Variant chosen for padded size Np: 0 matrix cores, 1 VALU, -1 not tuned yet. */
int negf_multi_variant(int Np);
int negf_multi_destroy(negf_multi *m);

/*
 * Extension, no reference counterpart: kernels of the learned adaptive energy grid (a
 * multilayer perceptron trained on T(E) curves this library computes).  Device pointers,
 * FP32 row-major storage with leading dimensions, work queued on `stream` (a hipStream_t;
 * NULL = the default stream).  Products round both operands to bfloat16 (nearest even)
 * and accumulate in FP32 on the matrix cores.
 *
 * negf_mlp_gemm: C (M x N) = op(A) (M x K) op(B) (K x N); op(A)(i,k) = transA ? A[k lda + i]
 *   : A[i lda + k], likewise B.  mode 0: C = product (+ C if accumulate); mode 1:
 *   Z = product + bias[j], C = GELU(Z) (exact erf form); mode 2: C = product + bias[j].
 * negf_mlp_gelu_backward: dZ = dY GELU'(Z) over n entries.
 * negf_mlp_colsum: out[j] = sum_i X[i ldx + j].
 * negf_mlp_nll: Gaussian negative log-likelihood of T (B x P) under O (B x 2P) = [mean |
 *   log-variance clamped to +-20]: *loss += mean over B P of 0.5 ((T - mu)^2 e^-s + s);
 *   G (B x 2P) = d loss / d O.
 * negf_mlp_adam: w -= lr_t m / (sqrt(v) + eps) after the moment updates (lr_t carries the
 *   bias correction).
 */
int negf_mlp_gemm(int M, int N, int K, const float *A, int lda, int transA, const float *B, int ldb, int transB,
                  float *C, int ldc, const float *bias, float *Z, int mode, int accumulate, void *stream);
int negf_mlp_gelu_backward(int n, const float *Z, const float *dY, float *dZ, void *stream);
int negf_mlp_colsum(int M, int N, const float *X, int ldx, float *out, void *stream);
int negf_mlp_nll(int B, int P, const float *O, const float *T, float *G, float *loss, void *stream);
int negf_mlp_adam(int n, float *w, const float *g, float *m, float *v, float lr_t, float b1, float b2, float eps,
                  void *stream);
/* Device buffers for the entry points above: allocate, free, and a synchronous copy
 * (kind 0 host -> device, 1 device -> host, 2 device -> device; it waits for the work
 * queued before it on the default stream). */
int negf_dev_alloc(void **ptr, size_t bytes);
int negf_dev_free(void *ptr);
int negf_dev_copy(void *dst, const void *src, size_t bytes, int kind);

#ifdef __cplusplus
}
#endif
#endif /* NEGF_GPU_H */
