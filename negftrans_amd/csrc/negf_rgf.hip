// negf_rgf.hip -- MI355X (gfx950) implementation of the NEGFtrans RGF energy loop
// behind the C-ABI declared in include/negf_gpu.h.
//
// Hot path: k_transmission, one persistent launch over every (energy, k-point) job.
// Each workgroup owns one job at a time and runs the whole backward sweep of
// GreensFunction_tri_solver "F" (src/math_kernel.f90:80-129) with O(N^2) state:
//     g_{N_z} = D_{N_z}^{-1},  g_m = (D_m - g_{m+1})^{-1},  Q = g_{N_z} g_{N_z-1} ... g_2,
//     G_{N_z,1} = (-1)^{N_z-1} Q (D_1 - g_2)^{-1}
// (the reference's generators in the same order; only the product is associated from
// the right instead of being propagated slice by slice), building each D_m on the fly
// from the resident V-hat table, and fusing the Caroli trace of Transmission_solver
// (src/grid.f90:335-385) into the last block product.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <queue>
#include <string>
#include <vector>

#include "../../include/negf_gpu.h"
#include "rgf_device.h"

// ------------------------------------------------------------------ errors
static thread_local std::string g_err;
static int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}
// Shared with the other translation units of the library (not exported).
__attribute__((visibility("hidden"))) int negf_fail_internal(int code, const std::string &msg) {
    return fail(code, msg);
}
#define HIPCHK(x)                                                                      \
    do {                                                                               \
        hipError_t e_ = (x);                                                           \
        if (e_ != hipSuccess)                                                          \
            return fail(-100 - (int)e_, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

static const double PI = 3.141592653589793;

// ------------------------------------------------------------------ kernels
// Only the open lead channels enter the trace: tau = sum gL_i gR_j |G_{N_z,1}(i,j)|^2 over
// rows i with gL_i != 0 and columns j with gR_j != 0 (src/grid.f90:355-383; closed
// channels have Q = 0 and sin(0) = 0).  The sweep therefore carries R = E_L^T g_{N_z} ...
// g_m (the open rows of the running product, n_L x N) instead of the full N x N product:
// R <- R g_m costs n_L N^2 per slice and is exact (no change of the arithmetic that
// reaches the trace), while the inversion keeps its full N^3 cost.
#define NEGF_RA 8   // rows of R accumulated per pass of the VALU row product
// The row product runs on the matrix cores at padded sizes 256..320 (the VALU form's
// strided loads cost more there than the 16-row MFMA tiles' padding), on the VALU
// elsewhere; single-device and uneven-batch kernels make the same choice, so a batch's T
// equals the single-device one-workgroup T to the last bit.
#define NEGF_RP_MFMA_LO 256
#define NEGF_RP_MFMA_HI 320

// Workgroups of at most 256 threads (padded sizes up to 128) are held to two waves per
// SIMD, i.e. at most 256 VGPRs + AGPRs per lane, so that a CU holds four of them (left
// to itself the compiler takes 224 VGPRs and 36-96 AGPRs there: one wave per SIMD).
#define NEGF_SMALL_WPE(MAXT) __attribute__((amdgpu_waves_per_eu((MAXT) <= 256 ? 2 : 1)))

struct TransArgs {
    SysView sv;           // the device of every job, unless svs is set
    const SysView *svs;   // uneven batches: per device, indexed by jdev[job] (same Np for all)
    const int *jdev;
    const int *sched_off, *sched_job;   // uneven batches: jobs of workgroup b are
                                        // sched_job[sched_off[b] .. sched_off[b + 1])
    int njobs;
    const cd *z;          // per job E + i eta
    const int *kidx;      // per job k-point index
    const double *kinh;   // per k-point, Np each
    const cd *sigL, *sigR;      // per job, Np each
    const int *open_off;  // per job offset into the open-channel lists
    const int *nL, *nR;   // per job open-row / open-column counts
    const int *Lrow, *Rcol;     // open row indices / open column indices (flat)
    const double *gL, *gR;      // matching gamma values (flat)
    int nLmax;            // max n_L over jobs (R buffers are nLmax x Np)
    cd *ws;               // per slot 2 Np^2 + 2 nLmax Np
    double *tau;          // per job
    int *status;          // per job: first zero pivot (1-based slice) or 0
    // k_transmission_split only: launch phase (0: both sweeps, 1: join), first job of
    // the launch, column stride of the open-channel chains, and per job and side the
    // first zero pivot of that side's sweep
    int phase, job_base, ldX;
    int *status2;
};

// Rn(a, c) = sum_k R(a, k) S(k, sig(c)), a < nL, c < Np (R, Rn: ld ldR) on the FP64
// matrix cores (v_mfma_f64_16x16x4f64, four real products per complex one).  Every wave
// takes 16-column tiles of Rn, 16 rows of R at a time: A(i, k) = R(a0 + i, k0 + k) in
// lane i + 16 k (zero for a0 + i >= nL), B(k, j) = S(k0 + k, sig(c0 + j)) in lane
// j + 16 k, so one load instruction reads four consecutive rows of 16 columns and the
// next k-step reuses the same cache lines.  D(i, j) comes back in lane j + 16 (i % 4),
// register i / 4.  Np % 16 == 0.
__device__ __forceinline__ void row_product_mfma(cd *__restrict__ Rn, const cd *__restrict__ R, int nL, int ldR,
                                                 const cd *__restrict__ S, const int *sig, int Np) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    for (int a0 = 0; a0 < nL; a0 += 16) {
        const bool arow = a0 + li < nL;
        const cd *ra = R + (a0 + li);
        for (int c0 = 16 * w; c0 < Np; c0 += 16 * nw) {
            const cd *sc = S + (size_t)sig[c0 + li] * Np + lk;
            negf_d4 re0 = {0.0, 0.0, 0.0, 0.0}, re1 = re0, im0 = re0, im1 = re0;
            for (int k0 = 0; k0 < Np; k0 += 16) {
                cd av[4], bv[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    bv[u] = sc[k0 + 4 * u];
                    av[u] = arow ? ra[(size_t)(k0 + 4 * u + lk) * ldR] : cmk(0.0, 0.0);
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    re0 = __builtin_amdgcn_mfma_f64_16x16x4f64(av[u].x, bv[u].x, re0, 0, 0, 0);
                    re1 = __builtin_amdgcn_mfma_f64_16x16x4f64(-av[u].y, bv[u].y, re1, 0, 0, 0);
                    im0 = __builtin_amdgcn_mfma_f64_16x16x4f64(av[u].x, bv[u].y, im0, 0, 0, 0);
                    im1 = __builtin_amdgcn_mfma_f64_16x16x4f64(av[u].y, bv[u].x, im1, 0, 0, 0);
                }
            }
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int a = a0 + 4 * r + lk;
                if (a < nL) Rn[a + (size_t)(c0 + li) * ldR] = cmk(re0[r] + re1[r], im0[r] + im1[r]);
            }
        }
    }
    __syncthreads();
}

struct RowProdShared {
    cd Rs[NEGF_RA * 64];
};

// The same product on the VALU, NEGF_RA rows of R at a time staged in LDS: for small
// blocks, where the matrix-core form's 16-row tiles and MFMA latency cost more than its
// better loads gain (Rn, R: ld ldR).
__device__ __forceinline__ void row_product(cd *__restrict__ Rn, const cd *__restrict__ R, int nL,
                                            int ldR, const cd *__restrict__ S, const int *sig,
                                            int Np, RowProdShared &rs) {
    // column c of the result belongs to a lane pair (two threads per column when the
    // workgroup has 2 Np threads, else one); lane h of the pair sums the k = h mod 2 terms
    const bool pair = (int)blockDim.x >= 2 * Np;
    const int c = pair ? (threadIdx.x >> 1) : threadIdx.x, h = pair ? (threadIdx.x & 1) : 0;
    const int step = pair ? 2 : 1;
    const cd *scol = S + (size_t)(c < Np ? sig[c] : 0) * Np;
    for (int a0 = 0; a0 < nL; a0 += NEGF_RA) {
        cd acc[NEGF_RA];
#pragma unroll
        for (int a = 0; a < NEGF_RA; ++a) acc[a] = cmk(0.0, 0.0);
        for (int k0 = 0; k0 < Np; k0 += 64) {
            for (int e = threadIdx.x; e < NEGF_RA * 64; e += blockDim.x) {
                const int a = e % NEGF_RA, k = k0 + e / NEGF_RA;
                rs.Rs[e] = (a0 + a < nL && k < Np) ? R[(a0 + a) + (size_t)k * ldR] : cmk(0.0, 0.0);
            }
            __syncthreads();
            if (c < Np) {
                const int kend = min(64, Np - k0);
                for (int kk = h; kk < kend; kk += step) {
                    const cd sv = scol[k0 + kk];
#pragma unroll
                    for (int a = 0; a < NEGF_RA; ++a) acc[a] = cfma(rs.Rs[a + kk * NEGF_RA], sv, acc[a]);
                }
            }
            __syncthreads();
        }
        if (pair) {
#pragma unroll
            for (int a = 0; a < NEGF_RA; ++a) acc[a] = cadd(acc[a], lane_swap(acc[a]));
        }
        if (c < Np && h == 0) {
#pragma unroll
            for (int a = 0; a < NEGF_RA; ++a)
                if (a0 + a < nL) Rn[(a0 + a) + (size_t)c * ldR] = acc[a];
        }
    }
    __syncthreads();
}

// TWO: the trailing update on the FP64 matrix cores -- gje_invert4 (two lanes per row,
// 32-column panels, Np <= 512) or, with 16-column panels, gje_invert4n (four-tile waves,
// Np <= NEGF_MAXNP, shared state sized for it); otherwise one thread per row on the VALU
// (gje_invert2, kept as the uneven batch's VALU variant).
template <int PB, bool TWO, int MAXNP = NEGF_MAXNP4, int TCH = NEGF_TCH>
using GjeSh = typename std::conditional<TWO, Gje4Shared<PB, double, (PB == 16 ? NEGF_MAXNP : MAXNP), TCH>,
                                        GjeShared<PB>>::type;

// (TWO with 16-column panels: the four-tile form for 512 < Np <= 1024, gje_invert4n)
template <int PB, bool TWO, class SH>
__device__ __forceinline__ int gje_any(cd *Z, int Np, SH &s, cd *scratch) {
    if constexpr (TWO && PB == 16) return gje_invert4n<16, 4, double, SH>(Z, Np, s, scratch);
    else if constexpr (TWO) return gje_invert4<PB, double, SH>(Z, Np, s, scratch);
    else return gje_invert2<PB>(Z, Np, s, scratch);
}

// MULTI: uneven batch -- jobs from a.sched_job, each with its device's SysView (a.svs);
// otherwise every job uses a.sv and workgroup b runs jobs b, b + gridDim.x, ...
// MAXNP / TCH: size limits of the Gauss-Jordan shared state (smaller for small Np, so
// that several workgroups fit on a CU).
template <int MAXT, int PB, bool TWO, bool MULTI = false, int MAXNP = NEGF_MAXNP4, int TCH = NEGF_TCH>
__global__ __launch_bounds__(MAXT) NEGF_SMALL_WPE(MAXT) void k_transmission(TransArgs a) {
    __shared__ GjeSh<PB, TWO, MAXNP, TCH> s;
    __shared__ RowProdShared rs;
    const int Np = a.sv.Np, ldR = a.nLmax;
    const size_t NN = (size_t)Np * Np;
    cd *base = a.ws + (size_t)blockIdx.x * (2 * NN + 2 * (size_t)ldR * Np + 2 * (size_t)PB * Np);
    cd *scratch = base + 2 * NN + 2 * (size_t)ldR * Np;
    const int i0 = MULTI ? a.sched_off[blockIdx.x] : blockIdx.x;
    const int i1 = MULTI ? a.sched_off[blockIdx.x + 1] : a.njobs;
    const int di = MULTI ? 1 : (int)gridDim.x;
    for (int i = i0; i < i1; i += di) {
        const int job = MULTI ? a.sched_job[i] : i;
        const SysView &sv = MULTI ? a.svs[a.jdev[job]] : a.sv;
        const int nL = a.nL[job], nR = a.nR[job], off = a.open_off[job];
        if (nL == 0 || nR == 0) {             // no open channel: tau = 0 exactly
            if (threadIdx.x == 0) {
                a.tau[job] = 0.0;
                a.status[job] = 0;
            }
            continue;
        }
        JobView jv;
        jv.z = a.z[job];
        jv.kinh = a.kinh + (size_t)a.kidx[job] * Np;
        jv.sigL = a.sigL + (size_t)job * Np;
        jv.sigR = a.sigR + (size_t)job * Np;
        cd *S = base, *M = base + NN;               // latest generator / block being inverted
        cd *R = base + 2 * NN, *Rn = R + (size_t)ldR * Np;
        int bad = 0;
        double part = 0.0;
        for (int m = sv.Nz; m >= 1; --m) {
            // M = D_m - g_{m+1} (m = N_z: M = D_{N_z}), inverted in place
            build_block(M, sv, jv, m - 1, m == sv.Nz ? nullptr : S, s.sigma);
            __syncthreads();
            if (gje_any<PB, TWO>(M, Np, s, scratch) && !bad) bad = m;
            if (m == sv.Nz) {
                // R = open rows of g_{N_z}
                for (int e = threadIdx.x; e < nL * Np; e += blockDim.x) {
                    const int ai = e % nL, c = e / nL;
                    R[ai + (size_t)c * ldR] = M[a.Lrow[off + ai] + (size_t)s.sigma[c] * Np];
                }
                __syncthreads();
            } else if (m >= 2) {
                if (Np >= NEGF_RP_MFMA_LO && Np <= NEGF_RP_MFMA_HI)
                    row_product_mfma(Rn, R, nL, ldR, M, s.sigma, Np);   // R <- R g_m
                else row_product(Rn, R, nL, ldR, M, s.sigma, Np, rs);
                cd *t = R;
                R = Rn;
                Rn = t;
            } else {
                // G_{N_z,1}(open rows, open cols) = +-R G_11(:, open cols); trace
                for (int e = threadIdx.x; e < nL * nR; e += blockDim.x) {
                    const int ai = e % nL, bi = e / nL;
                    const cd *gc = M + (size_t)s.sigma[a.Rcol[off + bi]] * Np;
                    cd acc = cmk(0.0, 0.0);
                    for (int k = 0; k < Np; ++k) acc = cfma(R[ai + (size_t)k * ldR], gc[k], acc);
                    part += a.gL[off + ai] * a.gR[off + bi] * (acc.x * acc.x + acc.y * acc.y);
                }
            }
            cd *t = S;
            S = M;
            M = t;
        }
        const double tot = block_sum2(part, s.red);
        if (threadIdx.x == 0) {
            a.tau[job] = tot * sv.scale * sv.scale;
            a.status[job] = bad;
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------ mixed precision
// NEGF_PREC_MIXED_FP32: the backward sweep runs in FP32 (build rounded from FP64, FP32
// Gauss-Jordan on the matrix cores) and keeps every generator g_m (and its column
// permutation) in HBM -- N_z Np^2 complex64 per workgroup, 0.82 GB at Np = 320,
// N_z = 1000.  G(:, 1) on the open columns then solves A X = E_1(:, open) for the scaled
// block-tridiagonal A = tridiag(I, D_m, I) of GreensFunction_tri_solver
// (src/math_kernel.f90:44-165): X_1 = g_1 E_1, X_m = -g_m X_{m-1}, followed by
// iterative refinement with FP64 residuals r = E_1 - A X (D_m rebuilt in FP64 from the
// V-hat table) and correction solves through the stored generators,
//     y_m = g_m (r_m - y_{m+1})  (m = N_z .. 1),  d_1 = y_1,  d_m = y_m - g_m d_{m-1},
// X += d.  The trace uses X_{N_z} = G_{N_z,1}(:, open).
#define NEGF_NRH 16   // right-hand-side columns of one solve (open channels at slice 1)

struct MixedArgs {
    TransArgs a;      // job tables (a.ws unused)
    int nref;         // refinement iterations
    double tol;       // stop when the estimated error of the job's T is below tol
    int ldb;          // columns of X and Y (max n_R over the jobs, <= NEGF_NRH)
    cfl *hist;        // per slot N_z Np^2: g_m, column-permuted like the Gauss-Jordan output
    int *hsig;        // per slot N_z Np: their column permutations
    cd *X;            // per slot N_z Np ldb: G(:, 1)(:, open), row-major per slice
    cfl *Y;           // per slot N_z Np ldb: correction sweep vectors
    int *next;        // job counter (zero at launch): workgroups take jobs in turn, so that
                      // a job with more refinement iterations holds up only its own slot
    const int *order; // the order in which the jobs are taken (likely slow ones first)
};

struct SolveShared {
    union {
        struct {
            cfl cL[NEGF_MAXNP4 * NEGF_NRH];   // one right-hand side, rows permuted: cL[sigma(i) 16 + b] = v(i, b)
            cd xL[64 * NEGF_NRH];             // 64 rows of X_m for the residual
        };
        cfl stage[NEGF_MAXNP4 * NEGF_NRH + 2 * 64 * NEGF_NRH];   // the sweep's: free outside the solves
    };
};

// out = g_m v for the 16 columns staged in cL, on the FP32 matrix cores.  Zg is the stored
// generator (g_m(:, c) = Zg(:, sigma_m(c))), so out(r, b) = sum_c Zg(r, c) cL(c, b).
// Wave w computes rows [32 w, 32 w + 32) as two 16-row tiles with A = Zg tile
// (A[i][k] at lane i + 16 k) and B = cL (B[k][b] at lane b + 16 k); lane b + 16 q,
// register r of tile t holds out(32 w + 16 t + 4 q + r, b).
// The generator is streamed from HBM in groups of 8 K-steps (32 columns) with the next
// group's loads issued before the current group's MFMAs (two register buffers), so the
// load latency hides behind the matrix cores; Np is a multiple of 32.  Complex products in
// the 3M form: p1 += Ar Br, p2 += Ai Bi, p3 += (Ar + Ai)(Br + Bi); re = p1 - p2,
// im = p3 - p1 - p2 (three MFMAs per tile and K-step instead of four).
__device__ __forceinline__ void gen_matvec(const cfl *Zg, const cfl *cL, int Np, negf_f4 (&ore)[2],
                                           negf_f4 (&oim)[2]) {
    const int tid = opaque_tid(), lane = tid & 63, w = tid >> 6;
    const int i = lane & 15, k = lane >> 4;
    const __amdgpu_buffer_rsrc_t zr = negf_rsrc(const_cast<cfl *>(Zg), (unsigned)(Np * Np) * 8u);
    const unsigned vo = (unsigned)(((w << 5) + i) + k * Np) * 8u;
    negf_f4 p1[2], p2[2], p3[2];
    for (int tt = 0; tt < 2; ++tt) {
        p1[tt] = negf_f4{0.f, 0.f, 0.f, 0.f};
        p2[tt] = negf_f4{0.f, 0.f, 0.f, 0.f};
        p3[tt] = negf_f4{0.f, 0.f, 0.f, 0.f};
    }
    constexpr int GK = 8;                       // K-steps (of 4 columns) per group
    const unsigned gstr = 4u * (unsigned)Np * 8u;
    auto load = [&](cfl (&A)[GK][2], int g) {
#pragma unroll
        for (int s = 0; s < GK; ++s) {
            const unsigned so = (unsigned)(g * GK + s) * gstr;
            A[s][0] = GjePrec<float>::load(zr, vo, so);
            A[s][1] = GjePrec<float>::load(zr, vo + 128u, so);
        }
    };
    auto compute = [&](const cfl (&A)[GK][2], int g) {
#pragma unroll
        for (int s = 0; s < GK; ++s) {
            const cfl b = cL[((g * GK + s) * 4 + k) * NEGF_NRH + i];
            const float bs = b.x + b.y;
#pragma unroll
            for (int tt = 0; tt < 2; ++tt) {
                p1[tt] = GjePrec<float>::mfma(A[s][tt].x, b.x, p1[tt]);
                p2[tt] = GjePrec<float>::mfma(A[s][tt].y, b.y, p2[tt]);
                p3[tt] = GjePrec<float>::mfma(A[s][tt].x + A[s][tt].y, bs, p3[tt]);
            }
        }
    };
    const int ng = Np / (4 * GK);
    cfl A0[GK][2], A1[GK][2];
    load(A0, 0);
    for (int g = 0; g < ng; g += 2) {
        if (g + 1 < ng) load(A1, g + 1);
        __builtin_amdgcn_sched_barrier(0);      // keep the prefetch ahead of the MFMAs
        compute(A0, g);
        if (g + 1 < ng) {
            if (g + 2 < ng) load(A0, g + 2);
            __builtin_amdgcn_sched_barrier(0);
            compute(A1, g + 1);
        }
    }
#pragma unroll
    for (int tt = 0; tt < 2; ++tt) {
        ore[tt] = p1[tt] - p2[tt];
        oim[tt] = p3[tt] - p1[tt] - p2[tt];
    }
}

// Correction right-hand side of slice m (1-based), scattered into cL by sigma_m:
//     c_m = E_1 delta_{m,1} - X_{m-1} - D_m X_m - X_{m+1} - y_{m+1}     (FP64, rounded)
// Row i belongs to the lane pair t = 2 i, 2 i + 1; lane h = t & 1 takes the columns
// j = h, h + 2, ... of D_m (each D entry is built once) for all NB right-hand sides, and
// the two partial sums meet through a DPP lane swap.  The slice's V-hat window is read
// through vs (LDS when staged) with the basis keys ckey, four columns per round.
template <int NB>
__device__ __forceinline__ void resid_scatter_nb(const SysView &sv, const JobView &jv, const int m, const cd *X,
                                                 const cfl *Y, const int ldb, const int nR, const int *Rc,
                                                 const int *sg, const cd *vs, const int *ckey, SolveShared &so) {
    const int Np = sv.Np, N = sv.N, Nz = sv.Nz;
    const int t = opaque_tid(), i = t >> 1, h = t & 1;
    const int slice = m - 1;
    const bool rin = i < N;
    const int rb = rin ? ckey[i] + sv.NGX + (2 * sv.NGX + 1) * sv.NGY : 0;
    const double kin = rin ? jv.kinh[i] : 0.0;
    cd dg = cmk(0.0, 0.0);
    if (rin) {
        if (slice == 0) dg = cadd(dg, jv.sigL[i]);
        if (slice == Nz - 1) dg = cadd(dg, jv.sigR[i]);
    }
    const cd *xm = X + (size_t)slice * Np * ldb;
    cd acc[NB];
#pragma unroll
    for (int b = 0; b < NB; ++b) acc[b] = cmk(0.0, 0.0);
    for (int j0 = 0; j0 < N; j0 += 64) {
        for (int e = t; e < 64 * NEGF_NRH; e += blockDim.x) {
            const int row = j0 + (e >> 4), b = e & 15;
            so.xL[e] = (row < N && b < ldb) ? xm[(size_t)row * ldb + b] : cmk(0.0, 0.0);
        }
        __syncthreads();
        if (rin) {
            const int jn = min(64, N - j0);
            // four columns per round: their window gathers are issued together
            for (int jj0 = h; jj0 < jn; jj0 += 8) {
                cd d[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int jj = jj0 + 2 * u, col = j0 + jj;
                    d[u] = jj < jn ? d_entry(vs[rb - ckey[col]], i == col, kin, jv.z, sv.scale, dg) : cmk(0.0, 0.0);
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int jj = min(jj0 + 2 * u, 63);
                    const cd *xr = so.xL + jj * NEGF_NRH;
#pragma unroll
                    for (int b = 0; b < NB; ++b) acc[b] = cfma(d[u], xr[b], acc[b]);
                }
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int b = 0; b < NB; ++b) acc[b] = cadd(acc[b], lane_swap(acc[b]));
    if (i < Np) {
        const int si = sg[i];
#pragma unroll
        for (int bb = 0; bb < NEGF_NRH / 2; ++bb) {
            const int b = 2 * bb + h;                 // lane h writes the columns b % 2 == h
            cfl v = cmk(0.0f, 0.0f);
            if (b < nR && rin) {
                cd ax = cmk(0.0, 0.0);
#pragma unroll
                for (int q = 0; q < NB; ++q)
                    if (q == b) ax = acc[q];
                cd r = cmk(-ax.x, -ax.y);
                if (m == 1 && i == Rc[b]) r.x += 1.0;
                if (m > 1) r = csub(r, X[((size_t)(slice - 1) * Np + i) * ldb + b]);
                if (m < Nz) {
                    r = csub(r, X[((size_t)(slice + 1) * Np + i) * ldb + b]);
                    r = csub(r, to_cd(Y[((size_t)(slice + 1) * Np + i) * ldb + b]));
                }
                v = from_cd<cfl>(r);
            }
            so.cL[si * NEGF_NRH + b] = v;
        }
    }
}

__device__ __forceinline__ void resid_scatter(const SysView &sv, const JobView &jv, const int m, const cd *X,
                                              const cfl *Y, const int ldb, const int nR, const int *Rc,
                                              const int *sg, const cd *vs, const int *ckey, SolveShared &so) {
    if (nR <= 2) resid_scatter_nb<2>(sv, jv, m, X, Y, ldb, nR, Rc, sg, vs, ckey, so);
    else if (nR <= 4) resid_scatter_nb<4>(sv, jv, m, X, Y, ldb, nR, Rc, sg, vs, ckey, so);
    else if (nR <= 6) resid_scatter_nb<6>(sv, jv, m, X, Y, ldb, nR, Rc, sg, vs, ckey, so);
    else if (nR <= 8) resid_scatter_nb<8>(sv, jv, m, X, Y, ldb, nR, Rc, sg, vs, ckey, so);
    else if (nR <= 10) resid_scatter_nb<10>(sv, jv, m, X, Y, ldb, nR, Rc, sg, vs, ckey, so);
    else if (nR <= 12) resid_scatter_nb<12>(sv, jv, m, X, Y, ldb, nR, Rc, sg, vs, ckey, so);
    else resid_scatter_nb<16>(sv, jv, m, X, Y, ldb, nR, Rc, sg, vs, ckey, so);
}

// cL[i 16 + b] = coef * V(i, b) (b < nR, else 0) for a row-major Np x ldb array V (the
// mixed sweep's column map is the identity).  blockDim.x == 2 Np, so every thread takes
// exactly 8 entries; their loads are all issued before the first store.
template <class C>
__device__ __forceinline__ void scatter_rhs(cfl *cL, const C *V, int Np, int ldb, int nR, float coef) {
    constexpr int PER = NEGF_NRH / 2;
    const int tid = opaque_tid();
    C x[PER];
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        const int e = tid + u * blockDim.x, i = e >> 4, b = e & 15;
        x[u] = b < nR ? V[(size_t)i * ldb + b] : C{};
    }
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        const int e = tid + u * blockDim.x, b = e & 15;
        const cd xv = to_cd(x[u]);
        cL[e] = b < nR ? cmk((float)xv.x * coef, (float)xv.y * coef) : cmk(0.0f, 0.0f);
    }
}

template <int MAXT>
__global__ __launch_bounds__(MAXT) void k_transmission_mixed(MixedArgs ma) {
    __shared__ BgjeShared<> s;
    __shared__ SolveShared so;
    __shared__ int ckey[NEGF_MAXNP4];
    static_assert(sizeof(so.stage) >= (BgjeShared<>::TSZ + 2 * BgjeShared<>::WLSZ) * sizeof(cfl),
                  "second staging buffer");
    const TransArgs &a = ma.a;
    const SysView &sv = a.sv;
    const int Np = sv.Np, Nz = sv.Nz, ldb = ma.ldb;
    const size_t NN = (size_t)Np * Np;
    const int t = opaque_tid(), lane = t & 63, w = t >> 6;
    const int jj = lane & 15, q = lane >> 4;
    // basis keys for the V-hat window index of the block build and the residual, and the
    // slice's window staged in LDS when it fits (the Gauss-Jordan's staging area, free
    // outside the inversion); otherwise read from HBM
    const int nvh = (2 * sv.NGX + 1) * (2 * sv.NGY + 1);
    const bool vh_lds = nvh <= BgjeShared<>::VHCAP;
    for (int c = t; c < Np; c += blockDim.x) ckey[c] = c < sv.N ? sv.nx[c] + (2 * sv.NGX + 1) * sv.ny[c] : 0;
    bgje_roles(s);                               // ends with a barrier
    auto stage_vh = [&](int slice) {             // the caller synchronizes
        if (vh_lds)
            for (int e = t; e < nvh; e += blockDim.x) s.vh[e] = sv.vrec[(size_t)slice * nvh + e];
    };
    auto vsrc = [&](int slice) -> const cd * { return vh_lds ? s.vh : sv.vrec + (size_t)slice * nvh; };
    cfl *hist = ma.hist + (size_t)blockIdx.x * Nz * NN;
    int *hsig = ma.hsig + (size_t)blockIdx.x * Nz * Np;
    cd *X = ma.X + (size_t)blockIdx.x * Nz * Np * ldb;
    cfl *Y = ma.Y + (size_t)blockIdx.x * Nz * Np * ldb;
    // Jobs are taken in turn from the counter.  The loop has one back edge, through the
    // fetch; the job body leaves early by returning (a `continue` back to the fetch has
    // compiled into a loop that skipped the atomic); the job index, the open-channel
    // counts and the growth flag are wave-uniform (readfirstlane), since the branches on
    // them enclose barriers.
    auto fetch = [&]() {
        if (t == 0) s.job = atomicAdd(ma.next, 1);
        __syncthreads();
        const int jb = __builtin_amdgcn_readfirstlane(s.job);
        __syncthreads();
        return jb;
    };
    auto run_job = [&](const int job) {
        const int nL = __builtin_amdgcn_readfirstlane(a.nL[job]), nR = __builtin_amdgcn_readfirstlane(a.nR[job]);
        const int off = __builtin_amdgcn_readfirstlane(a.open_off[job]);
        if (nL == 0 || nR == 0) {             // no open channel: tau = 0 exactly
            if (t == 0) {
                a.tau[job] = 0.0;
                a.status[job] = 0;
            }
            return;
        }
        JobView jv;
        jv.z = a.z[job];
        jv.kinh = a.kinh + (size_t)a.kidx[job] * Np;
        jv.sigL = a.sigL + (size_t)job * Np;
        jv.sigR = a.sigR + (size_t)job * Np;
        const int *Rc = a.Rcol + off;
        bool grew = false;
        // 1. FP32 sweep, every generator kept: g_m = (D_m - g_{m+1})^{-1} by the block
        // Gauss-Jordan (no column map: hsig is the identity)
        for (int i = t; i < Np; i += blockDim.x) s.sigma[i] = i;
        __syncthreads();
        for (int m = Nz; m >= 1 && !grew; --m) {
            cfl *M = hist + (size_t)(m - 1) * NN;
            stage_vh(m - 1);
            __syncthreads();
            build_block_keys(M, sv, jv, m - 1, m == Nz ? nullptr : hist + (size_t)m * NN, vsrc(m - 1), ckey);
            __syncthreads();
            if (bgje_invert64(M, Np, s, so.stage)) grew = true;    // the solves' staging area is free here
            for (int i = t; i < Np; i += blockDim.x) hsig[(size_t)(m - 1) * Np + i] = i;
        }
        __syncthreads();
        if (grew) {                           // growth check failed: the host redoes it in FP64
            if (t == 0) {
                a.tau[job] = 0.0;
                a.status[job] = -1;
            }
            __syncthreads();
            return;
        }
        // 2. first solve: X_1 = g_1(:, open) = Zg_1(:, sigma_1(open)), X_m = -g_m X_{m-1}
        for (int e = t; e < Np * ldb; e += blockDim.x) {
            const int i = e / ldb, b = e % ldb;
            X[e] = b < nR ? to_cd(hist[i + (size_t)hsig[Rc[b]] * Np]) : cmk(0.0, 0.0);
        }
        __syncthreads();
        for (int m = 2; m <= Nz; ++m) {
            scatter_rhs(so.cL, X + (size_t)(m - 2) * Np * ldb, Np, ldb, nR, -1.0f);
            __syncthreads();
            negf_f4 ore[2], oim[2];
            gen_matvec(hist + (size_t)(m - 1) * NN, so.cL, Np, ore, oim);
            const int tid_e = opaque_tid(), w = tid_e >> 6, jj = tid_e & 15, q = (tid_e >> 4) & 3;
            if (jj < nR) {
#pragma unroll
                for (int tt = 0; tt < 2; ++tt)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int row = (w << 5) + 16 * tt + 4 * q + r;
                        X[((size_t)(m - 1) * Np + row) * ldb + jj] = cmk((double)ore[tt][r], (double)oim[tt][r]);
                    }
            }
            __syncthreads();
        }
        // 3. iterative refinement with FP64 residuals, until the error left in X_{N_z} on the
        // open rows, estimated from the last correction and the observed contraction, is
        // below ma.tol of |X_{N_z}| there (at most ma.nref iterations; a job that does not get
        // there reports status -1)
        bool conv = false;
        double dprev = -1.0;
        for (int it = 0; it < ma.nref && !conv; ++it) {
            for (int m = Nz; m >= 1; --m) {           // y_m = g_m (r_m - y_{m+1})
                stage_vh(m - 1);
                __syncthreads();
                resid_scatter(sv, jv, m, X, Y, ldb, nR, Rc, hsig + (size_t)(m - 1) * Np, vsrc(m - 1), ckey, so);
                __syncthreads();
                negf_f4 ore[2], oim[2];
                gen_matvec(hist + (size_t)(m - 1) * NN, so.cL, Np, ore, oim);
                const int tid_e = opaque_tid(), w = tid_e >> 6, jj = tid_e & 15, q = (tid_e >> 4) & 3;
                if (jj < nR) {
#pragma unroll
                    for (int tt = 0; tt < 2; ++tt)
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            const int row = (w << 5) + 16 * tt + 4 * q + r;
                            Y[((size_t)(m - 1) * Np + row) * ldb + jj] = cmk(ore[tt][r], oim[tt][r]);
                        }
                }
                __syncthreads();
            }
            for (int e = t; e < Np * ldb; e += blockDim.x) X[e] = cadd(X[e], to_cd(Y[e]));   // d_1 = y_1
            __syncthreads();
            for (int m = 2; m <= Nz; ++m) {           // d_m = y_m - g_m d_{m-1}, X_m += d_m
                scatter_rhs(so.cL, Y + (size_t)(m - 2) * Np * ldb, Np, ldb, nR, 1.0f);
                __syncthreads();
                negf_f4 ore[2], oim[2];
                gen_matvec(hist + (size_t)(m - 1) * NN, so.cL, Np, ore, oim);
                const int tid_e = opaque_tid(), w = tid_e >> 6, jj = tid_e & 15, q = (tid_e >> 4) & 3;
                if (jj < nR) {
#pragma unroll
                    for (int tt = 0; tt < 2; ++tt)
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            const int row = (w << 5) + 16 * tt + 4 * q + r;
                            const size_t e = ((size_t)(m - 1) * Np + row) * ldb + jj;
                            const cfl y = Y[e];
                            const cfl d = cmk(y.x - ore[tt][r], y.y - oim[tt][r]);
                            Y[e] = d;
                            X[e] = cadd(X[e], to_cd(d));
                        }
                }
                __syncthreads();
            }
            // convergence, on the transmission itself: with rho = max|d| / max|d_prev| the
            // observed contraction, the error left in X_{N_z} is about e = |d| rho / (1 - rho),
            // so T = s^2 sum gL gR |X|^2 is off by at most dT = s^2 sum gL gR (2 |X| e + e^2)
            // (open rows and columns).  Stop when dT <= tol.  The first iteration (no rho yet)
            // stops only when the correction itself moves T by less than 1e-3 tol.
            double dmx = 0.0, sxd = 0.0, sdd = 0.0;
            for (int e = t; e < nL * nR; e += blockDim.x) {
                const int ai = e % nL, bi = e / nL;
                const size_t o = ((size_t)(Nz - 1) * Np + a.Lrow[off + ai]) * ldb + bi;
                const cfl d = Y[o];
                const cd x = X[o];
                const double ad = sqrt((double)d.x * d.x + (double)d.y * d.y), ax = sqrt(x.x * x.x + x.y * x.y);
                const double gg = a.gL[off + ai] * a.gR[off + bi];
                dmx = fmax(dmx, ad);
                sxd += gg * ax * ad;
                sdd += gg * ad * ad;
            }
            dmx = block_max2(dmx, s.red);
            sxd = block_sum2(sxd, s.red);
            sdd = block_sum2(sdd, s.red);
            const double s2 = sv.scale * sv.scale;
            const double rho = dprev > 0.0 ? dmx / dprev : 1.0;
            const double f = rho / (1.0 - rho);
            conv = rho < 0.5 ? s2 * (2.0 * f * sxd + f * f * sdd) <= ma.tol
                             : s2 * (2.0 * sxd + sdd) <= 1e-3 * ma.tol;
            conv = __builtin_amdgcn_readfirstlane((int)conv) != 0;
            dprev = dmx;
            // a contraction this slow will not reach the tolerance within the limit: the
            // job goes to the FP64 path at once
            if (!conv && it >= 3 && rho >= 0.9) break;
        }
        // 4. trace on X_{N_z} = G_{N_z,1}(:, open)
        double part = 0.0;
        for (int e = t; e < nL * nR; e += blockDim.x) {
            const int ai = e % nL, bi = e / nL;
            const cd v = X[((size_t)(Nz - 1) * Np + a.Lrow[off + ai]) * ldb + bi];
            part += a.gL[off + ai] * a.gR[off + bi] * (v.x * v.x + v.y * v.y);
        }
        const double tot = block_sum2(part, s.red);
        if (t == 0) {
            a.tau[job] = tot * sv.scale * sv.scale;
            a.status[job] = conv ? 0 : -1;
        }
        __syncthreads();
    };
    for (int c = fetch(); c < a.njobs; c = fetch()) run_job(__builtin_amdgcn_readfirstlane(ma.order[c]));
}

// D mode (diagonal blocks): backward generators stored per slot, then
//   G_11 = (D_1 - g_2)^{-1},  G_kk = g_k (G_{k-1,k-1} g_k + I)   (src/math_kernel.f90:94-109)
// and out(:, :, k, tgt) += wgt * G_kk for the N x N physical block.
struct DiagArgs {
    SysView sv;
    // uneven batches (MULTI): per-job device index into svs, per-workgroup job lists,
    // per-device output bases (N x N x nsel[d] per target), slice -> output slot maps
    // (smap[smap_off[d] + k], -1 = not kept) and the per-workgroup workspace stride
    const SysView *svs;
    const int *jdev;
    const int *sched_off, *sched_job;
    cd *const *outs;
    const int *nsel, *smap, *smap_off;
    size_t ws_stride;
    int njobs;
    const cd *z;
    const int *kidx;
    const double *kinh;
    const cd *sigL, *sigR;
    const cd *wgt;        // per job complex weight (includes 2 Delta^2)
    const int *target;    // per job output index (density: 0, LDOS: energy index)
    cd *ws;               // per slot (N_z + 2) Np^2
    cd *out;              // N*N*N_z per target, column-major ld N
    int *status;
    int phase, job_base;  // k_diagonal_split only
};

template <int MAXT, int PB, bool TWO, bool MULTI = false>
__global__ __launch_bounds__(MAXT) void k_diagonal(DiagArgs a) {
    __shared__ GjeSh<PB, TWO> s;
    const int Np = a.sv.Np;
    const size_t NN = (size_t)Np * Np;
    const size_t stride = MULTI ? a.ws_stride : (size_t)(a.sv.Nz + 2) * NN + 2 * (size_t)PB * Np;
    cd *G = a.ws + (size_t)blockIdx.x * stride;   // G[m] = g_{m+1}
    const int i0 = MULTI ? a.sched_off[blockIdx.x] : blockIdx.x;
    const int i1 = MULTI ? a.sched_off[blockIdx.x + 1] : a.njobs;
    const int di = MULTI ? 1 : (int)gridDim.x;
    for (int i = i0; i < i1; i += di) {
        const int job = MULTI ? a.sched_job[i] : i;
        const int dv = MULTI ? a.jdev[job] : 0;
        const SysView &sv = MULTI ? a.svs[dv] : a.sv;
        const int N = sv.N, Nz = sv.Nz;
        cd *scratch = G + (size_t)(Nz + 2) * NN;
        cd *X = G, *Y = G + (size_t)Nz * NN, *W = G + (size_t)(Nz + 1) * NN;
        JobView jv;
        jv.z = a.z[job];
        jv.kinh = a.kinh + (size_t)a.kidx[job] * Np;
        jv.sigL = a.sigL + (size_t)job * Np;
        jv.sigR = a.sigR + (size_t)job * Np;
        const cd wgt = a.wgt[job];
        const int nsel = MULTI ? a.nsel[dv] : Nz;
        cd *out = (MULTI ? a.outs[dv] : a.out) + (size_t)a.target[job] * N * N * nsel;
        auto slice_out = [&](int k) -> cd * {
            const int q = MULTI ? a.smap[a.smap_off[dv] + k] : k;
            return q < 0 ? nullptr : out + (size_t)q * N * N;
        };
        int bad = 0;
        // backward generators g_{N_z} .. g_2 stored unpermuted; the last pass leaves
        // G_11 = (D_1 - g_2)^{-1} in X
        for (int m = Nz; m >= 1; --m) {
            build_block(W, sv, jv, m - 1, m == Nz ? nullptr : G + (size_t)m * NN, nullptr);
            __syncthreads();
            if (gje_any<PB, TWO>(W, Np, s, scratch) && !bad) bad = m;
            copy_cols(m >= 2 ? G + (size_t)(m - 1) * NN : X, W, s.sigma, Np);
            __syncthreads();
        }
        if (cd *o0 = slice_out(0)) {
            for (int col = 0; col < N; ++col)
                for (int r = threadIdx.x; r < N; r += blockDim.x) {
                    const cd v = cmul(wgt, X[r + (size_t)col * Np]);
                    double *o = reinterpret_cast<double *>(o0 + r + (size_t)col * N);
                    atomicAdd(o, v.x);
                    atomicAdd(o + 1, v.y);
                }
        }
        __syncthreads();
        // G_kk = g_k (G_{k-1,k-1} g_k + I)  (two products per slice, one call site)
        for (int step = 2; step < 2 * Nz; ++step) {
            const int k = step >> 1;          // 0-based slice, generator g_{k+1} at G[k]
            const bool second = step & 1;
            const cd *gk = G + (size_t)k * NN;
            cd *ok = slice_out(k);
            gemm_nn(second ? gk : X, second ? Y : gk, nullptr, Np, [&](int r, int c, cd v) {
                if (!second) {
                    if (r == c) v.x += 1.0;
                    Y[r + (size_t)c * Np] = v;
                } else {
                    X[r + (size_t)c * Np] = v;
                    if (ok && r < N && c < N) {
                        const cd t = cmul(wgt, v);
                        double *o = reinterpret_cast<double *>(ok + r + (size_t)c * N);
                        atomicAdd(o, t.x);
                        atomicAdd(o + 1, t.y);
                    }
                }
            });
            __syncthreads();
        }
        if (threadIdx.x == 0) a.status[job] = bad;
        __syncthreads();
    }
}

// D mode cut in two (same blocks, same operation count).  With the forward generators
// from the left lead, f_1 = D_1^{-1}, f_k = (D_k - f_{k-1})^{-1}, beside the backward ones
// g_k above, the diagonal blocks at a cut m = N_z / 2 are
//   G_mm = (D_m - f_{m-1} - g_{m+1})^{-1},   G_{m+1,m+1} = (D_{m+1} - f_m - g_{m+2})^{-1},
// and the recursion runs outwards from the cut on either side:
//   G_kk = f_k (G_{k+1,k+1} f_k + I)  (k < m),   G_kk = g_k (G_{k-1,k-1} g_k + I)  (k > m + 1).
// Two workgroups per job, one per side.  Phase 0 stores the generators of its side
// (f_1..f_m or g_{N_z}..g_{m+1}: N_z / 2 inversions each); phase 1, a second launch, makes
// the cut block of its side from both sides' generators and runs the products.  A job's
// critical path is half of k_diagonal's, and a batch of fewer jobs than CUs fills twice
// as many.  Per job the workspace holds G[k - 1] = f_k (k <= m) or g_k (k > m), then per
// side X, Y, W (Np^2 each) and the Gauss-Jordan scratch.  Needs N_z >= 2.
template <int MAXT, int PB, bool TWO>
__global__ __launch_bounds__(MAXT) void k_diagonal_split(DiagArgs a) {
    __shared__ GjeSh<PB, TWO> s;
    const SysView &sv = a.sv;
    const int Np = sv.Np, N = sv.N, Nz = sv.Nz, m = Nz / 2;
    const size_t NN = (size_t)Np * Np;
    const size_t per_job = (size_t)(Nz + 6) * NN + 4 * (size_t)PB * Np;
    const int jl = blockIdx.x >> 1, side = blockIdx.x & 1;
    const int job = a.job_base + jl;
    cd *G = a.ws + (size_t)jl * per_job;
    cd *X = G + (size_t)(Nz + 3 * side) * NN, *Y = X + NN, *W = Y + NN;
    cd *scratch = G + (size_t)(Nz + 6) * NN + (size_t)side * 2 * PB * Np;
    JobView jv;
    jv.z = a.z[job];
    jv.kinh = a.kinh + (size_t)a.kidx[job] * Np;
    jv.sigL = a.sigL + (size_t)job * Np;
    jv.sigR = a.sigR + (size_t)job * Np;
    int *status = a.status + 2 * (size_t)job + side;
    if (a.phase == 0) {
        int bad = 0;
        if (side == 0) {
            for (int k = 1; k <= m; ++k) {
                build_block(W, sv, jv, k - 1, k == 1 ? nullptr : G + (size_t)(k - 2) * NN, nullptr);
                __syncthreads();
                if (gje_any<PB, TWO>(W, Np, s, scratch) && !bad) bad = k;
                copy_cols(G + (size_t)(k - 1) * NN, W, s.sigma, Np);
                __syncthreads();
            }
        } else {
            for (int k = Nz; k > m; --k) {
                build_block(W, sv, jv, k - 1, k == Nz ? nullptr : G + (size_t)k * NN, nullptr);
                __syncthreads();
                if (gje_any<PB, TWO>(W, Np, s, scratch) && !bad) bad = k;
                copy_cols(G + (size_t)(k - 1) * NN, W, s.sigma, Np);
                __syncthreads();
            }
        }
        if (threadIdx.x == 0) *status = bad;
        return;
    }
    const cd wgt = a.wgt[job];
    cd *out = a.out + (size_t)a.target[job] * N * N * Nz;
    // the cut block of this side: slice c (1-based), neighbours f_{c-1} and g_{c+1}
    const int c = side == 0 ? m : m + 1;
    const cd *fl = c >= 2 ? G + (size_t)(c - 2) * NN : nullptr;
    const cd *gr = c + 1 <= Nz ? G + (size_t)c * NN : nullptr;
    for (size_t e = threadIdx.x; e < NN; e += blockDim.x)
        Y[e] = fl && gr ? cadd(fl[e], gr[e]) : (fl ? fl[e] : gr[e]);
    __syncthreads();
    build_block(W, sv, jv, c - 1, Y, nullptr);
    __syncthreads();
    const int zp = gje_any<PB, TWO>(W, Np, s, scratch);
    copy_cols(X, W, s.sigma, Np);
    __syncthreads();
    auto accumulate = [&](cd *ok, int r, int col, cd v) {
        const cd t = cmul(wgt, v);
        double *o = reinterpret_cast<double *>(ok + r + (size_t)col * N);
        atomicAdd(o, t.x);
        atomicAdd(o + 1, t.y);
    };
    {
        cd *o0 = out + (size_t)(c - 1) * N * N;
        for (int col = 0; col < N; ++col)
            for (int r = threadIdx.x; r < N; r += blockDim.x) accumulate(o0, r, col, X[r + (size_t)col * Np]);
    }
    __syncthreads();
    // outwards from the cut: k = c - 1 .. 1 on the left, c + 1 .. N_z on the right
    const int nsteps = side == 0 ? c - 1 : Nz - c;
    for (int step = 0; step < 2 * nsteps; ++step) {
        const int k = side == 0 ? c - 1 - (step >> 1) : c + 1 + (step >> 1);   // 1-based
        const bool second = step & 1;
        const cd *gk = G + (size_t)(k - 1) * NN;
        cd *ok = out + (size_t)(k - 1) * N * N;
        gemm_nn(second ? gk : X, second ? Y : gk, nullptr, Np, [&](int r, int col, cd v) {
            if (!second) {
                if (r == col) v.x += 1.0;
                Y[r + (size_t)col * Np] = v;
            } else {
                X[r + (size_t)col * Np] = v;
                if (r < N && col < N) accumulate(ok, r, col, v);
            }
        });
        __syncthreads();
    }
    if (threadIdx.x == 0 && zp && *status == 0) *status = c;
}

// Vo(:, q) = add(:, q) - M(:, sig) Vi(:, q) for the nq columns q (add == nullptr: -M Vi);
// V blocks are Np x nq with column stride Np.  FP64 matrix cores: every wave takes 16-row
// tiles of Vo and 16 columns at a time; A(i, k) = M(r0 + i, sig(k0 + k)) in lane i + 16 k
// (16 consecutive rows of one column per load), B(k, j) = Vi(k0 + k, q0 + j) in lane
// j + 16 k (zero for q0 + j >= nq).  D(i, j) comes back in lane j + 16 (i % 4), register
// i / 4.  M is read once per product, not once per column.  Np % 16 == 0.
__device__ __forceinline__ void colprod_mfma(cd *__restrict__ Vo, const cd *__restrict__ M, const int *sig,
                                             const cd *__restrict__ Vi, const cd *add, int nq, int Np) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    for (int q0 = 0; q0 < nq; q0 += 16) {
        const bool bcol = q0 + li < nq;
        const cd *vb = Vi + (size_t)(q0 + li) * Np + lk;
        for (int r0 = 16 * w; r0 < Np; r0 += 16 * nw) {
            const cd *ma = M + (r0 + li);
            negf_d4 re0 = {0.0, 0.0, 0.0, 0.0}, re1 = re0, im0 = re0, im1 = re0;
            for (int k0 = 0; k0 < Np; k0 += 16) {
                cd av[4], bv[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int kk = k0 + 4 * u + lk;
                    av[u] = ma[(size_t)(sig ? sig[kk] : kk) * Np];
                    bv[u] = bcol ? vb[k0 + 4 * u] : cmk(0.0, 0.0);
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    re0 = __builtin_amdgcn_mfma_f64_16x16x4f64(av[u].x, bv[u].x, re0, 0, 0, 0);
                    re1 = __builtin_amdgcn_mfma_f64_16x16x4f64(-av[u].y, bv[u].y, re1, 0, 0, 0);
                    im0 = __builtin_amdgcn_mfma_f64_16x16x4f64(av[u].x, bv[u].y, im0, 0, 0, 0);
                    im1 = __builtin_amdgcn_mfma_f64_16x16x4f64(av[u].y, bv[u].x, im1, 0, 0, 0);
                }
            }
            if (bcol) {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const size_t e = (size_t)(q0 + li) * Np + r0 + 4 * r + lk;
                    const cd p = cmk(re0[r] + re1[r], im0[r] + im1[r]);
                    Vo[e] = add ? csub(add[e], p) : cmk(-p.x, -p.y);
                }
            }
        }
    }
}

// Transmission with two workgroups per job (same blocks; for calls with fewer jobs than
// half the workgroups the GPU holds).  With the cut at m = N_z / 2 the left workgroup
// sweeps the forward generators f_1..f_m and carries the left parts' first columns on the
// open columns of G_{N_z,1}, L_1 = f_1(:, open), L_k = -f_k L_{k-1}; the right one sweeps
// g_{N_z}..g_{m+1} and carries R = E^T g_{N_z} ... g_{m+1} on the open rows, as
// k_transmission does.  Then (second launch, one workgroup per job)
//   G_{m,1}(:, open) = (I - f_m g_{m+1})^{-1} L_m,   G_{N_z,1}(open, open) = +-R G_{m,1},
// and tau as in k_transmission (the sign drops out of |.|^2).  Chains and generators
// alternate between two buffers by step parity, so phase 1 finds L_m and R from m and
// N_z alone; each side leaves its last generator unpermuted for the product.
template <int MAXT, int PB, bool TWO, int MAXNP = NEGF_MAXNP4, int TCH = NEGF_TCH>
__global__ __launch_bounds__(MAXT) NEGF_SMALL_WPE(MAXT) void k_transmission_split(TransArgs a) {
    __shared__ GjeSh<PB, TWO, MAXNP, TCH> s;
    __shared__ RowProdShared rs;
    const SysView &sv = a.sv;
    const int Np = sv.Np, Nz = sv.Nz, m = Nz / 2, ldX = a.ldX;
    const size_t NN = (size_t)Np * Np, XX = (size_t)ldX * Np, SCR = 2 * (size_t)PB * Np;
    const size_t side_sz = 2 * NN + 2 * XX + SCR;
    const size_t per_job = 2 * side_sz + 3 * NN + XX + SCR;
    const int jl = a.phase == 0 ? (int)(blockIdx.x >> 1) : (int)blockIdx.x;
    const int side = a.phase == 0 ? (int)(blockIdx.x & 1) : 0;
    const int job = a.job_base + jl;
    cd *J0 = a.ws + (size_t)jl * per_job;
    cd *F = J0 + 2 * side_sz, *Gb = F + NN, *Wm = Gb + NN, *Gm = Wm + NN, *scr2 = Gm + XX;
    const int nL = a.nL[job], nR = a.nR[job], off = a.open_off[job];
    if (nL == 0 || nR == 0) {                 // no open channel: tau = 0 exactly
        if (a.phase == 1 && threadIdx.x == 0) {
            a.tau[job] = 0.0;
            a.status[job] = 0;
        }
        return;
    }
    JobView jv;
    jv.z = a.z[job];
    jv.kinh = a.kinh + (size_t)a.kidx[job] * Np;
    jv.sigL = a.sigL + (size_t)job * Np;
    jv.sigR = a.sigR + (size_t)job * Np;
    if (a.phase == 0) {
        cd *sb = J0 + (size_t)side * side_sz;
        cd *gen[2] = {sb, sb + NN};
        cd *buf[2] = {sb + 2 * NN, sb + 2 * NN + XX};
        cd *scratch = sb + 2 * NN + 2 * XX;
        int bad = 0;
        if (side == 0) {
            for (int k = 1; k <= m; ++k) {
                cd *M = gen[k & 1];
                build_block(M, sv, jv, k - 1, k == 1 ? nullptr : gen[(k - 1) & 1], s.sigma);
                __syncthreads();
                if (gje_any<PB, TWO>(M, Np, s, scratch) && !bad) bad = k;
                if (k == 1) {
                    for (int e = threadIdx.x; e < nR * Np; e += blockDim.x) {
                        const int q = e / Np, r = e % Np;
                        buf[0][e] = M[r + (size_t)s.sigma[a.Rcol[off + q]] * Np];
                    }
                } else {
                    colprod_mfma(buf[(k - 1) & 1], M, s.sigma, buf[(k - 2) & 1], nullptr, nR, Np);
                }
                __syncthreads();
            }
            copy_cols(F, gen[m & 1], s.sigma, Np);
        } else {
            for (int k = Nz; k > m; --k) {
                cd *M = gen[k & 1];
                build_block(M, sv, jv, k - 1, k == Nz ? nullptr : gen[(k + 1) & 1], s.sigma);
                __syncthreads();
                if (gje_any<PB, TWO>(M, Np, s, scratch) && !bad) bad = k;
                if (k == Nz) {
                    for (int e = threadIdx.x; e < nL * Np; e += blockDim.x) {
                        const int ai = e % nL, c = e / nL;
                        buf[0][ai + (size_t)c * ldX] = M[a.Lrow[off + ai] + (size_t)s.sigma[c] * Np];
                    }
                    __syncthreads();
                } else {
                    cd *Rn = buf[(Nz - k) & 1], *Ro = buf[(Nz - k - 1) & 1];
                    if (Np >= NEGF_RP_MFMA_LO && Np <= NEGF_RP_MFMA_HI) row_product_mfma(Rn, Ro, nL, ldX, M, s.sigma, Np);
                    else row_product(Rn, Ro, nL, ldX, M, s.sigma, Np, rs);
                }
            }
            copy_cols(Gb, gen[(m + 1) & 1], s.sigma, Np);
        }
        __syncthreads();
        if (threadIdx.x == 0) a.status2[2 * (size_t)job + side] = bad;
        return;
    }
    // phase 1: W = I - f_m g_{m+1} on the physical block (identity on the padding)
    const int N = sv.N;
    gemm_nn(F, Gb, nullptr, Np, [&](int r, int c, cd v) {
        Wm[r + (size_t)c * Np] = (r < N && c < N) ? cmk((r == c ? 1.0 : 0.0) - v.x, -v.y) : cmk(r == c ? 1.0 : 0.0, 0.0);
    });
    __syncthreads();
    const int zp = gje_any<PB, TWO>(Wm, Np, s, scr2);
    const cd *Lm = J0 + 2 * NN + (size_t)((m - 1) & 1) * XX;                       // left side
    const cd *Rm = J0 + side_sz + 2 * NN + (size_t)((Nz - m - 1) & 1) * XX;         // right side
    colprod_mfma(Gm, Wm, s.sigma, Lm, nullptr, nR, Np);                             // -G_{m,1}(:, open)
    __syncthreads();
    double part = 0.0;
    for (int e = threadIdx.x; e < nL * nR; e += blockDim.x) {
        const int ai = e % nL, bi = e / nL;
        const cd *gc = Gm + (size_t)bi * Np;
        cd acc = cmk(0.0, 0.0);
        for (int k = 0; k < Np; ++k) acc = cfma(Rm[ai + (size_t)k * ldX], gc[k], acc);
        part += a.gL[off + ai] * a.gR[off + bi] * (acc.x * acc.x + acc.y * acc.y);
    }
    const double tot = block_sum2(part, s.red);
    if (threadIdx.x == 0) {
        const int b0 = a.status2[2 * (size_t)job], b1 = a.status2[2 * (size_t)job + 1];
        a.tau[job] = tot * sv.scale * sv.scale;
        a.status[job] = b0 ? b0 : (b1 ? b1 : (zp ? m : 0));
    }
}

// Bias-window spectral blocks for the non-equilibrium density (extension, no reference
// counterpart; the reference rejects VDS != 0, src/main.f90:129-132).  Per job the
// backward generators are stored as in D mode, then the first block column on the open
// channels of the left lead, X_1 = G_11(:, open), X_k = -g_k X_{k-1} (the F recursion
// of src/math_kernel.f90:120-128), gives
//     out_k += wgt * sum_a gamma^L_a X_k(:, a) X_k(:, a)^dagger      (r, c < N),
// i.e. the diagonal blocks of G Gamma_L G^dagger with Gamma_L on slice 1.
struct NeqArgs {
    SysView sv;
    int njobs;
    const cd *z;
    const int *kidx;
    const double *kinh;
    const cd *sigL, *sigR;
    const cd *wgt;          // per job: window weight * k weight * (2 Delta^2)^2
    const int *open_off, *nL, *Lcol;   // open channels of the left lead (flat lists)
    const double *gL;
    int nLmax;
    cd *ws;                 // per slot (N_z + 1) Np^2 + 2 nLmax Np + 2 PB Np
    cd *out;                // N*N*N_z, column-major ld N
    int *status;
    int phase, job_base;    // k_neq_split only
};

template <int MAXT, int PB, bool TWO>
__global__ __launch_bounds__(MAXT) void k_neq(NeqArgs a) {
    __shared__ GjeSh<PB, TWO> s;
    const SysView &sv = a.sv;
    const int Np = sv.Np, N = sv.N, Nz = sv.Nz, ldV = a.nLmax;
    const size_t NN = (size_t)Np * Np;
    cd *G = a.ws + (size_t)blockIdx.x * ((Nz + 1) * NN + 2 * (size_t)ldV * Np + 2 * (size_t)PB * Np);
    cd *W = G + (size_t)Nz * NN;                  // G[0] = G_11, G[m - 1] = g_m (m >= 2)
    cd *V0 = W + NN, *V1 = V0 + (size_t)ldV * Np;
    cd *scratch = V1 + (size_t)ldV * Np;
    for (int job = blockIdx.x; job < a.njobs; job += gridDim.x) {
        const int nL = a.nL[job], off = a.open_off[job];
        if (nL == 0) {                            // no open channel of the left lead
            if (threadIdx.x == 0) a.status[job] = 0;
            continue;
        }
        JobView jv;
        jv.z = a.z[job];
        jv.kinh = a.kinh + (size_t)a.kidx[job] * Np;
        jv.sigL = a.sigL + (size_t)job * Np;
        jv.sigR = a.sigR + (size_t)job * Np;
        const cd wgt = a.wgt[job];
        int bad = 0;
        for (int m = Nz; m >= 1; --m) {
            build_block(W, sv, jv, m - 1, m == Nz ? nullptr : G + (size_t)m * NN, nullptr);
            __syncthreads();
            if (gje_any<PB, TWO>(W, Np, s, scratch) && !bad) bad = m;
            copy_cols(G + (size_t)(m - 1) * NN, W, s.sigma, Np);
            __syncthreads();
        }
        auto accumulate = [&](int k, const cd *V) {
            cd *ok = a.out + (size_t)k * N * N;
            for (int e = threadIdx.x; e < N * N; e += blockDim.x) {
                const int r = e % N, c = e / N;
                cd v = cmk(0.0, 0.0);
                for (int q = 0; q < nL; ++q) {
                    const cd x = V[(size_t)q * Np + r], y = V[(size_t)q * Np + c];
                    const double gq = a.gL[off + q];
                    v.x += gq * (x.x * y.x + x.y * y.y);       // x conj(y)
                    v.y += gq * (x.y * y.x - x.x * y.y);
                }
                const cd t = cmul(wgt, v);
                double *o = reinterpret_cast<double *>(ok + e);
                atomicAdd(o, t.x);
                atomicAdd(o + 1, t.y);
            }
        };
        for (int e = threadIdx.x; e < nL * Np; e += blockDim.x) {
            const int q = e / Np, r = e % Np;
            V0[e] = G[r + (size_t)a.Lcol[off + q] * Np];
        }
        __syncthreads();
        accumulate(0, V0);
        for (int k = 2; k <= Nz; ++k) {
            colprod_mfma(V1, G + (size_t)(k - 1) * NN, nullptr, V0, nullptr, nL, Np);
            __syncthreads();
            accumulate(k - 1, V1);
            __syncthreads();
            cd *t = V0;
            V0 = V1;
            V1 = t;
        }
        if (threadIdx.x == 0) a.status[job] = bad;
        __syncthreads();
    }
}

// The same bias-window blocks cut in two, as k_diagonal_split.  The left workgroup keeps
// the forward generators f_1..f_m (m = N_z / 2) and the first column of each left part,
//   L_1 = f_1(:, open),  L_k = -f_k L_{k-1}       (L_k = [(A restricted to 1..k)^{-1}]_{k,1}),
// the right one the backward generators g_{N_z}..g_{m+1}.  In the second launch both invert
// the cut block G_mm = (D_m - f_{m-1} - g_{m+1})^{-1}, take G_{m,1} = -G_mm L_{m-1}
// (G_11 itself when m = 1), and run outwards:
//   G_{i,1} = L_i - f_i G_{i+1,1}  (i < m; Dyson at the cut between i and i + 1),
//   G_{k,1} = -g_k G_{k-1,1}       (k > m, as in k_neq),
// the left one accumulating slices 1..m, the right one m+1..N_z.  Per job the workspace
// holds G[k - 1] = f_k or g_k, L_1..L_{m-1} (ldV x Np each), then per side W (Np^2),
// two column blocks and the Gauss-Jordan scratch.
template <int MAXT, int PB, bool TWO>
__global__ __launch_bounds__(MAXT) void k_neq_split(NeqArgs a) {
    __shared__ GjeSh<PB, TWO> s;
    const SysView &sv = a.sv;
    const int Np = sv.Np, N = sv.N, Nz = sv.Nz, ldV = a.nLmax, m = Nz / 2;
    const size_t NN = (size_t)Np * Np, VV = (size_t)ldV * Np;
    const size_t per_side = NN + 2 * VV + 2 * (size_t)PB * Np;
    const size_t per_job = (size_t)Nz * NN + (size_t)m * VV + 2 * per_side;
    const int jl = blockIdx.x >> 1, side = blockIdx.x & 1;
    const int job = a.job_base + jl;
    cd *G = a.ws + (size_t)jl * per_job;
    cd *Lst = G + (size_t)Nz * NN;
    cd *W = Lst + (size_t)m * VV + side * per_side;
    cd *V0 = W + NN, *V1 = V0 + VV;
    cd *scratch = V1 + VV;
    int *status = a.status + 2 * (size_t)job + side;
    const int nL = a.nL[job], off = a.open_off[job];
    if (nL == 0) {                                // no open channel of the left lead
        if (threadIdx.x == 0) *status = 0;
        return;
    }
    JobView jv;
    jv.z = a.z[job];
    jv.kinh = a.kinh + (size_t)a.kidx[job] * Np;
    jv.sigL = a.sigL + (size_t)job * Np;
    jv.sigR = a.sigR + (size_t)job * Np;
    // Vo = add - M(:, sig) Vi on the nL open columns (add == nullptr: -M(:, sig) Vi)
    auto matvec = [&](cd *Vo, const cd *M, const int *sig, const cd *Vi, const cd *add) {
        colprod_mfma(Vo, M, sig, Vi, add, nL, Np);
    };
    if (a.phase == 0) {
        int bad = 0;
        if (side == 0) {
            for (int k = 1; k <= m; ++k) {
                build_block(W, sv, jv, k - 1, k == 1 ? nullptr : G + (size_t)(k - 2) * NN, nullptr);
                __syncthreads();
                if (gje_any<PB, TWO>(W, Np, s, scratch) && !bad) bad = k;
                copy_cols(G + (size_t)(k - 1) * NN, W, s.sigma, Np);
                __syncthreads();
                if (k < m) {
                    cd *Lk = Lst + (size_t)(k - 1) * VV;
                    if (k == 1) {
                        for (int e = threadIdx.x; e < nL * Np; e += blockDim.x) {
                            const int q = e / Np, r = e % Np;
                            Lk[e] = G[r + (size_t)a.Lcol[off + q] * Np];
                        }
                    } else {
                        matvec(Lk, G + (size_t)(k - 1) * NN, nullptr, Lst + (size_t)(k - 2) * VV, nullptr);
                    }
                    __syncthreads();
                }
            }
        } else {
            for (int k = Nz; k > m; --k) {
                build_block(W, sv, jv, k - 1, k == Nz ? nullptr : G + (size_t)k * NN, nullptr);
                __syncthreads();
                if (gje_any<PB, TWO>(W, Np, s, scratch) && !bad) bad = k;
                copy_cols(G + (size_t)(k - 1) * NN, W, s.sigma, Np);
                __syncthreads();
            }
        }
        if (threadIdx.x == 0) *status = bad;
        return;
    }
    const cd wgt = a.wgt[job];
    auto accumulate = [&](int k, const cd *V) {
        cd *ok = a.out + (size_t)k * N * N;
        for (int e = threadIdx.x; e < N * N; e += blockDim.x) {
            const int r = e % N, c = e / N;
            cd v = cmk(0.0, 0.0);
            for (int q = 0; q < nL; ++q) {
                const cd x = V[(size_t)q * Np + r], y = V[(size_t)q * Np + c];
                const double gq = a.gL[off + q];
                v.x += gq * (x.x * y.x + x.y * y.y);       // x conj(y)
                v.y += gq * (x.y * y.x - x.x * y.y);
            }
            const cd t = cmul(wgt, v);
            double *o = reinterpret_cast<double *>(ok + e);
            atomicAdd(o, t.x);
            atomicAdd(o + 1, t.y);
        }
    };
    // cut block: W = D_m - g_{m+1} - f_{m-1} on the physical part, inverted in place
    build_block(W, sv, jv, m - 1, G + (size_t)m * NN, nullptr);
    __syncthreads();
    if (m >= 2) {
        const cd *fl = G + (size_t)(m - 2) * NN;
        for (int e = threadIdx.x; e < N * N; e += blockDim.x) {
            const size_t o = (size_t)(e % N) + (size_t)(e / N) * Np;
            W[o] = csub(W[o], fl[o]);
        }
    }
    __syncthreads();
    const int zp = gje_any<PB, TWO>(W, Np, s, scratch);
    if (m >= 2) {
        matvec(V0, W, s.sigma, Lst + (size_t)(m - 2) * VV, nullptr);
    } else {
        for (int e = threadIdx.x; e < nL * Np; e += blockDim.x) {
            const int q = e / Np, r = e % Np;
            V0[e] = W[r + (size_t)s.sigma[a.Lcol[off + q]] * Np];
        }
    }
    __syncthreads();
    if (side == 0) {
        accumulate(m - 1, V0);
        __syncthreads();
        for (int i = m - 1; i >= 1; --i) {
            matvec(V1, G + (size_t)(i - 1) * NN, nullptr, V0, Lst + (size_t)(i - 1) * VV);
            __syncthreads();
            accumulate(i - 1, V1);
            __syncthreads();
            cd *tt = V0;
            V0 = V1;
            V1 = tt;
        }
    } else {
        for (int k = m + 1; k <= Nz; ++k) {
            matvec(V1, G + (size_t)(k - 1) * NN, nullptr, V0, nullptr);
            __syncthreads();
            accumulate(k - 1, V1);
            __syncthreads();
            cd *tt = V0;
            V0 = V1;
            V1 = tt;
        }
    }
    if (threadIdx.x == 0 && zp && *status == 0) *status = m;
}

// Batched primitives (drop-ins for inverse / matrix_product, src/math_kernel.f90:8-42,
// and the building blocks of the generic tri-diagonal solver).
// inv[b] = (A[b] - (Sub ? Sub[b] : 0))^{-1} for count matrices of N x N (ld N on both
// sides); the work copy is padded to Np.
template <int MAXT, int PB, bool TWO>
__global__ __launch_bounds__(MAXT) void k_inverse_batch(int N, int Np, const cd *A, const cd *Sub,
                                                        cd *work, cd *inv, int *status) {
    __shared__ GjeSh<PB, TWO> s;
    const int b = blockIdx.x;
    const size_t NN = (size_t)Np * Np;
    cd *Wk = work + b * (NN + 2 * (size_t)PB * Np);
    const cd *Ab = A + (size_t)b * N * N;
    const cd *Sb = Sub ? Sub + (size_t)b * N * N : nullptr;
    for (int col = 0; col < Np; ++col)
        for (int r = threadIdx.x; r < Np; r += blockDim.x) {
            cd v = cmk(r == col ? 1.0 : 0.0, 0.0);
            if (r < N && col < N) {
                v = Ab[r + (size_t)col * N];
                if (Sb) v = csub(v, Sb[r + (size_t)col * N]);
            }
            Wk[r + (size_t)col * Np] = v;
        }
    __syncthreads();
    const int zp = gje_any<PB, TWO>(Wk, Np, s, Wk + NN);
    cd *ob = inv + (size_t)b * N * N;
    for (int col = 0; col < N; ++col) {
        const cd *sc = Wk + (size_t)s.sigma[col] * Np;
        for (int r = threadIdx.x; r < N; r += blockDim.x) ob[r + (size_t)col * N] = sc[r];
    }
    if (threadIdx.x == 0) status[b] = zp;
}

// C[b] = alpha * A[ia(b)] * B[ib(b)] (+ I if add_identity), N x N, ld N; operands are
// padded into the work area so every product runs at the padded size Np.
template <int MAXT>
__global__ __launch_bounds__(MAXT) void k_gemm_batch(int N, int Np, const cd *A, long strideA,
                                                     const cd *B, long strideB, cd *C,
                                                     long strideC, double alpha, int add_identity,
                                                     cd *work) {
    const int b = blockIdx.x;
    const size_t NN = (size_t)Np * Np;
    cd *Ap = work + (size_t)b * 2 * NN, *Bp = Ap + NN;
    const cd *Ab = A + b * strideA, *Bb = B + b * strideB;
    for (int col = 0; col < Np; ++col)
        for (int r = threadIdx.x; r < Np; r += blockDim.x) {
            const bool in = r < N && col < N;
            Ap[r + (size_t)col * Np] = in ? Ab[r + (size_t)col * N] : cmk(0.0, 0.0);
            Bp[r + (size_t)col * Np] = in ? Bb[r + (size_t)col * N] : cmk(0.0, 0.0);
        }
    __syncthreads();
    cd *Cb = C + b * strideC;
    gemm_nn(Ap, Bp, nullptr, Np, [&](int r, int c, cd v) {
        if (r < N && c < N) {
            v.x *= alpha;
            v.y *= alpha;
            if (add_identity && r == c) v.x += 1.0;
            Cb[r + (size_t)c * N] = v;
        }
    });
}

// Plane-wave -> real space (src/grid.f90:71-135) by the equivalent Q-grid form:
// rho(Q) = sum_{n_i - n_j = Q mod grid} M_ij, out = unnormalised inverse 2-D DFT of rho.
// One workgroup per matrix; rho and the half-transformed grid live in global scratch.
__global__ __launch_bounds__(256) void k_pw_to_real(int N, const int *nx, const int *ny,
                                                    const cd *mats, int Nx, int Ny,
                                                    const cd *twx, const cd *twy, cd *scratch,
                                                    cd *out) {
    const int b = blockIdx.x;
    const cd *M = mats + (size_t)b * N * N;
    cd *rho = scratch + (size_t)b * 2 * Nx * Ny, *half = rho + (size_t)Nx * Ny;
    for (int i = threadIdx.x; i < Nx * Ny; i += blockDim.x) rho[i] = cmk(0.0, 0.0);
    __syncthreads();
    for (int e = threadIdx.x; e < N * N; e += blockDim.x) {
        const int i = e % N, j = e / N;
        int qx = (nx[i] - nx[j]) % Nx, qy = (ny[i] - ny[j]) % Ny;
        if (qx < 0) qx += Nx;
        if (qy < 0) qy += Ny;
        const cd v = M[e];
        double *o = reinterpret_cast<double *>(rho + qx + (size_t)qy * Nx);
        atomicAdd(o, v.x);
        atomicAdd(o + 1, v.y);
    }
    __syncthreads();
    // half(mx, qy) = sum_qx rho(qx, qy) e^{+2 pi i qx mx / Nx}
    for (int e = threadIdx.x; e < Nx * Ny; e += blockDim.x) {
        const int mx = e % Nx, qy = e / Nx;
        cd acc = cmk(0.0, 0.0);
        for (int qx = 0; qx < Nx; ++qx) acc = cfma(rho[qx + (size_t)qy * Nx], twx[(qx * mx) % Nx], acc);
        half[e] = acc;
    }
    __syncthreads();
    cd *ob = out + (size_t)b * Nx * Ny;
    for (int e = threadIdx.x; e < Nx * Ny; e += blockDim.x) {
        const int mx = e % Nx, my = e / Nx;
        cd acc = cmk(0.0, 0.0);
        for (int qy = 0; qy < Ny; ++qy) acc = cfma(half[mx + (size_t)qy * Nx], twy[(qy * my) % Ny], acc);
        ob[e] = acc;
    }
}

// V-hat of every slice (src/grid.f90:12-69 + src/main.f90:238-247):
//   V(p, q) = (1/(Nx Ny)) sum_{x,y} V(x, y) e^{-2 pi i (p x/Nx + q y/Ny)}, p = 0..NGX,
// p < 0 by conjugate symmetry, copied into the window |p| <= hx, |q| <= hy (zero elsewhere).
// One workgroup per slice; twx/twy hold e^{-2 pi i j/N}.
// One window-restricted DFT pass of the non-local transform: the input has column-major
// dims d[0..3]; along `axis` (length N) the output keeps frequencies f = -R..R,
//     out[.., f + R, ..] = scale * sum_n in[.., n, ..] tw[(f n) mod N],
// tw[k] = exp(+-2 pi i k / N).  One thread per output entry (grid-stride).
__global__ __launch_bounds__(256) void k_dft_axis(const cd *in, cd *out, int d0, int d1, int d2, int d3, int axis,
                                                  int R, const cd *tw, double scale) {
    const int din[4] = {d0, d1, d2, d3};
    int dout[4] = {d0, d1, d2, d3};
    dout[axis] = 2 * R + 1;
    const int N = din[axis];
    const size_t str[4] = {1, (size_t)d0, (size_t)d0 * d1, (size_t)d0 * d1 * d2};
    const size_t total = (size_t)dout[0] * dout[1] * dout[2] * dout[3];
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        size_t rest = e, base = 0;
        int f = 0;
        for (int a = 0; a < 4; ++a) {
            const int o = (int)(rest % dout[a]);
            rest /= dout[a];
            if (a == axis) f = o - R;
            else base += (size_t)o * str[a];
        }
        cd acc = cmk(0.0, 0.0);
        for (int n = 0; n < N; ++n) {
            int k = (int)(((long long)f * n) % N);
            if (k < 0) k += N;
            acc = cfma(in[base + (size_t)n * str[axis]], tw[k], acc);
        }
        out[e] = cmk(acc.x * scale, acc.y * scale);
    }
}

__global__ __launch_bounds__(256) void k_vhat(int Nx, int Ny, const double *V, int NGX, int NGY,
                                              int hx, int hy, const cd *twx, const cd *twy,
                                              cd *scratch, cd *vrec) {
    const int k = blockIdx.x;
    const double *Vs = V + (size_t)k * Nx * Ny;
    cd *A = scratch + (size_t)k * (NGX + 1) * Ny;   // A(p, y) = sum_x V(x,y) e^{-2 pi i p x/Nx}
    for (int e = threadIdx.x; e < (NGX + 1) * Ny; e += blockDim.x) {
        const int p = e % (NGX + 1), y = e / (NGX + 1);
        cd acc = cmk(0.0, 0.0);
        for (int x = 0; x < Nx; ++x) {
            const cd t = twx[(p * x) % Nx];
            const double v = Vs[x + (size_t)y * Nx];
            acc.x = fma(v, t.x, acc.x);
            acc.y = fma(v, t.y, acc.y);
        }
        A[e] = acc;
    }
    __syncthreads();
    const int W = 2 * NGX + 1, H = 2 * NGY + 1;
    cd *out = vrec + (size_t)k * W * H;
    const double norm = 1.0 / ((double)Nx * Ny);
    for (int e = threadIdx.x; e < W * H; e += blockDim.x) {
        const int p = e % W - NGX, q = e / W - NGY;
        cd v = cmk(0.0, 0.0);
        if (p >= -hx && p <= hx && q >= -hy && q <= hy) {
            const int pp = p >= 0 ? p : -p, qq = p >= 0 ? q : -q;
            const int qm = ((qq % Ny) + Ny) % Ny;
            cd acc = cmk(0.0, 0.0);
            for (int y = 0; y < Ny; ++y) acc = cfma(A[pp + (size_t)y * (NGX + 1)], twy[(qm * y) % Ny], acc);
            v = cmk(acc.x * norm, (p >= 0 ? acc.y : -acc.y) * norm);
        }
        out[e] = v;
    }
}

// Materialised slice blocks: one workgroup per slice.  mode 0: H_k (src/grid.f90:248-259);
// mode 1: D_k (src/grid.f90:291-331).  out is N x N per slice (ld N).
__global__ __launch_bounds__(256) void k_blocks(SysView sv, JobView jv, int mode, cd *work, cd *out) {
    const int k = blockIdx.x, N = sv.N, Np = sv.Np;
    const int W = 2 * sv.NGX + 1;
    const cd *vs = sv.vrec + (size_t)k * W * (2 * sv.NGY + 1);
    cd *ob = out + (size_t)k * N * N;
    if (mode == 1) {
        cd *M = work + (size_t)k * Np * Np;
        build_block(M, sv, jv, k, nullptr, nullptr);
        __syncthreads();
        for (int e = threadIdx.x; e < N * N; e += blockDim.x) ob[e] = M[(e % N) + (size_t)(e / N) * Np];
        return;
    }
    for (int e = threadIdx.x; e < N * N; e += blockDim.x) {
        const int r = e % N, c = e / N;
        cd h = vs[(sv.nx[r] - sv.nx[c] + sv.NGX) + W * (sv.ny[r] - sv.ny[c] + sv.NGY)];
        if (r == c) h.x += jv.kinh[r];
        ob[e] = h;
    }
}

// D_k from a given H (E_minus_H_construction's own argument list): one workgroup per
// slice; lead terms (precomputed per basis function) on the first and last slice.
__global__ __launch_bounds__(256) void k_emh_from_h(int N, int Nz, const cd *H, cd z, double scale,
                                                    const cd *sigL, const cd *sigR, cd *out) {
    const int k = blockIdx.x;
    const cd *Hk = H + (size_t)k * N * N;
    cd *ok = out + (size_t)k * N * N;
    for (int e = threadIdx.x; e < N * N; e += blockDim.x) {
        const int r = e % N, c = e / N;
        cd d = cmk(-Hk[e].x, -Hk[e].y);
        if (r == c) d = cadd(d, z);
        d.x *= scale;
        d.y *= scale;
        if (r == c) {
            d.x -= 2.0;
            if (k == 0) d = cadd(d, sigL[r]);
            if (k == Nz - 1) d = cadd(d, sigR[r]);
        }
        ok[e] = d;
    }
}

// tau = sum_ij gL_i gR_j |G_ij|^2 over one N x N block (single workgroup).
__global__ __launch_bounds__(256) void k_trace(int N, const cd *G, const double *gL, const double *gR,
                                               double *tau) {
    __shared__ double red[4];
    double part = 0.0;
    for (int e = threadIdx.x; e < N * N; e += blockDim.x) {
        const int r = e % N, c = e / N;
        part += gL[r] * gR[c] * (G[e].x * G[e].x + G[e].y * G[e].y);
    }
    for (int off = 32; off > 0; off >>= 1) part += __shfl_xor(part, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = part;
    __syncthreads();
    if (threadIdx.x == 0) *tau = red[0] + red[1] + red[2] + red[3];
}

// End-slice term of a semi-infinite flat lead on the device's own finite-difference
// lattice, by decimation (extension; the reference adds exp(i q Delta) with the continuum
// q, src/grid.f90:284-289,323-331).  One thread per (lead, job, basis function).  The
// scaled lead is the chain tridiag(1, d, 1) with d = 2 Delta^2 (z - V - |G + k|^2 / 2) - 2;
// Sancho-Rubio doubling steps in FP64,
//     g = 1/e;  e_s -= a g b;  e -= a g b + b g a;  a <- -a g a;  b <- -b g b,
// run from e_s = e = d, a = b = 1 until |re a| + |im a| < tol.  out = -1/e_s, the term
// added to the end block; iters = steps taken, or -1 when max_iter was reached (an open
// channel at eta = 0 never converges).  out and iters are [2][nj][Np] (left, then right).
__global__ __launch_bounds__(256) void k_lead_decimation(int nj, int N, int Np, double two_d2, double VL,
                                                         double VR, const cd *z, const int *kidx,
                                                         const double *kinh, double tol, int max_iter,
                                                         cd *out, int *iters) {
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= 2 * nj * N) return;
    const int lead = id / (nj * N), rem = id % (nj * N), job = rem / N, i = rem % N;
    const cd zz = z[job];
    const double kh = kinh[(size_t)kidx[job] * Np + i];
    const cd d = cmk(two_d2 * (zz.x - (lead ? VR : VL) - kh) - 2.0, two_d2 * zz.y);
    cd es = d, e = d, a = cmk(1.0, 0.0), b = cmk(1.0, 0.0);
    int done = -1;
    for (int n = 1; n <= max_iter; ++n) {
        const cd g = crecip(e);
        const cd agb = cmul(cmul(a, g), b), bga = cmul(cmul(b, g), a);
        es = csub(es, agb);
        e = csub(csub(e, agb), bga);
        const cd an = cmul(cmul(a, g), a), bn = cmul(cmul(b, g), b);
        a = cmk(-an.x, -an.y);
        b = cmk(-bn.x, -bn.y);
        if (fabs(a.x) + fabs(a.y) < tol) {
            done = n;
            break;
        }
    }
    const cd ge = crecip(es);
    const size_t o = ((size_t)lead * nj + job) * Np + i;
    out[o] = cmk(-ge.x, -ge.y);
    iters[o] = done;
}

// p[0 .. n) = 0 (a kernel rather than a memset, so that it may also sit in a captured graph)
__global__ __launch_bounds__(64) void k_zero_int(int *p, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}

// ------------------------------------------------------------------ host side
struct negf_ctx {
    int device = 0;
    int N = 0, Nz = 0, Np = 0, NGX = 0, NGY = 0, flags = 0;
    double Lx = 0, Ly = 0, Lz = 0, VL = 0, VR = 0, eta = 0;
    std::vector<int> nx, ny;
    int *d_nx = nullptr, *d_ny = nullptr;
    cd *d_vrec = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_kernel_s = 0.0;
    int n_cu = 0;
    // reusable device buffers
    void *ws = nullptr;
    size_t ws_bytes = 0;
    int mixed_refine = 24;  // refinement iteration limit of NEGF_PREC_MIXED_FP32
    double mixed_tol = 1e-7; // its stopping tolerance (estimated absolute error of T)
    int last_fallbacks = 0; // energies of the last mixed call redone in FP64
    bool diag_one_sided = false; // transmission, D mode, bias window: never the two-sided forms (NEGF_DIAG_ONE_SIDED=1)
    bool mixed_fallback = true;  // mixed path: recompute unconverged energies in FP64 (NEGF_MIXED_FALLBACK=0 keeps
                                 // the unconverged mixed values, for measuring the refinement)
    // NEGF_AUTOTUNE: FP64 transmission variant per job-count class (ceil(log2 jobs)),
    // chosen by timing every variant on the first call of the class
    std::map<int, int> tuned;
    double tune_s[4] = {-1.0, -1.0, -1.0, -1.0};   // device seconds of each variant in the last tuning call
    int last_variant = -1;                          // variant of the last FP64 transmission call
    int force_variant = -1;                         // NEGF_TRANS_VARIANT: one variant for every call
};

namespace {

int round_up(int a, int b) { return (a + b - 1) / b * b; }

template <class T>
int dmalloc(T **p, size_t count) {
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void **)p, count * sizeof(T));
    if (e != hipSuccess) return fail(-100 - (int)e, std::string("hipMalloc: ") + hipGetErrorString(e));
    return 0;
}

int ensure_ws(negf_ctx *c, size_t bytes) {
    if (c->ws_bytes >= bytes) return 0;
    if (c->ws) HIPCHK(hipFree(c->ws));
    c->ws = nullptr;
    c->ws_bytes = 0;
    HIPCHK(hipMalloc(&c->ws, bytes));
    c->ws_bytes = bytes;
    return 0;
}

// RAII list of temporary device allocations.
struct Temps {
    std::vector<void *> ptrs;
    ~Temps() {
        for (void *p : ptrs) (void)hipFree(p);
    }
    template <class T>
    int alloc(T **p, size_t count) {
        int st = dmalloc(p, count);
        if (st == 0) ptrs.push_back((void *)*p);
        return st;
    }
    template <class T>
    int upload(T **p, const T *host, size_t count, hipStream_t s) {
        int st = alloc(p, count);
        if (st) return st;
        hipError_t e = hipMemcpyAsync(*p, host, count * sizeof(T), hipMemcpyHostToDevice, s);
        if (e != hipSuccess) return fail(-100 - (int)e, std::string("upload: ") + hipGetErrorString(e));
        return 0;
    }
};

int threads_for(int Np) { return std::min(1024, round_up(Np, 64)); }

// Padded block size: a multiple of the Gauss-Jordan panel width (32 up to 512, 16 above,
// where one thread per row needs the 1024-thread workgroup and its smaller register file).
int padded_size(int N) {
    const int p32 = round_up(N, 32);
    return p32 <= 512 ? p32 : round_up(N, 16);
}
int panel_width(int Np) { return Np <= 512 ? 32 : 16; }

// Kernel shape per padded size: two threads per row (640- or 1024-thread groups) up to
// Np = 512, one thread per row with panel 16 above.
enum Shape { SHAPE_640, SHAPE_1024_TWO, SHAPE_1024_ONE };
Shape shape_for(int Np) { return Np <= 320 ? SHAPE_640 : (Np <= 512 ? SHAPE_1024_TWO : SHAPE_1024_ONE); }
int threads_for_shape(int Np) {
    return shape_for(Np) == SHAPE_1024_ONE ? std::min(1024, round_up(Np, 64)) : round_up(2 * Np, 64);
}

// Per-job host tables: kinetic/2 per k-point, lead terms and trace weights per job.
struct JobTables {
    std::vector<cd> z, sigL, sigR;
    std::vector<double> kinh, gamL, gamR;
    std::vector<int> kidx;
};

int lead_decimation(const negf_ctx *c, JobTables &t, const std::vector<double> &realE,
                    const std::vector<int> &job_e, const std::vector<double> &kin, bool with_gamma);

int job_tables(const negf_ctx *c, const std::vector<std::complex<double>> &energies,
               const std::vector<double> &realE, const double *kx, const double *ky, int nk,
               const std::vector<int> &job_e, const std::vector<int> &job_k, bool with_gamma,
               JobTables &t) {
    const int Np = c->Np, N = c->N;
    const double delta = c->Lz / c->Nz;
    t.kinh.assign((size_t)nk * Np, 0.0);
    std::vector<double> kin((size_t)nk * N);
    for (int ik = 0; ik < nk; ++ik)
        for (int i = 0; i < N; ++i) {
            const double gx = 2 * c->nx[i] * PI / c->Lx + kx[ik];
            const double gy = 2 * c->ny[i] * PI / c->Ly + ky[ik];
            kin[(size_t)ik * N + i] = gx * gx + gy * gy;
            t.kinh[(size_t)ik * Np + i] = (gx * gx + gy * gy) / 2;
        }
    const size_t nj = job_e.size();
    t.z.resize(nj);
    t.kidx.resize(nj);
    t.sigL.assign(nj * Np, make_double2(0, 0));
    t.sigR.assign(nj * Np, make_double2(0, 0));
    if (with_gamma) {
        t.gamL.assign(nj * Np, 0.0);
        t.gamR.assign(nj * Np, 0.0);
    }
    for (size_t j = 0; j < nj; ++j) {
        const std::complex<double> E = energies[job_e[j]] + std::complex<double>(0.0, c->eta);
        t.z[j] = make_double2(E.real(), E.imag());
        t.kidx[j] = job_k[j];
        for (int i = 0; i < N; ++i) {
            const double kn = kin[(size_t)job_k[j] * N + i];
            // src/grid.f90:284-289 (principal sqrt) and :323-331
            const std::complex<double> qR = std::sqrt(2.0 * (E - c->VL) - kn);
            const std::complex<double> qT = std::sqrt(2.0 * (E - c->VR) - kn);
            const std::complex<double> sl = std::exp(std::complex<double>(0, 1) * qR * delta);
            const std::complex<double> sr = std::exp(std::complex<double>(0, 1) * qT * delta);
            t.sigL[j * Np + i] = make_double2(sl.real(), sl.imag());
            t.sigR[j * Np + i] = make_double2(sr.real(), sr.imag());
            if (with_gamma) {
                // src/grid.f90:355-360,365,372: real E, closed channels clamped to 0
                const double Er = realE[job_e[j]];
                const double QR = std::sqrt(std::max(2 * (Er - c->VL) - kn, 0.0));
                const double QT = std::sqrt(std::max(2 * (Er - c->VR) - kn, 0.0));
                t.gamL[j * Np + i] = std::sin(QR * delta) / (delta * delta);
                t.gamR[j * Np + i] = std::sin(QT * delta) / (delta * delta);
            }
        }
    }
    if (c->flags & NEGF_LEAD_DECIMATION) return lead_decimation(c, t, realE, job_e, kin, with_gamma);
    return 0;
}

// NEGF_LEAD_DECIMATION: replace the closed-form lead terms by k_lead_decimation's and
// the trace weights by Im(term) / Delta^2 on the lattice-open channels
// (0 < Delta^2 (E - V - |G + k|^2 / 2) < 2 at real E), 0 elsewhere.
int lead_decimation(const negf_ctx *c, JobTables &t, const std::vector<double> &realE,
                    const std::vector<int> &job_e, const std::vector<double> &kin, bool with_gamma) {
    const int Np = c->Np, N = c->N, nj = (int)t.z.size();
    const int nk = (int)(t.kinh.size() / Np);
    const double delta = c->Lz / c->Nz;
    Temps tmp;
    cd *d_z, *d_out;
    int *d_k, *d_it;
    double *d_kin;
    int st;
    const size_t n = (size_t)2 * nj * Np;
    if ((st = tmp.upload(&d_z, t.z.data(), nj, c->stream)) ||
        (st = tmp.upload(&d_k, t.kidx.data(), nj, c->stream)) ||
        (st = tmp.upload(&d_kin, t.kinh.data(), (size_t)nk * Np, c->stream)) || (st = tmp.alloc(&d_out, n)) ||
        (st = tmp.alloc(&d_it, n)))
        return st;
    const int threads = 2 * nj * N;
    hipLaunchKernelGGL(k_lead_decimation, dim3((threads + 255) / 256), dim3(256), 0, c->stream, nj, N, Np,
                       2.0 * delta * delta, c->VL, c->VR, d_z, d_k, d_kin, 1e-14, 200, d_out, d_it);
    HIPCHK(hipGetLastError());
    std::vector<cd> out(n);
    std::vector<int> it(n);
    HIPCHK(hipMemcpyAsync(out.data(), d_out, n * sizeof(cd), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(it.data(), d_it, n * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int j = 0; j < nj; ++j)
        for (int i = 0; i < N; ++i) {
            const size_t oL = (size_t)j * Np + i, oR = (size_t)(nj + j) * Np + i;
            if (it[oL] < 0 || it[oR] < 0)
                return fail(-21, "lead decimation did not converge in 200 steps (eta = 0 on an open channel?)");
            t.sigL[oL] = out[oL];
            t.sigR[oL] = out[oR];
            if (with_gamma) {
                const double Er = realE[job_e[j]], kn = kin[(size_t)t.kidx[j] * N + i];
                const double xL = delta * delta * (Er - c->VL - kn / 2), xR = delta * delta * (Er - c->VR - kn / 2);
                t.gamL[oL] = (xL > 0 && xL < 2) ? out[oL].y / (delta * delta) : 0.0;
                t.gamR[oL] = (xR > 0 && xR < 2) ? out[oR].y / (delta * delta) : 0.0;
            }
        }
    return 0;
}

int check_ctx(const negf_ctx *c) {
    if (!c) return fail(-1, "null context");
    return 0;
}

SysView sysview(const negf_ctx *c) {
    SysView sv;
    sv.N = c->N;
    sv.Nz = c->Nz;
    sv.Np = c->Np;
    sv.NGX = c->NGX;
    sv.NGY = c->NGY;
    const double delta = c->Lz / c->Nz;
    sv.scale = 2.0 * delta * delta;
    sv.nx = c->d_nx;
    sv.ny = c->d_ny;
    sv.vrec = c->d_vrec;
    return sv;
}

int slots_for(negf_ctx *c, const void *kernel, int threads, int njobs, size_t per_slot_bytes) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0) != hipSuccess ||
        per_cu < 1)
        per_cu = 1;
    long slots = (long)c->n_cu * per_cu;
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) == hipSuccess && per_slot_bytes > 0) {
        free_b += c->ws_bytes;   // the context's own workspace is reused (or released first)
        // keep 4 % (at least 4 GiB) free: at 256 slots of ~0.95 GB (the mixed path's
        // generator history at N = 317, N_z = 1000) a 20 % margin cut the slots to ~240,
        // and 1024 jobs then took five rounds on some workgroups instead of four
        const size_t keep = std::max((size_t)4 << 30, free_b / 25);
        const long fit = free_b > keep ? (long)((free_b - keep) / per_slot_bytes) : 1L;
        slots = std::min(slots, std::max(1L, fit));
    }
    return (int)std::max(1L, std::min<long>(slots, njobs));
}

}  // namespace

// FP64 transmission kernel for padded size Np: the matrix-core Gauss-Jordan with its
// shared state sized for Np <= 64 (5 workgroups per CU), Np <= 128 (3 per CU) or up to
// 512 (1 per CU), and above 512 its four-tile, 16-column form (ceil(Np / 64) waves).
template <bool MULTI>
static const void *trans_kernel(int Np) {
    if (Np <= 64) return (const void *)k_transmission<128, 32, true, MULTI, 64, 32>;
    if (Np <= 128) return (const void *)k_transmission<256, 32, true, MULTI, 128, 64>;
    if (Np <= 320) return (const void *)k_transmission<640, 32, true, MULTI>;
    if (Np <= 512) return (const void *)k_transmission<1024, 32, true, MULTI>;
    return (const void *)k_transmission<1024, 16, true, MULTI, NEGF_MAXNP>;
}
// The two-sided transmission kernel for padded size Np (the same shapes as trans_kernel).
static const void *trans_split_kernel(int Np) {
    if (Np <= 64) return (const void *)k_transmission_split<128, 32, true, 64, 32>;
    if (Np <= 128) return (const void *)k_transmission_split<256, 32, true, 128, 64>;
    if (Np <= 320) return (const void *)k_transmission_split<640, 32, true>;
    if (Np <= 512) return (const void *)k_transmission_split<1024, 32, true>;
    return (const void *)k_transmission_split<1024, 16, true, NEGF_MAXNP>;
}
static void launch_trans_split(int Np, int grid, int threads, hipStream_t st, const TransArgs &a) {
    if (Np <= 64) hipLaunchKernelGGL((k_transmission_split<128, 32, true, 64, 32>), dim3(grid), dim3(threads), 0, st, a);
    else if (Np <= 128) hipLaunchKernelGGL((k_transmission_split<256, 32, true, 128, 64>), dim3(grid), dim3(threads), 0, st, a);
    else if (Np <= 320) hipLaunchKernelGGL((k_transmission_split<640, 32, true>), dim3(grid), dim3(threads), 0, st, a);
    else if (Np <= 512) hipLaunchKernelGGL((k_transmission_split<1024, 32, true>), dim3(grid), dim3(threads), 0, st, a);
    else hipLaunchKernelGGL((k_transmission_split<1024, 16, true, NEGF_MAXNP>), dim3(grid), dim3(threads), 0, st, a);
}
template <bool MULTI>
static void launch_trans(int Np, int slots, int threads, hipStream_t st, const TransArgs &a) {
    if (Np <= 64) hipLaunchKernelGGL((k_transmission<128, 32, true, MULTI, 64, 32>), dim3(slots), dim3(threads), 0, st, a);
    else if (Np <= 128) hipLaunchKernelGGL((k_transmission<256, 32, true, MULTI, 128, 64>), dim3(slots), dim3(threads), 0, st, a);
    else if (Np <= 320) hipLaunchKernelGGL((k_transmission<640, 32, true, MULTI>), dim3(slots), dim3(threads), 0, st, a);
    else if (Np <= 512) hipLaunchKernelGGL((k_transmission<1024, 32, true, MULTI>), dim3(slots), dim3(threads), 0, st, a);
    else hipLaunchKernelGGL((k_transmission<1024, 16, true, MULTI, NEGF_MAXNP>), dim3(slots), dim3(threads), 0, st, a);
}

// FP64 transmission variants: 0 one workgroup per job with the 32-column two-tile
// Gauss-Jordan (Np <= 512), 1 two workgroups per job (two-sided) with it, 2 one workgroup
// per job with the 16-column four-tile form (ceil(Np / 64) waves, any Np), 3 two-sided with
// it.  Above Np = 512 only 2 and 3 exist.
static bool tvar_four(int v) { return v >= 2; }
static bool tvar_split(int v) { return v & 1; }
static int tvar_threads(int Np, int v) {
    return tvar_four(v) ? std::min(1024, round_up(Np, 64)) : threads_for_shape(Np);
}
static const void *tvar_kernel(int Np, int v) {
    if (!tvar_four(v)) return tvar_split(v) ? trans_split_kernel(Np) : trans_kernel<false>(Np);
    return tvar_split(v) ? (const void *)k_transmission_split<1024, 16, true, NEGF_MAXNP>
                            : (const void *)k_transmission<1024, 16, true, false, NEGF_MAXNP>;
}
static void launch_tvar(int Np, int v, int grid, int threads, hipStream_t st, const TransArgs &a) {
    if (!tvar_four(v)) {
        if (tvar_split(v)) launch_trans_split(Np, grid, threads, st, a);
        else launch_trans<false>(Np, grid, threads, st, a);
    } else if (tvar_split(v)) {
        hipLaunchKernelGGL((k_transmission_split<1024, 16, true, NEGF_MAXNP>), dim3(grid), dim3(threads), 0, st, a);
    } else {
        hipLaunchKernelGGL((k_transmission<1024, 16, true, false, NEGF_MAXNP>), dim3(grid), dim3(threads), 0, st, a);
    }
}

static void launch_inverse(int Np, int count, hipStream_t st, int N, const cd *A, const cd *Sub, cd *work,
                           cd *inv, int *status) {
    const int th = threads_for_shape(Np);
    switch (shape_for(Np)) {
    case SHAPE_640:
        hipLaunchKernelGGL((k_inverse_batch<640, 32, true>), dim3(count), dim3(th), 0, st, N, Np, A, Sub, work, inv, status);
        break;
    case SHAPE_1024_TWO:
        hipLaunchKernelGGL((k_inverse_batch<1024, 32, true>), dim3(count), dim3(th), 0, st, N, Np, A, Sub, work, inv, status);
        break;
    default:
        hipLaunchKernelGGL((k_inverse_batch<1024, 16, true>), dim3(count), dim3(th), 0, st, N, Np, A, Sub, work, inv, status);
    }
}

extern "C" {

const char *negf_version(void) { return "negf_gpu 0.1 (gfx950, FP64 RGF)"; }
const char *negf_last_error(void) { return g_err.c_str(); }

int negf_last_fallbacks(const negf_ctx *c) { return c ? c->last_fallbacks : 0; }
int negf_last_variant(const negf_ctx *c) { return c ? c->last_variant : -1; }
double negf_tune_seconds(const negf_ctx *c, int variant) {
    return (c && variant >= 0 && variant < 4) ? c->tune_s[variant] : -1.0;
}

int negf_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int negf_ctx_create(negf_ctx **out, int N, int N_z, const int *nx_grid, const int *ny_grid,
                    double Lx, double Ly, double Lz, double V_L, double V_R, double eta,
                    const double *vrec, int NGX, int NGY, int flags) {
    if (!out) return fail(-1, "out is null");
    *out = nullptr;
    if (N < 1 || N_z < 2) return fail(-2, "need N >= 1 and N_z >= 2");
    if (!nx_grid || !ny_grid || !vrec) return fail(-3, "null array argument");
    if (NGX < 0 || NGY < 0) return fail(-4, "NGX/NGY must be >= 0");
    const int Np = padded_size(N);
    if (Np > NEGF_MAXNP) return fail(-5, "N too large for one workgroup (N <= 1024)");
    for (int i = 0; i < N; ++i) {
        for (int j = 0; j < N; ++j) {
            if (std::abs(nx_grid[i] - nx_grid[j]) > NGX || std::abs(ny_grid[i] - ny_grid[j]) > NGY)
                return fail(-6, "basis differences exceed the V_reciprocal window (NGX/NGY too small)");
        }
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(-7, "no GPU visible");
    auto *c = new negf_ctx();
    c->N = N;
    c->Nz = N_z;
    c->Np = Np;
    c->NGX = NGX;
    c->NGY = NGY;
    c->flags = flags;
    if (const char *r = std::getenv("NEGF_MIXED_REFINE")) c->mixed_refine = std::max(0, std::atoi(r));
    if (const char *r = std::getenv("NEGF_MIXED_TOL")) c->mixed_tol = std::max(1e-15, std::atof(r));
    if (const char *r = std::getenv("NEGF_DIAG_ONE_SIDED")) c->diag_one_sided = std::atoi(r) != 0;
    if (const char *r = std::getenv("NEGF_TRANS_VARIANT")) c->force_variant = std::atoi(r);
    if (const char *r = std::getenv("NEGF_MIXED_FALLBACK")) c->mixed_fallback = std::atoi(r) != 0;
    c->Lx = Lx;
    c->Ly = Ly;
    c->Lz = Lz;
    c->VL = V_L;
    c->VR = V_R;
    c->eta = eta;
    c->nx.assign(nx_grid, nx_grid + N);
    c->ny.assign(ny_grid, ny_grid + N);
    auto bail = [&](int st) {
        negf_ctx_destroy(c);
        return st;
    };
    if (hipGetDevice(&c->device) != hipSuccess) return bail(fail(-8, "hipGetDevice failed"));
    if (hipDeviceGetAttribute(&c->n_cu, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess)
        return bail(fail(-8, "device attribute query failed"));
    if (hipStreamCreate(&c->stream) != hipSuccess) return bail(fail(-9, "hipStreamCreate failed"));
    if (hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess)
        return bail(fail(-9, "hipEventCreate failed"));
    const size_t nv = (size_t)(2 * NGX + 1) * (2 * NGY + 1) * N_z;
    int st;
    if ((st = dmalloc(&c->d_nx, N)) || (st = dmalloc(&c->d_ny, N)) || (st = dmalloc(&c->d_vrec, nv)))
        return bail(st);
    if (hipMemcpy(c->d_nx, nx_grid, N * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->d_ny, ny_grid, N * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->d_vrec, vrec, nv * sizeof(cd), hipMemcpyHostToDevice) != hipSuccess)
        return bail(fail(-10, "upload of basis / V_reciprocal failed"));
    *out = c;
    return 0;
}

int negf_ctx_destroy(negf_ctx *c) {
    if (!c) return 0;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->d_nx) (void)hipFree(c->d_nx);
    if (c->d_ny) (void)hipFree(c->d_ny);
    if (c->d_vrec) (void)hipFree(c->d_vrec);
    if (c->ws) (void)hipFree(c->ws);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return 0;
}

double negf_last_kernel_seconds(negf_ctx *c) { return c ? c->last_kernel_s : -1.0; }

}  // extern "C"

// Copies tau and the per-job status back, records the launch time and sums
// tau(E) += w_k tau(E, k) (src/negf.f90:172).
static int finish_transmission(negf_ctx *c, int nj, const double *d_tau, const int *d_st, const std::vector<int> &je,
                               const std::vector<int> &jk, const double *kw, double *tau, double *tau_jobs) {
    std::vector<double> h_tau(nj);
    std::vector<int> h_st(nj);
    HIPCHK(hipMemcpyAsync(h_tau.data(), d_tau, nj * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(h_st.data(), d_st, nj * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_kernel_s = ms * 1e-3;
    int status = 0;
    for (int j = 0; j < nj; ++j) {
        tau[je[j]] += h_tau[j] * kw[jk[j]];
        if (tau_jobs) tau_jobs[j] = h_tau[j];
        if (h_st[j] && !status) status = h_st[j];
    }
    return status;
}

extern "C" {

int negf_transmission_batch(negf_ctx *c, const double *E, int nE, const double *kx,
                            const double *ky, const double *kw, int nk, double *tau,
                            double *tau_jobs) {
    int st = check_ctx(c);
    if (st) return st;
    if (nE < 0 || nk < 1 || (nE > 0 && (!E || !tau)) || !kx || !ky || !kw)
        return fail(-2, "bad arguments to negf_transmission_batch");
    if (nE == 0) return 0;
    HIPCHK(hipSetDevice(c->device));
    const int nj = nE * nk;
    std::vector<std::complex<double>> energies(nE);
    std::vector<double> realE(E, E + nE);
    std::vector<int> je(nj), jk(nj);
    for (int i = 0; i < nE; ++i) energies[i] = E[i];
    for (int j = 0; j < nj; ++j) {
        je[j] = j % nE;   // i_energy = 1 + mod(i_job - 1, N_E)   (src/negf.f90:132-133)
        jk[j] = j / nE;
    }
    JobTables t;
    if ((st = job_tables(c, energies, realE, kx, ky, nk, je, jk, true, t))) return st;
    // open-channel lists per job
    std::vector<int> off(nj), nL(nj), nR(nj), Lrow, Rcol;
    std::vector<double> gLv, gRv;
    int nLmax = 1;
    for (int j = 0; j < nj; ++j) {
        off[j] = (int)Lrow.size();
        int l = 0, r = 0;
        std::vector<int> rc;
        std::vector<double> rg;
        for (int i = 0; i < c->N; ++i) {
            const double gl = t.gamL[(size_t)j * c->Np + i], gr = t.gamR[(size_t)j * c->Np + i];
            if (gl != 0.0) { Lrow.push_back(i); gLv.push_back(gl); ++l; }
            if (gr != 0.0) { rc.push_back(i); rg.push_back(gr); ++r; }
        }
        // Lrow/Rcol share one offset: pad the shorter list so both start at off[j]
        const int n = std::max(l, r);
        for (int q = l; q < n; ++q) { Lrow.push_back(0); gLv.push_back(0.0); }
        for (int q = 0; q < n; ++q) {
            Rcol.push_back(q < r ? rc[q] : 0);
            gRv.push_back(q < r ? rg[q] : 0.0);
        }
        nL[j] = l;
        nR[j] = r;
        nLmax = std::max(nLmax, l);
    }
    if (Lrow.empty()) { Lrow.push_back(0); Rcol.push_back(0); gLv.push_back(0); gRv.push_back(0); }
    Temps tmp;
    TransArgs a{};
    a.sv = sysview(c);
    a.njobs = nj;
    cd *d_z, *d_sl, *d_sr;
    int *d_k, *d_st, *d_off, *d_nL, *d_nR, *d_Lr, *d_Rc;
    double *d_kin, *d_gl, *d_gr, *d_tau;
    const size_t njp = (size_t)nj * c->Np;
    if ((st = tmp.upload(&d_z, t.z.data(), nj, c->stream)) ||
        (st = tmp.upload(&d_k, t.kidx.data(), nj, c->stream)) ||
        (st = tmp.upload(&d_kin, t.kinh.data(), (size_t)nk * c->Np, c->stream)) ||
        (st = tmp.upload(&d_sl, t.sigL.data(), njp, c->stream)) ||
        (st = tmp.upload(&d_sr, t.sigR.data(), njp, c->stream)) ||
        (st = tmp.upload(&d_off, off.data(), nj, c->stream)) ||
        (st = tmp.upload(&d_nL, nL.data(), nj, c->stream)) ||
        (st = tmp.upload(&d_nR, nR.data(), nj, c->stream)) ||
        (st = tmp.upload(&d_Lr, Lrow.data(), Lrow.size(), c->stream)) ||
        (st = tmp.upload(&d_Rc, Rcol.data(), Rcol.size(), c->stream)) ||
        (st = tmp.upload(&d_gl, gLv.data(), gLv.size(), c->stream)) ||
        (st = tmp.upload(&d_gr, gRv.data(), gRv.size(), c->stream)) ||
        (st = tmp.alloc(&d_tau, nj)) || (st = tmp.alloc(&d_st, nj)))
        return st;
    const int threads = threads_for_shape(c->Np);
    const Shape shp = shape_for(c->Np);
    if ((c->flags & 3) == NEGF_PREC_MIXED_FP32) {
        int nRmax = 1;
        for (int j = 0; j < nj; ++j) nRmax = std::max(nRmax, nR[j]);
        if (shp == SHAPE_1024_ONE)
            return fail(-2, "mixed precision needs N <= 512 (padded block " + std::to_string(c->Np) + ")");
        if (nRmax > NEGF_NRH)
            return fail(-2, "mixed precision handles at most " + std::to_string(NEGF_NRH) + " open channels, got " +
                                std::to_string(nRmax));
        auto al = [](size_t b) { return (b + 255) / 256 * 256; };
        const size_t Np = c->Np, Nz = c->Nz, ldb = nRmax;
        const size_t b_hist = al(Nz * Np * Np * sizeof(cfl)), b_sig = al(Nz * Np * sizeof(int)),
                     b_x = al(Nz * Np * ldb * sizeof(cd)),
                     b_y = al(Nz * Np * ldb * sizeof(cfl));
        const size_t per_slot_m = b_hist + b_sig + b_x + b_y;
        const void *kfm = shp == SHAPE_640 ? (const void *)k_transmission_mixed<640> : (const void *)k_transmission_mixed<1024>;
        const int slots = slots_for(c, kfm, threads, nj, per_slot_m);
        if ((st = ensure_ws(c, per_slot_m * slots))) return st;
        MixedArgs ma;
        ma.a = a;
        ma.a.z = d_z;
        ma.a.kidx = d_k;
        ma.a.kinh = d_kin;
        ma.a.sigL = d_sl;
        ma.a.sigR = d_sr;
        ma.a.open_off = d_off;
        ma.a.nL = d_nL;
        ma.a.nR = d_nR;
        ma.a.Lrow = d_Lr;
        ma.a.Rcol = d_Rc;
        ma.a.gL = d_gl;
        ma.a.gR = d_gr;
        ma.a.nLmax = nLmax;
        ma.a.ws = nullptr;
        ma.a.tau = d_tau;
        ma.a.status = d_st;
        ma.nref = c->mixed_refine;
        ma.tol = c->mixed_tol;
        ma.ldb = nRmax;
        char *base = (char *)c->ws;
        ma.hist = (cfl *)base;
        ma.hsig = (int *)(base + b_hist * slots);
        ma.X = (cd *)(base + (b_hist + b_sig) * slots);
        ma.Y = (cfl *)(base + (b_hist + b_sig + b_x) * slots);
        // Jobs near a subband threshold (an open channel with small q, or a closed one about
        // to open) need the most refinement iterations; the kernel takes them first, so that
        // none of them is started last.  Key: min over channels of |2 (E - V) - |G + k|^2|.
        std::vector<int> order(nj);
        {
            std::vector<double> key(nj);
            for (int j = 0; j < nj; ++j) {
                const double Er = t.z[j].x;
                double dmin = INFINITY;
                for (int i = 0; i < c->N; ++i) {
                    const double kn = 2.0 * t.kinh[(size_t)t.kidx[j] * c->Np + i];
                    dmin = std::min(dmin, std::min(std::fabs(2.0 * (Er - c->VL) - kn), std::fabs(2.0 * (Er - c->VR) - kn)));
                }
                key[j] = dmin;
                order[j] = j;
            }
            std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return key[x] < key[y]; });
        }
        int *d_order;
        if ((st = tmp.upload(&d_order, order.data(), nj, c->stream))) return st;
        ma.order = d_order;
        if ((st = tmp.alloc(&ma.next, 1))) return st;
        hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, c->stream, ma.next, 1);
        // each region holds `slots` padded per-slot blocks; the kernel steps through a region
        // with the unpadded per-slot size, which is never larger
        HIPCHK(hipEventRecord(c->ev0, c->stream));
        if (shp == SHAPE_640) hipLaunchKernelGGL((k_transmission_mixed<640>), dim3(slots), dim3(threads), 0, c->stream, ma);
        else hipLaunchKernelGGL((k_transmission_mixed<1024>), dim3(slots), dim3(threads), 0, c->stream, ma);
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(c->ev1, c->stream));
        std::vector<double> h_tau(nj);
        std::vector<int> h_st(nj);
        HIPCHK(hipMemcpyAsync(h_tau.data(), d_tau, nj * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipMemcpyAsync(h_st.data(), d_st, nj * sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
        // energies with a job whose refinement did not converge run again on the FP64 path
        std::vector<int> redo;
        {
            std::vector<char> mark(nE, 0);
            for (int j = 0; j < nj; ++j)
                if (h_st[j] < 0) mark[je[j]] = 1;
            for (int i = 0; i < nE; ++i)
                if (mark[i]) redo.push_back(i);
        }
        double extra_s = 0.0;
        if (!c->mixed_fallback) redo.clear();
        if (!redo.empty()) {
            const int n = (int)redo.size();
            std::vector<double> Er(n), tr(n, 0.0), tj((size_t)n * nk, 0.0);
            for (int q = 0; q < n; ++q) Er[q] = E[redo[q]];
            const int saved = c->flags;
            c->flags &= ~3;
            const int st2 = negf_transmission_batch(c, Er.data(), n, kx, ky, kw, nk, tr.data(), tj.data());
            c->flags = saved;
            if (st2 < 0) return st2;
            extra_s = c->last_kernel_s;
            for (int q = 0; q < n; ++q)
                for (int k = 0; k < nk; ++k) {
                    const int j = k * nE + redo[q];
                    h_tau[j] = tj[(size_t)k * n + q];
                    h_st[j] = 0;
                }
            if (st2 > 0) h_st[(size_t)0 * nE + redo[0]] = st2;
        }
        c->last_kernel_s = ms * 1e-3 + extra_s;
        c->last_fallbacks = (int)redo.size();
        int status = 0;
        for (int j = 0; j < nj; ++j) {
            tau[je[j]] += h_tau[j] * kw[jk[j]];
            if (tau_jobs) tau_jobs[j] = h_tau[j];
            if (h_st[j] > 0 && !status) status = h_st[j];
        }
        return status;
    }
    a.z = d_z;
    a.kidx = d_k;
    a.kinh = d_kin;
    a.sigL = d_sl;
    a.sigR = d_sr;
    a.open_off = d_off;
    a.nL = d_nL;
    a.nR = d_nR;
    a.Lrow = d_Lr;
    a.Rcol = d_Rc;
    a.gL = d_gl;
    a.gR = d_gr;
    a.nLmax = nLmax;
    a.tau = d_tau;
    a.status = d_st;
    // The variant by the fixed rule: the two-sided form when two workgroups per job fit
    // the GPU at once (few jobs), the two-tile Gauss-Jordan up to Np = 512.
    const bool big = c->Np > 512;
    auto split_fits = [&](int v) {
        int per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, tvar_kernel(c->Np, v), tvar_threads(c->Np, v),
                                                         0) != hipSuccess || per_cu < 1)
            per_cu = 1;
        return 2L * nj <= (long)c->n_cu * per_cu;
    };
    int rule = big ? 2 : 0;
    if (c->Nz >= 2 && !c->diag_one_sided && split_fits(rule + 1)) rule += 1;
    // Launches variant v over all jobs between the context's events.
    auto run_variant = [&](int v) -> int {
        const int th = tvar_threads(c->Np, v);
        const int pw = std::max(panel_width(c->Np), tvar_four(v) ? 16 : 32);
        if (tvar_split(v)) {
            int ldX = 1;
            for (int j = 0; j < nj; ++j) ldX = std::max(ldX, std::max(nL[j], nR[j]));
            const size_t NN = (size_t)c->Np * c->Np, XX = (size_t)ldX * c->Np, SCR = (size_t)2 * pw * c->Np;
            const size_t per_job = (2 * (2 * NN + 2 * XX + SCR) + 3 * NN + XX + SCR) * sizeof(cd);
            int *d_st2;
            int e;
            if ((e = tmp.alloc(&d_st2, 2 * (size_t)nj))) return e;
            const int P = slots_for(c, tvar_kernel(c->Np, v), th, nj, per_job);
            if ((e = ensure_ws(c, per_job * P))) return e;
            a.ws = (cd *)c->ws;
            a.ldX = ldX;
            a.status2 = d_st2;
            HIPCHK(hipEventRecord(c->ev0, c->stream));
            for (int base = 0; base < nj; base += P) {
                a.job_base = base;
                const int n = std::min(P, nj - base);
                for (int ph = 0; ph < 2; ++ph) {
                    a.phase = ph;
                    launch_tvar(c->Np, v, ph == 0 ? 2 * n : n, th, c->stream, a);
                    HIPCHK(hipGetLastError());
                }
            }
            HIPCHK(hipEventRecord(c->ev1, c->stream));
            return 0;
        }
        const size_t per_slot = ((size_t)2 * c->Np * c->Np + (size_t)2 * nLmax * c->Np + (size_t)2 * pw * c->Np) * sizeof(cd);
        const int slots = slots_for(c, tvar_kernel(c->Np, v), th, nj, per_slot);
        int e;
        if ((e = ensure_ws(c, per_slot * slots))) return e;
        a.ws = (cd *)c->ws;
        HIPCHK(hipEventRecord(c->ev0, c->stream));
        launch_tvar(c->Np, v, slots, th, c->stream, a);
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(c->ev1, c->stream));
        return 0;
    };
    int variant = rule;
    if (c->force_variant >= 0) {
        if (c->force_variant > 3 || (big && c->force_variant < 2) || (tvar_split(c->force_variant) && c->Nz < 2))
            return fail(-2, "NEGF_TRANS_VARIANT " + std::to_string(c->force_variant) + " does not exist for this size");
        variant = c->force_variant;
    } else if ((c->flags & NEGF_AUTOTUNE) && !c->diag_one_sided) {
        // Job-count class ceil(log2 jobs): on its first call every variant that exists
        // for this Np runs over the call's own jobs, the fastest (HIP events) is kept for
        // the class, and its results are returned.
        int cls = 0;
        while ((1 << cls) < nj) ++cls;
        auto it = c->tuned.find(cls);
        if (it != c->tuned.end()) {
            variant = it->second;
        } else {
            std::vector<double> best_tau(nj);
            std::vector<int> best_st(nj);
            float best_ms = 0.f;
            int best = -1;
            for (int v = 0; v < 4; ++v) c->tune_s[v] = -1.0;
            for (int v = big ? 2 : 0; v < 4; ++v) {
                if (tvar_split(v) && c->Nz < 2) continue;
                if ((st = run_variant(v))) return st;
                std::vector<double> h_tau(nj);
                std::vector<int> h_st(nj);
                HIPCHK(hipMemcpyAsync(h_tau.data(), d_tau, nj * sizeof(double), hipMemcpyDeviceToHost, c->stream));
                HIPCHK(hipMemcpyAsync(h_st.data(), d_st, nj * sizeof(int), hipMemcpyDeviceToHost, c->stream));
                HIPCHK(hipStreamSynchronize(c->stream));
                float ms = 0.f;
                HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
                c->tune_s[v] = ms * 1e-3;
                if (best < 0 || ms < best_ms) {
                    best = v;
                    best_ms = ms;
                    best_tau.swap(h_tau);
                    best_st.swap(h_st);
                }
            }
            c->tuned[cls] = best;
            c->last_variant = best;
            c->last_kernel_s = best_ms * 1e-3;
            int status = 0;
            for (int j = 0; j < nj; ++j) {
                tau[je[j]] += best_tau[j] * kw[jk[j]];
                if (tau_jobs) tau_jobs[j] = best_tau[j];
                if (best_st[j] && !status) status = best_st[j];
            }
            return status;
        }
    }
    if ((st = run_variant(variant))) return st;
    c->last_variant = variant;
    return finish_transmission(c, nj, d_tau, d_st, je, jk, kw, tau, tau_jobs);
}

static int diagonal_jobs(negf_ctx *c, const std::vector<std::complex<double>> &energies,
                         const std::vector<int> &je, const std::vector<int> &jk,
                         const std::vector<std::complex<double>> &wgt,
                         const std::vector<int> &target, int ntarget, const double *kx,
                         const double *ky, int nk, double *host_out, bool accumulate) {
    int st;
    HIPCHK(hipSetDevice(c->device));
    const int nj = (int)je.size();
    std::vector<double> realE(energies.size());
    for (size_t i = 0; i < energies.size(); ++i) realE[i] = energies[i].real();
    JobTables t;
    if ((st = job_tables(c, energies, realE, kx, ky, nk, je, jk, false, t))) return st;
    Temps tmp;
    DiagArgs a{};
    a.sv = sysview(c);
    a.njobs = nj;
    cd *d_z, *d_sl, *d_sr, *d_w, *d_out;
    int *d_k, *d_t, *d_st;
    double *d_kin;
    const size_t njp = (size_t)nj * c->Np;
    const size_t nout = (size_t)ntarget * c->N * c->N * c->Nz;
    std::vector<cd> hw(nj);
    for (int j = 0; j < nj; ++j) hw[j] = make_double2(wgt[j].real(), wgt[j].imag());
    if ((st = tmp.upload(&d_z, t.z.data(), nj, c->stream)) ||
        (st = tmp.upload(&d_k, t.kidx.data(), nj, c->stream)) ||
        (st = tmp.upload(&d_kin, t.kinh.data(), (size_t)nk * c->Np, c->stream)) ||
        (st = tmp.upload(&d_sl, t.sigL.data(), njp, c->stream)) ||
        (st = tmp.upload(&d_sr, t.sigR.data(), njp, c->stream)) ||
        (st = tmp.upload(&d_w, hw.data(), nj, c->stream)) ||
        (st = tmp.upload(&d_t, target.data(), nj, c->stream)) || (st = tmp.alloc(&d_st, 2 * (size_t)nj)))
        return st;
    if (accumulate) {
        if ((st = tmp.upload(&d_out, reinterpret_cast<const cd *>(host_out), nout, c->stream))) return st;
    } else {
        if ((st = tmp.alloc(&d_out, nout))) return st;
        HIPCHK(hipMemsetAsync(d_out, 0, nout * sizeof(cd), c->stream));
    }
    const int threads = threads_for_shape(c->Np);
    const Shape shp = shape_for(c->Np);
    a.z = d_z;
    a.kidx = d_k;
    a.kinh = d_kin;
    a.sigL = d_sl;
    a.sigR = d_sr;
    a.wgt = d_w;
    a.target = d_t;
    a.out = d_out;
    a.status = d_st;
    // Few jobs (at most half the workgroups the GPU holds at once): the two-sided form,
    // two workgroups per job
    const void *kfs = shp == SHAPE_640 ? (const void *)k_diagonal_split<640, 32, true>
                    : shp == SHAPE_1024_TWO ? (const void *)k_diagonal_split<1024, 32, true>
                                            : (const void *)k_diagonal_split<1024, 16, true>;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kfs, threads, 0) != hipSuccess || per_cu < 1)
        per_cu = 1;
    if (c->Nz >= 2 && nj >= 1 && 2L * nj <= (long)c->n_cu * per_cu && !c->diag_one_sided) {
        const size_t per_job = ((size_t)(c->Nz + 6) * c->Np * c->Np + (size_t)4 * panel_width(c->Np) * c->Np) * sizeof(cd);
        const int P = slots_for(c, kfs, threads, nj, per_job);
        if ((st = ensure_ws(c, per_job * P))) return st;
        a.ws = (cd *)c->ws;
        HIPCHK(hipEventRecord(c->ev0, c->stream));
        for (int base = 0; base < nj; base += P) {
            a.job_base = base;
            const int grid = 2 * std::min(P, nj - base);
            for (int ph = 0; ph < 2; ++ph) {
                a.phase = ph;
                if (shp == SHAPE_640) hipLaunchKernelGGL((k_diagonal_split<640, 32, true>), dim3(grid), dim3(threads), 0, c->stream, a);
                else if (shp == SHAPE_1024_TWO) hipLaunchKernelGGL((k_diagonal_split<1024, 32, true>), dim3(grid), dim3(threads), 0, c->stream, a);
                else hipLaunchKernelGGL((k_diagonal_split<1024, 16, true>), dim3(grid), dim3(threads), 0, c->stream, a);
                HIPCHK(hipGetLastError());
            }
        }
        HIPCHK(hipEventRecord(c->ev1, c->stream));
        std::vector<int> h_st(2 * (size_t)nj);
        HIPCHK(hipMemcpyAsync(host_out, d_out, nout * sizeof(cd), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipMemcpyAsync(h_st.data(), d_st, 2 * (size_t)nj * sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
        c->last_kernel_s = ms * 1e-3;
        for (int j = 0; j < nj; ++j)
            if (h_st[2 * j] || h_st[2 * j + 1]) return h_st[2 * j] ? h_st[2 * j] : h_st[2 * j + 1];
        return 0;
    }
    const size_t per_slot =
        ((size_t)(c->Nz + 2) * c->Np * c->Np + (size_t)2 * panel_width(c->Np) * c->Np) * sizeof(cd);
    const void *kfn = shp == SHAPE_640 ? (const void *)k_diagonal<640, 32, true>
                    : shp == SHAPE_1024_TWO ? (const void *)k_diagonal<1024, 32, true>
                                            : (const void *)k_diagonal<1024, 16, true>;
    const int slots = slots_for(c, kfn, threads, nj, per_slot);
    if ((st = ensure_ws(c, per_slot * slots))) return st;
    a.ws = (cd *)c->ws;
    HIPCHK(hipEventRecord(c->ev0, c->stream));
    if (shp == SHAPE_640) hipLaunchKernelGGL((k_diagonal<640, 32, true>), dim3(slots), dim3(threads), 0, c->stream, a);
    else if (shp == SHAPE_1024_TWO) hipLaunchKernelGGL((k_diagonal<1024, 32, true>), dim3(slots), dim3(threads), 0, c->stream, a);
    else hipLaunchKernelGGL((k_diagonal<1024, 16, true>), dim3(slots), dim3(threads), 0, c->stream, a);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(c->ev1, c->stream));
    std::vector<int> h_st(nj);
    HIPCHK(hipMemcpyAsync(host_out, d_out, nout * sizeof(cd), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(h_st.data(), d_st, nj * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_kernel_s = ms * 1e-3;
    for (int j = 0; j < nj; ++j)
        if (h_st[j]) return h_st[j];
    return 0;
}

static int blocks_common(negf_ctx *c, int mode, std::complex<double> E, double kx, double ky,
                         double *out) {
    int st = check_ctx(c);
    if (st) return st;
    if (!out) return fail(-2, "out is null");
    HIPCHK(hipSetDevice(c->device));
    // job_tables adds +i*eta; the block builders take the energy as given
    std::vector<std::complex<double>> en{E - std::complex<double>(0.0, c->eta)};
    std::vector<double> re{E.real()};
    std::vector<int> je{0}, jk{0};
    JobTables t;
    if ((st = job_tables(c, en, re, &kx, &ky, 1, je, jk, false, t))) return st;
    t.z[0] = make_double2(E.real(), E.imag());
    if (!(c->flags & NEGF_LEAD_DECIMATION)) {
        const double delta = c->Lz / c->Nz;
        for (int i = 0; i < c->N; ++i) {
            const double gx = 2 * c->nx[i] * PI / c->Lx + kx, gy = 2 * c->ny[i] * PI / c->Ly + ky;
            const double kn = gx * gx + gy * gy;
            const std::complex<double> qR = std::sqrt(2.0 * (E - c->VL) - kn);
            const std::complex<double> qT = std::sqrt(2.0 * (E - c->VR) - kn);
            const std::complex<double> sl = std::exp(std::complex<double>(0, 1) * qR * delta);
            const std::complex<double> sr = std::exp(std::complex<double>(0, 1) * qT * delta);
            t.sigL[i] = make_double2(sl.real(), sl.imag());
            t.sigR[i] = make_double2(sr.real(), sr.imag());
        }
    }
    Temps tmp;
    double *d_kin;
    cd *d_sl, *d_sr, *d_w = nullptr, *d_out;
    const size_t nout = (size_t)c->N * c->N * c->Nz;
    if ((st = tmp.upload(&d_kin, t.kinh.data(), (size_t)c->Np, c->stream)) ||
        (st = tmp.upload(&d_sl, t.sigL.data(), (size_t)c->Np, c->stream)) ||
        (st = tmp.upload(&d_sr, t.sigR.data(), (size_t)c->Np, c->stream)) ||
        (st = tmp.alloc(&d_out, nout)))
        return st;
    if (mode == 1 && (st = tmp.alloc(&d_w, (size_t)c->Np * c->Np * c->Nz))) return st;
    JobView jv;
    jv.z = t.z[0];
    jv.kinh = d_kin;
    jv.sigL = d_sl;
    jv.sigR = d_sr;
    hipLaunchKernelGGL(k_blocks, dim3(c->Nz), dim3(256), 0, c->stream, sysview(c), jv, mode, d_w, d_out);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, d_out, nout * sizeof(cd), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

int negf_hamiltonian(negf_ctx *c, double kx, double ky, double *out) {
    return blocks_common(c, 0, 0.0, kx, ky, out);
}

int negf_e_minus_h(negf_ctx *c, double E_re, double E_im, double kx, double ky, double *out) {
    return blocks_common(c, 1, std::complex<double>(E_re, E_im), kx, ky, out);
}

int negf_hamiltonian_direct(int N, int N_z, const int *nx_grid, const int *ny_grid, double Lx,
                            double Ly, double kx, double ky, const double *vrec, int NGX, int NGY,
                            double *out) {
    negf_ctx *c = nullptr;
    int st = negf_ctx_create(&c, N, N_z, nx_grid, ny_grid, Lx, Ly, 1.0, 0.0, 0.0, 0.0, vrec, NGX, NGY, 0);
    if (st) return st;
    st = negf_hamiltonian(c, kx, ky, out);
    negf_ctx_destroy(c);
    return st;
}

// exp(i q Delta) and sin(Q Delta)/Delta^2 per basis function (src/grid.f90:284-289, 354-360)
static void lead_terms(int N, const int *nx, const int *ny, double Lx, double Ly, double delta,
                       double kx, double ky, double VL, double VR, std::complex<double> E,
                       std::vector<cd> *sl, std::vector<cd> *sr, std::vector<double> *gl,
                       std::vector<double> *gr) {
    for (int i = 0; i < N; ++i) {
        const double gx = 2 * nx[i] * PI / Lx + kx, gy = 2 * ny[i] * PI / Ly + ky;
        const double kn = gx * gx + gy * gy;
        if (sl) {
            const std::complex<double> I(0, 1);
            const std::complex<double> a = std::exp(I * std::sqrt(2.0 * (E - VL) - kn) * delta);
            const std::complex<double> b = std::exp(I * std::sqrt(2.0 * (E - VR) - kn) * delta);
            (*sl)[i] = make_double2(a.real(), a.imag());
            (*sr)[i] = make_double2(b.real(), b.imag());
        }
        if (gl) {
            const double Er = E.real();
            (*gl)[i] = std::sin(std::sqrt(std::max(2 * (Er - VL) - kn, 0.0)) * delta) / (delta * delta);
            (*gr)[i] = std::sin(std::sqrt(std::max(2 * (Er - VR) - kn, 0.0)) * delta) / (delta * delta);
        }
    }
}

int negf_e_minus_h_from_h(int N, int N_z, const double *H, double E_re, double E_im,
                          const int *nx_grid, const int *ny_grid, double Lx, double Ly, double Lz,
                          double kx, double ky, double V_L, double V_R, double *out) {
    if (N < 1 || N_z < 1 || !H || !out || !nx_grid || !ny_grid) return fail(-2, "bad arguments to negf_e_minus_h_from_h");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(-7, "no GPU visible");
    const double delta = Lz / N_z;
    std::vector<cd> sl(N), sr(N);
    lead_terms(N, nx_grid, ny_grid, Lx, Ly, delta, kx, ky, V_L, V_R, std::complex<double>(E_re, E_im),
               &sl, &sr, nullptr, nullptr);
    Temps tmp;
    cd *dH, *dO, *dl, *dr;
    int st;
    const size_t n = (size_t)N * N * N_z;
    if ((st = tmp.upload(&dH, reinterpret_cast<const cd *>(H), n, nullptr)) || (st = tmp.alloc(&dO, n)) ||
        (st = tmp.upload(&dl, sl.data(), N, nullptr)) || (st = tmp.upload(&dr, sr.data(), N, nullptr)))
        return st;
    hipLaunchKernelGGL(k_emh_from_h, dim3(N_z), dim3(256), 0, nullptr, N, N_z, dH, make_double2(E_re, E_im),
                       2.0 * delta * delta, dl, dr, dO);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpy(out, dO, n * sizeof(cd), hipMemcpyDeviceToHost));
    return 0;
}

int negf_transmission_trace(int N, int N_z, const double *G, double E, const int *nx_grid,
                            const int *ny_grid, double Lx, double Ly, double Lz, double kx,
                            double ky, double V_L, double V_R, double *tau) {
    if (N < 1 || N_z < 1 || !G || !tau || !nx_grid || !ny_grid) return fail(-2, "bad arguments to negf_transmission_trace");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(-7, "no GPU visible");
    std::vector<double> gl(N), gr(N);
    lead_terms(N, nx_grid, ny_grid, Lx, Ly, Lz / N_z, kx, ky, V_L, V_R, std::complex<double>(E, 0.0),
               nullptr, nullptr, &gl, &gr);
    Temps tmp;
    cd *dG;
    double *dl, *dr, *dt;
    int st;
    if ((st = tmp.upload(&dG, reinterpret_cast<const cd *>(G), (size_t)N * N, nullptr)) ||
        (st = tmp.upload(&dl, gl.data(), N, nullptr)) || (st = tmp.upload(&dr, gr.data(), N, nullptr)) ||
        (st = tmp.alloc(&dt, 1)))
        return st;
    hipLaunchKernelGGL(k_trace, dim3(1), dim3(256), 0, nullptr, N, dG, dl, dr, dt);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpy(tau, dt, sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

int negf_density_batch(negf_ctx *c, const double *z, const double *w, int nnode, const double *kx,
                       const double *ky, const double *kw, int nk, double *density_matrix) {
    int st = check_ctx(c);
    if (st) return st;
    if (nnode < 0 || nk < 1 || !kx || !ky || !kw || (nnode > 0 && (!z || !w || !density_matrix)))
        return fail(-2, "bad arguments to negf_density_batch");
    if (nnode == 0) return 0;
    const double delta = c->Lz / c->Nz, s = 2 * delta * delta;
    std::vector<std::complex<double>> en(nnode), wg;
    std::vector<int> je, jk, tg;
    for (int n = 0; n < nnode; ++n) en[n] = std::complex<double>(z[2 * n], z[2 * n + 1]);
    for (int ik = 0; ik < nk; ++ik)
        for (int n = 0; n < nnode; ++n) {
            je.push_back(n);
            jk.push_back(ik);
            tg.push_back(0);
            wg.push_back(std::complex<double>(w[2 * n], w[2 * n + 1]) * kw[ik] * s);
        }
    return diagonal_jobs(c, en, je, jk, wg, tg, 1, kx, ky, nk, density_matrix, true);
}

int negf_neq_density_batch(negf_ctx *c, const double *E, const double *w, int nE, const double *kx,
                           const double *ky, const double *kw, int nk, double *density_matrix) {
    int st = check_ctx(c);
    if (st) return st;
    if (nE < 0 || nk < 1 || !kx || !ky || !kw || (nE > 0 && (!E || !w || !density_matrix)))
        return fail(-2, "bad arguments to negf_neq_density_batch");
    if (nE == 0) return 0;
    HIPCHK(hipSetDevice(c->device));
    const double delta = c->Lz / c->Nz, sc = 2 * delta * delta;
    const int nj = nE * nk;
    std::vector<std::complex<double>> energies(nE);
    std::vector<double> realE(E, E + nE);
    std::vector<int> je(nj), jk(nj);
    std::vector<cd> hw(nj);
    for (int i = 0; i < nE; ++i) energies[i] = E[i];
    for (int j = 0; j < nj; ++j) {
        je[j] = j % nE;
        jk[j] = j / nE;
        const std::complex<double> wj = std::complex<double>(w[2 * je[j]], w[2 * je[j] + 1]) * kw[jk[j]] * sc * sc;
        hw[j] = make_double2(wj.real(), wj.imag());
    }
    JobTables t;
    if ((st = job_tables(c, energies, realE, kx, ky, nk, je, jk, true, t))) return st;
    std::vector<int> off(nj), nL(nj), Lcol;
    std::vector<double> gLv;
    int nLmax = 1;
    for (int j = 0; j < nj; ++j) {
        off[j] = (int)Lcol.size();
        int l = 0;
        for (int i = 0; i < c->N; ++i) {
            const double g = t.gamL[(size_t)j * c->Np + i];
            if (g != 0.0) {
                Lcol.push_back(i);
                gLv.push_back(g);
                ++l;
            }
        }
        nL[j] = l;
        nLmax = std::max(nLmax, l);
    }
    if (Lcol.empty()) {
        Lcol.push_back(0);
        gLv.push_back(0.0);
    }
    Temps tmp;
    NeqArgs a{};
    a.sv = sysview(c);
    a.njobs = nj;
    cd *d_z, *d_sl, *d_sr, *d_w, *d_out;
    int *d_k, *d_st, *d_off, *d_nL, *d_Lc;
    double *d_kin, *d_gl;
    const size_t njp = (size_t)nj * c->Np;
    const size_t nout = (size_t)c->N * c->N * c->Nz;
    if ((st = tmp.upload(&d_z, t.z.data(), nj, c->stream)) ||
        (st = tmp.upload(&d_k, t.kidx.data(), nj, c->stream)) ||
        (st = tmp.upload(&d_kin, t.kinh.data(), (size_t)nk * c->Np, c->stream)) ||
        (st = tmp.upload(&d_sl, t.sigL.data(), njp, c->stream)) ||
        (st = tmp.upload(&d_sr, t.sigR.data(), njp, c->stream)) ||
        (st = tmp.upload(&d_w, hw.data(), nj, c->stream)) ||
        (st = tmp.upload(&d_off, off.data(), nj, c->stream)) ||
        (st = tmp.upload(&d_nL, nL.data(), nj, c->stream)) ||
        (st = tmp.upload(&d_Lc, Lcol.data(), Lcol.size(), c->stream)) ||
        (st = tmp.upload(&d_gl, gLv.data(), gLv.size(), c->stream)) ||
        (st = tmp.upload(&d_out, reinterpret_cast<const cd *>(density_matrix), nout, c->stream)) ||
        (st = tmp.alloc(&d_st, 2 * (size_t)nj)))
        return st;
    const int threads = threads_for_shape(c->Np);
    const Shape shp = shape_for(c->Np);
    a.z = d_z;
    a.kidx = d_k;
    a.kinh = d_kin;
    a.sigL = d_sl;
    a.sigR = d_sr;
    a.wgt = d_w;
    a.open_off = d_off;
    a.nL = d_nL;
    a.Lcol = d_Lc;
    a.gL = d_gl;
    a.nLmax = nLmax;
    a.out = d_out;
    a.status = d_st;
    // few jobs: two workgroups per job (k_neq_split), as in D mode
    const void *kfs = shp == SHAPE_640 ? (const void *)k_neq_split<640, 32, true>
                    : shp == SHAPE_1024_TWO ? (const void *)k_neq_split<1024, 32, true>
                                            : (const void *)k_neq_split<1024, 16, true>;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kfs, threads, 0) != hipSuccess || per_cu < 1)
        per_cu = 1;
    if (c->Nz >= 2 && 2L * nj <= (long)c->n_cu * per_cu && !c->diag_one_sided) {
        const size_t NN = (size_t)c->Np * c->Np, VV = (size_t)nLmax * c->Np;
        const size_t per_job = ((size_t)c->Nz * NN + (size_t)(c->Nz / 2) * VV +
                                2 * (NN + 2 * VV + (size_t)2 * panel_width(c->Np) * c->Np)) * sizeof(cd);
        const int P = slots_for(c, kfs, threads, nj, per_job);
        if ((st = ensure_ws(c, per_job * P))) return st;
        a.ws = (cd *)c->ws;
        HIPCHK(hipEventRecord(c->ev0, c->stream));
        for (int base = 0; base < nj; base += P) {
            a.job_base = base;
            const int grid = 2 * std::min(P, nj - base);
            for (int ph = 0; ph < 2; ++ph) {
                a.phase = ph;
                if (shp == SHAPE_640) hipLaunchKernelGGL((k_neq_split<640, 32, true>), dim3(grid), dim3(threads), 0, c->stream, a);
                else if (shp == SHAPE_1024_TWO) hipLaunchKernelGGL((k_neq_split<1024, 32, true>), dim3(grid), dim3(threads), 0, c->stream, a);
                else hipLaunchKernelGGL((k_neq_split<1024, 16, true>), dim3(grid), dim3(threads), 0, c->stream, a);
                HIPCHK(hipGetLastError());
            }
        }
        HIPCHK(hipEventRecord(c->ev1, c->stream));
        std::vector<int> h_st(2 * (size_t)nj);
        HIPCHK(hipMemcpyAsync(density_matrix, d_out, nout * sizeof(cd), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipMemcpyAsync(h_st.data(), d_st, 2 * (size_t)nj * sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
        c->last_kernel_s = ms * 1e-3;
        for (int j = 0; j < nj; ++j)
            if (h_st[2 * j] || h_st[2 * j + 1]) return h_st[2 * j] ? h_st[2 * j] : h_st[2 * j + 1];
        return 0;
    }
    const size_t per_slot = ((size_t)(c->Nz + 1) * c->Np * c->Np + (size_t)2 * nLmax * c->Np +
                             (size_t)2 * panel_width(c->Np) * c->Np) * sizeof(cd);
    const void *kfn = shp == SHAPE_640 ? (const void *)k_neq<640, 32, true>
                    : shp == SHAPE_1024_TWO ? (const void *)k_neq<1024, 32, true>
                                            : (const void *)k_neq<1024, 16, true>;
    const int slots = slots_for(c, kfn, threads, nj, per_slot);
    if ((st = ensure_ws(c, per_slot * slots))) return st;
    a.ws = (cd *)c->ws;
    HIPCHK(hipEventRecord(c->ev0, c->stream));
    if (shp == SHAPE_640) hipLaunchKernelGGL((k_neq<640, 32, true>), dim3(slots), dim3(threads), 0, c->stream, a);
    else if (shp == SHAPE_1024_TWO) hipLaunchKernelGGL((k_neq<1024, 32, true>), dim3(slots), dim3(threads), 0, c->stream, a);
    else hipLaunchKernelGGL((k_neq<1024, 16, true>), dim3(slots), dim3(threads), 0, c->stream, a);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(c->ev1, c->stream));
    std::vector<int> h_st(nj);
    HIPCHK(hipMemcpyAsync(density_matrix, d_out, nout * sizeof(cd), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(h_st.data(), d_st, nj * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_kernel_s = ms * 1e-3;
    for (int j = 0; j < nj; ++j)
        if (h_st[j]) return h_st[j];
    return 0;
}

int negf_ldos_batch(negf_ctx *c, const double *E, int nE, const double *kx, const double *ky,
                    const double *kw, int nk, double *ldos_matrix) {
    int st = check_ctx(c);
    if (st) return st;
    if (nE < 0 || nk < 1 || !kx || !ky || !kw || (nE > 0 && (!E || !ldos_matrix)))
        return fail(-2, "bad arguments to negf_ldos_batch");
    if (nE == 0) return 0;
    const double delta = c->Lz / c->Nz, s = 2 * delta * delta;
    std::vector<std::complex<double>> en(nE), wg;
    std::vector<int> je, jk, tg;
    for (int i = 0; i < nE; ++i) en[i] = E[i];
    for (int ik = 0; ik < nk; ++ik)
        for (int i = 0; i < nE; ++i) {
            je.push_back(i);
            jk.push_back(ik);
            tg.push_back(i);
            wg.push_back(-kw[ik] / PI * s);   // src/negf.f90:230-231
        }
    return diagonal_jobs(c, en, je, jk, wg, tg, nE, kx, ky, nk, ldos_matrix, false);
}

int negf_inverse_batch(int N, int count, const double *A, double *inv) {
    if (N < 1 || count < 0 || (count > 0 && (!A || !inv))) return fail(-2, "bad arguments to negf_inverse_batch");
    if (count == 0) return 0;
    const int Np = padded_size(N);
    if (Np > NEGF_MAXNP) return fail(-5, "N too large (N <= 1024)");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(-7, "no GPU visible");
    Temps tmp;
    cd *dA, *dW, *dI;
    int *dS;
    int st;
    const size_t nn = (size_t)N * N * count;
    if ((st = tmp.upload(&dA, reinterpret_cast<const cd *>(A), nn, nullptr)) ||
        (st = tmp.alloc(&dW, ((size_t)Np * Np + 2 * (size_t)panel_width(Np) * Np) * count)) || (st = tmp.alloc(&dI, nn)) ||
        (st = tmp.alloc(&dS, count)))
        return st;
    launch_inverse(Np, count, nullptr, N, dA, nullptr, dW, dI, dS);
    HIPCHK(hipGetLastError());
    std::vector<int> hs(count);
    HIPCHK(hipMemcpy(inv, dI, nn * sizeof(cd), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(hs.data(), dS, count * sizeof(int), hipMemcpyDeviceToHost));
    for (int b = 0; b < count; ++b)
        if (hs[b]) return b + 1;
    return 0;
}

}  // extern "C"

// ------------------------------------------------------------------ allocation-free inverse
// negf_inverse_batch_noalloc takes the Gauss-Jordan work areas of k_inverse_batch from
// `scratch`, a file-scope array in the device memory of every GPU the library runs on
// (NEGF_NOALLOC_SCRATCH_BYTES, held until the process ends).  In host code the name
// `scratch` is a host-side stand-in, not a device address: every host use goes through
// the address hipGetSymbolAddress returned for the current device, kept per device.
// k_inverse_batch writes each work area in full (the padded matrix, then the pivot-row
// staging) before reading it, so nothing needs clearing between calls; the array keeps
// the last call's contents, and calls that overlap on one GPU race on it.
#define NEGF_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
static __device__ __align__(256) unsigned char scratch[NEGF_NOALLOC_SCRATCH_BYTES];

namespace {
std::mutex g_scratch_mu;
void *g_scratch_dev[64];   // per device: the address of `scratch`, or null before the first call

int scratch_address(void **out) {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return fail(-7, "device index out of range for the scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(-100 - (int)e, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_dev[dev] = p;
    }
    *out = g_scratch_dev[dev];
    return 0;
}
}  // namespace

extern "C" {

int negf_inverse_batch_noalloc(int N, int count, const double *A, double *inv, int *status, void *stream) {
    if (N < 1 || count < 0 || (count > 0 && (!A || !inv || !status)))
        return fail(-2, "bad arguments to negf_inverse_batch_noalloc");
    const int Np = padded_size(N);
    if (Np > NEGF_MAXNP) return fail(-5, "N too large (N <= 1024)");
    const size_t per = ((size_t)Np * Np + 2 * (size_t)panel_width(Np) * Np) * sizeof(cd);
    if ((size_t)count * per > NEGF_NOALLOC_SCRATCH_BYTES)
        return fail(-6, "negf_inverse_batch_noalloc: " + std::to_string(count) + " matrices of N = " +
                            std::to_string(N) + " need " + std::to_string((size_t)count * per) +
                            " bytes of work area; the scratch holds " + std::to_string(NEGF_NOALLOC_SCRATCH_BYTES) +
                            " (at most " + std::to_string(NEGF_NOALLOC_SCRATCH_BYTES / per) + " per call)");
    if (count == 0) return 0;
    void *w = nullptr;
    if (int st = scratch_address(&w)) return st;
    launch_inverse(Np, count, (hipStream_t)stream, N, reinterpret_cast<const cd *>(A), nullptr,
                   static_cast<cd *>(w), reinterpret_cast<cd *>(inv), status);
    HIPCHK(hipGetLastError());
    return 0;
}

int negf_matrix_product_batch(int N, int count, const double *A, const double *B, double *C) {
    if (N < 1 || count < 0 || (count > 0 && (!A || !B || !C))) return fail(-2, "bad arguments to negf_matrix_product_batch");
    if (count == 0) return 0;
    const int Np = padded_size(N);
    if (Np > NEGF_MAXNP) return fail(-5, "N too large (N <= 1024)");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(-7, "no GPU visible");
    Temps tmp;
    cd *dA, *dB, *dC, *dW;
    int st;
    const size_t nn = (size_t)N * N * count;
    if ((st = tmp.upload(&dA, reinterpret_cast<const cd *>(A), nn, nullptr)) ||
        (st = tmp.upload(&dB, reinterpret_cast<const cd *>(B), nn, nullptr)) ||
        (st = tmp.alloc(&dC, nn)) || (st = tmp.alloc(&dW, (size_t)2 * Np * Np * count)))
        return st;
    const long s = (long)N * N;
    if (threads_for(Np) <= 512)
        hipLaunchKernelGGL(k_gemm_batch<512>, dim3(count), dim3(threads_for(Np)), 0, nullptr, N, Np, dA, s,
                           dB, s, dC, s, 1.0, 0, dW);
    else
        hipLaunchKernelGGL(k_gemm_batch<1024>, dim3(count), dim3(threads_for(Np)), 0, nullptr, N, Np, dA, s,
                           dB, s, dC, s, 1.0, 0, dW);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpy(C, dC, nn * sizeof(cd), hipMemcpyDeviceToHost));
    return 0;
}

// Generic tri-diagonal solver with identity off-diagonal blocks, host-orchestrated over
// the batched primitives (one launch per slice step), following
// src/math_kernel.f90:80-162 operation by operation.
int negf_rgf_tri_solve(int N, int N_z, const double *G_inv_D, int mode, double *diag,
                       double *first, double *last) {
    if (N < 1 || N_z < 2 || !G_inv_D) return fail(-2, "bad arguments to negf_rgf_tri_solve");
    if ((mode & 1) && !diag) return fail(-2, "mode D needs diag");
    if ((mode & 2) && !first) return fail(-2, "mode F needs first");
    if ((mode & 4) && !last) return fail(-2, "mode L needs last");
    const int Np = padded_size(N);
    if (Np > NEGF_MAXNP) return fail(-5, "N too large (N <= 1024)");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(-7, "no GPU visible");
    Temps tmp;
    const size_t NN = (size_t)N * N;
    const int th = threads_for(Np);
    cd *dD, *gr, *gl, *dOut, *dW, *dM;
    int *dS;
    int st;
    if ((st = tmp.upload(&dD, reinterpret_cast<const cd *>(G_inv_D), NN * N_z, nullptr)) ||
        (st = tmp.alloc(&gr, NN * (N_z - 1))) || (st = tmp.alloc(&gl, NN * (N_z - 1))) ||
        (st = tmp.alloc(&dOut, NN * N_z)) || (st = tmp.alloc(&dW, (size_t)2 * Np * Np + 2 * (size_t)panel_width(Np) * Np)) ||
        (st = tmp.alloc(&dM, NN)) || (st = tmp.alloc(&dS, 1)))
        return st;
    int status = 0;
    std::vector<int> hs(1);
    auto inv = [&](const cd *A, const cd *Sub, cd *o, int slice) -> int {
        launch_inverse(Np, 1, nullptr, N, A, Sub, dW, o, dS);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpy(hs.data(), dS, sizeof(int), hipMemcpyDeviceToHost));
        if (hs[0] && !status) status = slice;
        return 0;
    };
    auto gemm = [&](const cd *A, const cd *B, cd *C, double alpha, int addI) -> int {
        if (th <= 512)
            hipLaunchKernelGGL(k_gemm_batch<512>, dim3(1), dim3(th), 0, nullptr, N, Np, A, 0L, B, 0L, C, 0L,
                               alpha, addI, dW);
        else
            hipLaunchKernelGGL(k_gemm_batch<1024>, dim3(1), dim3(th), 0, nullptr, N, Np, A, 0L, B, 0L, C, 0L,
                               alpha, addI, dW);
        HIPCHK(hipGetLastError());
        return 0;
    };
    const bool D = mode & 1, F = mode & 2, L = mode & 4;
    if (D || F) {
        if ((st = inv(dD + NN * (N_z - 1), nullptr, gr + NN * (N_z - 2), N_z))) return st;
        for (int k = 2; k <= N_z - 1; ++k)
            if ((st = inv(dD + NN * (N_z - k), gr + NN * (N_z - k), gr + NN * (N_z - k - 1), N_z - k + 1)))
                return st;
    }
    if (D) {
        if ((st = inv(dD, gr, dOut, 1))) return st;
        for (int k = 2; k <= N_z; ++k) {
            if ((st = gemm(dOut + NN * (k - 2), gr + NN * (k - 2), dM, 1.0, 1))) return st;
            if ((st = gemm(gr + NN * (k - 2), dM, dOut + NN * (k - 1), 1.0, 0))) return st;
        }
        HIPCHK(hipMemcpy(diag, dOut, NN * N_z * sizeof(cd), hipMemcpyDeviceToHost));
    }
    if (F) {
        if ((st = inv(dD, gr, dOut, 1))) return st;
        for (int k = 2; k <= N_z; ++k)
            if ((st = gemm(gr + NN * (k - 2), dOut + NN * (k - 2), dOut + NN * (k - 1), -1.0, 0))) return st;
        HIPCHK(hipMemcpy(first, dOut, NN * N_z * sizeof(cd), hipMemcpyDeviceToHost));
    }
    if (L) {
        if ((st = inv(dD, nullptr, gl, 1))) return st;
        for (int k = 2; k <= N_z - 1; ++k)
            if ((st = inv(dD + NN * (k - 1), gl + NN * (k - 2), gl + NN * (k - 1), k))) return st;
        if ((st = inv(dD + NN * (N_z - 1), gl + NN * (N_z - 2), dOut + NN * (N_z - 1), N_z))) return st;
        for (int k = 1; k <= N_z - 1; ++k)
            if ((st = gemm(gl + NN * (N_z - k - 1), dOut + NN * (N_z - k), dOut + NN * (N_z - k - 1), -1.0, 0)))
                return st;
        HIPCHK(hipMemcpy(last, dOut, NN * N_z * sizeof(cd), hipMemcpyDeviceToHost));
    }
    HIPCHK(hipDeviceSynchronize());
    return status;
}

static void twiddles(int n, double sign, std::vector<cd> &t) {
    t.resize(n);
    for (int j = 0; j < n; ++j) {
        // angle 2 pi j / n, evaluated with the index reduced so cos/sin stay accurate
        const double ang = 2.0 * PI * (double)j / (double)n;
        t[j] = make_double2(std::cos(ang), sign * std::sin(ang));
    }
}

int negf_pw_to_real(int N, const int *nx_grid, const int *ny_grid, int nmat, const double *mats,
                    int N_x, int N_y, double *out) {
    if (N < 1 || nmat < 0 || N_x < 1 || N_y < 1 || !nx_grid || !ny_grid || (nmat > 0 && (!mats || !out)))
        return fail(-2, "bad arguments to negf_pw_to_real");
    if (nmat == 0) return 0;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(-7, "no GPU visible");
    std::vector<cd> twx, twy;
    twiddles(N_x, +1.0, twx);
    twiddles(N_y, +1.0, twy);
    Temps tmp;
    int *dnx, *dny;
    cd *dM, *dtx, *dty, *dS, *dO;
    int st;
    const size_t g = (size_t)N_x * N_y;
    if ((st = tmp.upload(&dnx, nx_grid, N, nullptr)) || (st = tmp.upload(&dny, ny_grid, N, nullptr)) ||
        (st = tmp.upload(&dM, reinterpret_cast<const cd *>(mats), (size_t)N * N * nmat, nullptr)) ||
        (st = tmp.upload(&dtx, twx.data(), N_x, nullptr)) || (st = tmp.upload(&dty, twy.data(), N_y, nullptr)) ||
        (st = tmp.alloc(&dS, 2 * g * nmat)) || (st = tmp.alloc(&dO, g * nmat)))
        return st;
    hipLaunchKernelGGL(k_pw_to_real, dim3(nmat), dim3(256), 0, nullptr, N, dnx, dny, dM, N_x, N_y, dtx,
                       dty, dS, dO);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpy(out, dO, g * nmat * sizeof(cd), hipMemcpyDeviceToHost));
    return 0;
}

// NonLocalPotential_RealToPlanewave (src/grid.f90:137-234): four window-restricted DFT
// passes -- backward over x' and y', forward over x and y -- with the reference's
// (-1)^(i+j-ii-jj) shift and 1/(N_x N_y)^2 folded into the frequencies and the scale.
int negf_nonlocal_vhat_from_real(int N_x, int N_y, const double *V_real, int Rx, int Ry, double *vrec) {
    if (N_x < 1 || N_y < 1 || Rx < 0 || Ry < 0 || !V_real || !vrec)
        return fail(-2, "bad arguments to negf_nonlocal_vhat_from_real");
    // src/grid.f90:157-161
    if (2 * Rx > N_x || 2 * Ry > N_y) return fail(-3, "size of V_reciprocal minus one larger than V_real");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(-7, "no GPU visible");
    std::vector<cd> tpx, tpy, tmx, tmy;
    twiddles(N_x, +1.0, tpx);
    twiddles(N_y, +1.0, tpy);
    twiddles(N_x, -1.0, tmx);
    twiddles(N_y, -1.0, tmy);
    const int fx = 2 * Rx + 1, fy = 2 * Ry + 1;
    const size_t n0 = (size_t)N_x * N_y * N_x * N_y;
    const size_t n1 = (size_t)N_x * N_y * fx * N_y, n2 = (size_t)N_x * N_y * fx * fy;
    const size_t n3 = (size_t)fx * N_y * fx * fy, n4 = (size_t)fx * fy * fx * fy;
    Temps tmp;
    cd *dV, *d1, *d2, *d3, *d4, *dpx, *dpy, *dmx, *dmy;
    int st;
    if ((st = tmp.upload(&dV, reinterpret_cast<const cd *>(V_real), n0, nullptr)) ||
        (st = tmp.upload(&dpx, tpx.data(), N_x, nullptr)) || (st = tmp.upload(&dpy, tpy.data(), N_y, nullptr)) ||
        (st = tmp.upload(&dmx, tmx.data(), N_x, nullptr)) || (st = tmp.upload(&dmy, tmy.data(), N_y, nullptr)) ||
        (st = tmp.alloc(&d1, n1)) || (st = tmp.alloc(&d2, n2)) || (st = tmp.alloc(&d3, n3)) || (st = tmp.alloc(&d4, n4)))
        return st;
    auto grid = [](size_t n) { return dim3((unsigned)std::min<size_t>((n + 255) / 256, 65535)); };
    const double sc = 1.0 / ((double)N_x * N_y * (double)N_x * N_y);
    hipLaunchKernelGGL(k_dft_axis, grid(n1), dim3(256), 0, nullptr, dV, d1, N_x, N_y, N_x, N_y, 2, Rx, dpx, 1.0);
    hipLaunchKernelGGL(k_dft_axis, grid(n2), dim3(256), 0, nullptr, d1, d2, N_x, N_y, fx, N_y, 3, Ry, dpy, 1.0);
    hipLaunchKernelGGL(k_dft_axis, grid(n3), dim3(256), 0, nullptr, d2, d3, N_x, N_y, fx, fy, 0, Rx, dmx, 1.0);
    hipLaunchKernelGGL(k_dft_axis, grid(n4), dim3(256), 0, nullptr, d3, d4, fx, N_y, fx, fy, 1, Ry, dmy, sc);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpy(vrec, d4, n4 * sizeof(cd), hipMemcpyDeviceToHost));
    return 0;
}

int negf_vhat_from_real(int N_x, int N_y, int N_z, const double *V_real, int NGX, int NGY,
                        int full_window, double *vrec) {
    if (N_x < 1 || N_y < 1 || N_z < 1 || !V_real || !vrec || NGX < 0 || NGY < 0)
        return fail(-2, "bad arguments to negf_vhat_from_real");
    // src/grid.f90:33-37 and the window bound of :58-60
    if (2 * NGX > N_x || 2 * NGY > N_y) return fail(-3, "size of V_reciprocal minus one larger than V_real");
    if (NGX > N_x / 2 || NGY + 1 + N_y / 2 > N_y) return fail(-3, "V_reciprocal window does not fit the real-space grid");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(-7, "no GPU visible");
    std::vector<cd> twx, twy;
    twiddles(N_x, -1.0, twx);
    twiddles(N_y, -1.0, twy);
    Temps tmp;
    double *dV;
    cd *dtx, *dty, *dS, *dO;
    int st;
    const size_t nv = (size_t)(2 * NGX + 1) * (2 * NGY + 1) * N_z;
    if ((st = tmp.upload(&dV, V_real, (size_t)N_x * N_y * N_z, nullptr)) ||
        (st = tmp.upload(&dtx, twx.data(), N_x, nullptr)) || (st = tmp.upload(&dty, twy.data(), N_y, nullptr)) ||
        (st = tmp.alloc(&dS, (size_t)(NGX + 1) * N_y * N_z)) || (st = tmp.alloc(&dO, nv)))
        return st;
    const int hx = full_window ? NGX : NGX / 2, hy = full_window ? NGY : NGY / 2;
    hipLaunchKernelGGL(k_vhat, dim3(N_z), dim3(256), 0, nullptr, N_x, N_y, dV, NGX, NGY, hx, hy, dtx,
                       dty, dS, dO);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpy(vrec, dO, nv * sizeof(cd), hipMemcpyDeviceToHost));
    return 0;
}

}  // extern "C"

// ------------------------------------------------------------------ uneven batches
// Several devices (contexts of different N and N_z) in one submission.  Devices with the
// same padded block size Np share one k_transmission launch: a job reads its device's
// SysView (basis, V-hat, N_z, 2 Delta^2) through jdev, and the host hands jobs to
// workgroups longest-first (cost N_z) so that uneven N_z finish together.  The launches
// of the different Np run on forked streams captured into one hipGraph; negf_multi_replay
// launches that graph again (after negf_ctx_set_potential, for instance).
namespace {

enum Variant { VAR_MFMA = 0, VAR_VALU = 1 };
bool variant_ok(int v, int Np) { (void)v; return Np <= NEGF_MAXNP; }
int variant_threads(int v, int Np) { return v == VAR_VALU ? std::min(1024, round_up(Np, 64)) : threads_for_shape(Np); }
int variant_pb(int v) { return v == VAR_VALU ? 16 : 32; }
const void *variant_kernel(int v, int Np) {
    if (v == VAR_VALU) return (const void *)k_transmission<1024, 16, false, true>;
    return trans_kernel<true>(Np);
}
void launch_variant(int v, int Np, int slots, hipStream_t st, const TransArgs &a) {
    const int th = variant_threads(v, Np);
    if (v == VAR_VALU) hipLaunchKernelGGL((k_transmission<1024, 16, false, true>), dim3(slots), dim3(th), 0, st, a);
    else launch_trans<true>(Np, slots, th, st, a);
}

// Work per slice of one job for the CU shares of an uneven batch: Np^3, with Np counted as
// at least 160.  Below that the sweep is bound by the latency of the pivot steps, not by
// arithmetic, so a plain Np^3 share left the small-block groups a few workgroups each, and
// they finished last (48 devices of N 61-509: 28.5 s; with the floor at 128 / 160 / 224:
// 17.8 / 15.2 / 15.1 s).
double share_cost(int Np) { return std::pow((double)std::max(Np, 160), 3); }

// Variant picked per Np by the probe of tune_group, shared by every negf_multi.
std::mutex g_tune_mu;
std::map<int, int> g_tuned;
std::map<int, std::string> g_tune_log;

}  // namespace

struct negf_multi {
    std::vector<negf_ctx *> ctxs;
    std::vector<int> npv;                  // distinct Np, one stream each
    int device = 0, n_cu = 0;
    hipStream_t s0 = nullptr;
    std::vector<hipStream_t> gs;
    std::vector<hipEvent_t> gev;
    hipEvent_t fork = nullptr, ev0 = nullptr, ev1 = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<void *> bufs;              // device tables of the last call
    struct Group {
        int Np = 0, variant = 0, slots = 0, njobs = 0;
        double *d_tau = nullptr;
        int *d_st = nullptr;
        std::vector<int> dev, je, jk;      // per job: device, energy and k index
    };
    std::vector<Group> groups;
    std::vector<std::vector<double>> kw;   // per device k weights of the last call
    double last_s = 0.0;
    // LDOS at selected slices (negf_multi_ldos): its own graph and buffers
    hipGraph_t lgraph = nullptr;
    hipGraphExec_t lexec = nullptr;
    std::vector<void *> lbufs;
};

namespace {

template <class T>
int mbuf(negf_multi *m, T **p, size_t count) {
    int st = dmalloc(p, count);
    if (st == 0) m->bufs.push_back((void *)*p);
    return st;
}
template <class T>
int mupload(negf_multi *m, T **p, const std::vector<T> &h) {
    int st = mbuf(m, p, h.size());
    if (st) return st;
    if (!h.empty()) HIPCHK(hipMemcpyAsync(*p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, m->s0));
    return 0;
}
void multi_release_ldos(negf_multi *m) {
    if (m->s0) (void)hipStreamSynchronize(m->s0);
    if (m->lexec) (void)hipGraphExecDestroy(m->lexec);
    if (m->lgraph) (void)hipGraphDestroy(m->lgraph);
    m->lexec = nullptr;
    m->lgraph = nullptr;
    for (void *p : m->lbufs) (void)hipFree(p);
    m->lbufs.clear();
}

// Zero n complex entries (a kernel rather than a memset node, so it can sit in a graph).
__global__ __launch_bounds__(256) void k_zero_cd(cd *p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = make_double2(0.0, 0.0);
}

void multi_release(negf_multi *m) {
    if (m->s0) (void)hipStreamSynchronize(m->s0);
    if (m->exec) (void)hipGraphExecDestroy(m->exec);
    if (m->graph) (void)hipGraphDestroy(m->graph);
    m->exec = nullptr;
    m->graph = nullptr;
    for (void *p : m->bufs) (void)hipFree(p);
    m->bufs.clear();
    m->groups.clear();
}

// Probe each valid variant on the group's first jobs with every device cut to its first
// 8 slices (same Np, same per-slice work) and keep the fastest.  NEGF_MULTI_VARIANT
// (mfma | valu) forces a choice.
int tune_group(negf_multi *m, negf_multi::Group &g, const TransArgs &a, const std::vector<SysView> &svs,
               int nLmax) {
    if (const char *f = std::getenv("NEGF_MULTI_VARIANT")) {
        const int v = std::string(f) == "valu" ? VAR_VALU : VAR_MFMA;
        if (!variant_ok(v, g.Np)) return fail(-2, "NEGF_MULTI_VARIANT names a variant that cannot run this Np");
        g.variant = v;
        return 0;
    }
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tuned.find(g.Np);
        if (it != g_tuned.end()) {
            g.variant = it->second;
            return 0;
        }
    }
    std::vector<int> cand;
    for (int v : {VAR_MFMA, VAR_VALU})
        if (variant_ok(v, g.Np)) cand.push_back(v);
    int best = cand[0];
    std::string log = "Np=" + std::to_string(g.Np) + ":";
    if (cand.size() > 1) {
        std::vector<SysView> cut = svs;
        for (auto &sv : cut) sv.Nz = std::min(sv.Nz, 8);
        SysView *d_cut;
        int st;
        std::vector<void *> keep;
        if ((st = dmalloc(&d_cut, cut.size()))) return st;
        keep.push_back(d_cut);
        auto free_all = [&]() {
            (void)hipStreamSynchronize(m->s0);
            for (void *p : keep) (void)hipFree(p);
        };
        if (hipMemcpyAsync(d_cut, cut.data(), cut.size() * sizeof(SysView), hipMemcpyHostToDevice, m->s0) !=
            hipSuccess) {
            free_all();
            return fail(-11, "autotune probe upload failed");
        }
        const int probe_jobs = std::min(g.njobs, m->n_cu);
        std::vector<int> poff(probe_jobs + 1), pjob(probe_jobs);
        for (int j = 0; j <= probe_jobs; ++j) poff[j] = j;
        for (int j = 0; j < probe_jobs; ++j) pjob[j] = j;
        int *d_poff, *d_pjob;
        if ((st = dmalloc(&d_poff, poff.size())) || (keep.push_back(d_poff), false) ||
            (st = dmalloc(&d_pjob, pjob.size())) || (keep.push_back(d_pjob), false)) {
            free_all();
            return st;
        }
        if (hipMemcpyAsync(d_poff, poff.data(), poff.size() * sizeof(int), hipMemcpyHostToDevice, m->s0) != hipSuccess ||
            hipMemcpyAsync(d_pjob, pjob.data(), pjob.size() * sizeof(int), hipMemcpyHostToDevice, m->s0) != hipSuccess) {
            free_all();
            return fail(-11, "autotune probe upload failed");
        }
        double best_ms = 1e300;
        for (int v : cand) {
            const size_t per_slot = ((size_t)2 * g.Np * g.Np + (size_t)2 * nLmax * g.Np +
                                     (size_t)2 * variant_pb(v) * g.Np) * sizeof(cd);
            cd *ws;
            double *tau;
            int *sts;
            if ((st = dmalloc(&ws, per_slot / sizeof(cd) * probe_jobs)) || (keep.push_back(ws), false) ||
                (st = dmalloc(&tau, probe_jobs)) || (keep.push_back(tau), false) || (st = dmalloc(&sts, probe_jobs)) ||
                (keep.push_back(sts), false)) {
                free_all();
                return st;
            }
            TransArgs p = a;
            p.svs = d_cut;
            p.sched_off = d_poff;
            p.sched_job = d_pjob;
            p.njobs = probe_jobs;
            p.ws = ws;
            p.tau = tau;
            p.status = sts;
            float ms = 0.f;
            for (int rep = 0; rep < 2; ++rep) {       // first run warms up, second is timed
                if (hipEventRecord(m->ev0, m->s0) != hipSuccess) break;
                launch_variant(v, g.Np, probe_jobs, m->s0, p);
                if (hipGetLastError() != hipSuccess || hipEventRecord(m->ev1, m->s0) != hipSuccess ||
                    hipEventSynchronize(m->ev1) != hipSuccess || hipEventElapsedTime(&ms, m->ev0, m->ev1) != hipSuccess) {
                    free_all();
                    return fail(-11, "autotune probe launch failed");
                }
            }
            log += std::string(v == VAR_MFMA ? " mfma " : " valu ") + std::to_string(ms) + " ms";
            if (ms < best_ms) {
                best_ms = ms;
                best = v;
            }
        }
        free_all();
    }
    log += std::string(" -> ") + (best == VAR_MFMA ? "mfma" : "valu");
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tuned[g.Np] = best;
    g_tune_log[g.Np] = log;
    g.variant = best;
    return 0;
}

int multi_collect(negf_multi *m, double *const *tau) {
    HIPCHK(hipEventSynchronize(m->ev1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, m->ev0, m->ev1));
    m->last_s = ms * 1e-3;
    int status = 0;
    for (auto &g : m->groups) {
        std::vector<double> h_tau(g.njobs);
        std::vector<int> h_st(g.njobs);
        HIPCHK(hipMemcpy(h_tau.data(), g.d_tau, g.njobs * sizeof(double), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(h_st.data(), g.d_st, g.njobs * sizeof(int), hipMemcpyDeviceToHost));
        for (int j = 0; j < g.njobs; ++j) {
            const int d = g.dev[j];
            tau[d][g.je[j]] += h_tau[j] * m->kw[d][g.jk[j]];
            if (h_st[j] > 0 && !status) status = h_st[j];
        }
    }
    return status;
}

}  // namespace

extern "C" {

int negf_ctx_set_potential(negf_ctx *c, const double *vrec) {
    int st = check_ctx(c);
    if (st) return st;
    if (!vrec) return fail(-2, "vrec is null");
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    const size_t nv = (size_t)(2 * c->NGX + 1) * (2 * c->NGY + 1) * c->Nz;
    HIPCHK(hipMemcpy(c->d_vrec, vrec, nv * sizeof(cd), hipMemcpyHostToDevice));
    return 0;
}

int negf_multi_create(negf_multi **out, int ndev, negf_ctx *const *ctxs) {
    if (!out) return fail(-1, "out is null");
    *out = nullptr;
    if (ndev < 1 || !ctxs) return fail(-2, "need at least one context");
    for (int d = 0; d < ndev; ++d) {
        if (!ctxs[d]) return fail(-1, "null context in the list");
        if (ctxs[d]->device != ctxs[0]->device) return fail(-2, "contexts of one negf_multi must share a GPU");
        if ((ctxs[d]->flags & 3) != NEGF_PREC_FP64) return fail(-2, "uneven batches run the FP64 path only");
    }
    auto *m = new negf_multi();
    m->ctxs.assign(ctxs, ctxs + ndev);
    m->device = ctxs[0]->device;
    m->n_cu = ctxs[0]->n_cu;
    for (auto *c : m->ctxs)
        if (std::find(m->npv.begin(), m->npv.end(), c->Np) == m->npv.end()) m->npv.push_back(c->Np);
    std::sort(m->npv.begin(), m->npv.end());
    auto bail = [&](int st) {
        negf_multi_destroy(m);
        return st;
    };
    if (hipSetDevice(m->device) != hipSuccess) return bail(fail(-8, "hipSetDevice failed"));
    if (hipStreamCreate(&m->s0) != hipSuccess) return bail(fail(-9, "hipStreamCreate failed"));
    for (size_t g = 0; g < m->npv.size(); ++g) {
        hipStream_t s;
        hipEvent_t e;
        if (hipStreamCreate(&s) != hipSuccess) return bail(fail(-9, "hipStreamCreate failed"));
        m->gs.push_back(s);
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return bail(fail(-9, "hipEventCreate failed"));
        m->gev.push_back(e);
    }
    if (hipEventCreateWithFlags(&m->fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreate(&m->ev0) != hipSuccess || hipEventCreate(&m->ev1) != hipSuccess)
        return bail(fail(-9, "hipEventCreate failed"));
    *out = m;
    return 0;
}

int negf_multi_destroy(negf_multi *m) {
    if (!m) return 0;
    multi_release(m);
    multi_release_ldos(m);
    for (auto e : m->gev) (void)hipEventDestroy(e);
    for (auto s : m->gs) (void)hipStreamDestroy(s);
    if (m->fork) (void)hipEventDestroy(m->fork);
    if (m->ev0) (void)hipEventDestroy(m->ev0);
    if (m->ev1) (void)hipEventDestroy(m->ev1);
    if (m->s0) (void)hipStreamDestroy(m->s0);
    delete m;
    return 0;
}

int negf_multi_transmission(negf_multi *m, const double *const *E, const int *nE, const double *const *kx,
                            const double *const *ky, const double *const *kw, const int *nk,
                            double *const *tau) {
    if (!m) return fail(-1, "null negf_multi");
    const int ndev = (int)m->ctxs.size();
    if (!E || !nE || !kx || !ky || !kw || !nk || !tau) return fail(-2, "null argument to negf_multi_transmission");
    for (int d = 0; d < ndev; ++d)
        if (nE[d] < 0 || nk[d] < 1 || (nE[d] > 0 && (!E[d] || !tau[d])) || !kx[d] || !ky[d] || !kw[d])
            return fail(-2, "bad per-device arguments to negf_multi_transmission");
    HIPCHK(hipSetDevice(m->device));
    multi_release(m);
    int st;
    m->kw.assign(ndev, {});
    // per device job tables, as negf_transmission_batch builds them
    std::vector<JobTables> tabs(ndev);
    std::vector<std::vector<int>> dje(ndev), djk(ndev);
    for (int d = 0; d < ndev; ++d) {
        negf_ctx *c = m->ctxs[d];
        m->kw[d].assign(kw[d], kw[d] + nk[d]);
        const int nj = nE[d] * nk[d];
        std::vector<std::complex<double>> energies(nE[d]);
        std::vector<double> realE(E[d], E[d] + nE[d]);
        for (int i = 0; i < nE[d]; ++i) energies[i] = E[d][i];
        dje[d].resize(nj);
        djk[d].resize(nj);
        for (int j = 0; j < nj; ++j) {
            dje[d][j] = j % nE[d];
            djk[d][j] = j / nE[d];
        }
        if ((st = job_tables(c, energies, realE, kx[d], ky[d], nk[d], dje[d], djk[d], true, tabs[d]))) return st;
    }
    // one group per Np; work estimate N_z share_cost(Np) per job gives each group its share of CUs
    const int ng = (int)m->npv.size();
    std::vector<double> work(ng, 0.0);
    for (int d = 0; d < ndev; ++d) {
        const int g = (int)(std::find(m->npv.begin(), m->npv.end(), m->ctxs[d]->Np) - m->npv.begin());
        work[g] += (double)nE[d] * nk[d] * m->ctxs[d]->Nz * share_cost(m->ctxs[d]->Np);
    }
    double wsum = 0.0;
    for (double w : work) wsum += w;
    std::vector<TransArgs> args(ng);
    for (int gi = 0; gi < ng; ++gi) {
        negf_multi::Group g;
        g.Np = m->npv[gi];
        const int Np = g.Np;
        std::vector<SysView> svs;
        std::vector<cd> z, sl, sr;
        std::vector<double> kinh, gLv, gRv;
        std::vector<int> kidx, jdev, off, nL, nR, Lrow, Rcol;
        std::vector<double> cost;
        int nLmax = 1;
        for (int d = 0; d < ndev; ++d) {
            negf_ctx *c = m->ctxs[d];
            if (c->Np != Np) continue;
            const JobTables &t = tabs[d];
            const int dv = (int)svs.size(), krow = (int)(kinh.size() / Np);
            svs.push_back(sysview(c));
            kinh.insert(kinh.end(), t.kinh.begin(), t.kinh.end());
            for (size_t j = 0; j < t.z.size(); ++j) {
                g.dev.push_back(d);
                g.je.push_back(dje[d][j]);
                g.jk.push_back(djk[d][j]);
                jdev.push_back(dv);
                z.push_back(t.z[j]);
                kidx.push_back(krow + t.kidx[j]);
                sl.insert(sl.end(), t.sigL.begin() + j * Np, t.sigL.begin() + (j + 1) * Np);
                sr.insert(sr.end(), t.sigR.begin() + j * Np, t.sigR.begin() + (j + 1) * Np);
                off.push_back((int)Lrow.size());
                std::vector<int> rc;
                std::vector<double> rg;
                int l = 0, r = 0;
                for (int i = 0; i < c->N; ++i) {
                    const double gl = t.gamL[j * Np + i], gr = t.gamR[j * Np + i];
                    if (gl != 0.0) { Lrow.push_back(i); gLv.push_back(gl); ++l; }
                    if (gr != 0.0) { rc.push_back(i); rg.push_back(gr); ++r; }
                }
                const int n = std::max(l, r);
                for (int q = l; q < n; ++q) { Lrow.push_back(0); gLv.push_back(0.0); }
                for (int q = 0; q < n; ++q) {
                    Rcol.push_back(q < r ? rc[q] : 0);
                    gRv.push_back(q < r ? rg[q] : 0.0);
                }
                nL.push_back(l);
                nR.push_back(r);
                nLmax = std::max(nLmax, l);
                cost.push_back((l == 0 || r == 0) ? 0.0 : (double)c->Nz);
            }
        }
        if (Lrow.empty()) { Lrow.push_back(0); Rcol.push_back(0); gLv.push_back(0); gRv.push_back(0); }
        g.njobs = (int)z.size();
        TransArgs &a = args[gi];
        a = TransArgs{};
        a.sv = svs[0];
        a.njobs = g.njobs;
        a.nLmax = nLmax;
        if (g.njobs == 0) {
            m->groups.push_back(g);
            continue;
        }
        SysView *d_svs;
        cd *d_z, *d_sl, *d_sr;
        int *d_k, *d_jd, *d_off, *d_nL, *d_nR, *d_Lr, *d_Rc;
        double *d_kin, *d_gl, *d_gr;
        if ((st = mupload(m, &d_svs, svs)) || (st = mupload(m, &d_z, z)) || (st = mupload(m, &d_k, kidx)) ||
            (st = mupload(m, &d_jd, jdev)) || (st = mupload(m, &d_kin, kinh)) || (st = mupload(m, &d_sl, sl)) ||
            (st = mupload(m, &d_sr, sr)) || (st = mupload(m, &d_off, off)) || (st = mupload(m, &d_nL, nL)) ||
            (st = mupload(m, &d_nR, nR)) || (st = mupload(m, &d_Lr, Lrow)) || (st = mupload(m, &d_Rc, Rcol)) ||
            (st = mupload(m, &d_gl, gLv)) || (st = mupload(m, &d_gr, gRv)) || (st = mbuf(m, &g.d_tau, g.njobs)) ||
            (st = mbuf(m, &g.d_st, g.njobs)))
            return st;
        a.svs = d_svs;
        a.jdev = d_jd;
        a.z = d_z;
        a.kidx = d_k;
        a.kinh = d_kin;
        a.sigL = d_sl;
        a.sigR = d_sr;
        a.open_off = d_off;
        a.nL = d_nL;
        a.nR = d_nR;
        a.Lrow = d_Lr;
        a.Rcol = d_Rc;
        a.gL = d_gl;
        a.gR = d_gr;
        a.tau = g.d_tau;
        a.status = g.d_st;
        if ((st = tune_group(m, g, a, svs, nLmax))) return st;
        // workgroups: this group's share of the CUs, at most one per job
        int per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, variant_kernel(g.variant, Np),
                                                         variant_threads(g.variant, Np), 0) != hipSuccess || per_cu < 1)
            per_cu = 1;
        const double share = wsum > 0 ? work[gi] / wsum : 1.0;
        g.slots = std::max(1, std::min(g.njobs, (int)std::lround(share * m->n_cu * per_cu)));
        // longest job first onto the least-loaded workgroup
        std::vector<int> order(g.njobs);
        for (int j = 0; j < g.njobs; ++j) order[j] = j;
        std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return cost[x] > cost[y]; });
        typedef std::pair<double, int> Load;
        std::priority_queue<Load, std::vector<Load>, std::greater<Load>> heap;
        for (int b = 0; b < g.slots; ++b) heap.push(Load(0.0, b));
        std::vector<std::vector<int>> lists(g.slots);
        for (int j : order) {
            Load l = heap.top();
            heap.pop();
            lists[l.second].push_back(j);
            heap.push(Load(l.first + cost[j] + 1e-9, l.second));
        }
        std::vector<int> soff(1, 0), sjob;
        for (auto &l : lists) {
            sjob.insert(sjob.end(), l.begin(), l.end());
            soff.push_back((int)sjob.size());
        }
        const size_t per_slot = (size_t)2 * Np * Np + (size_t)2 * nLmax * Np + (size_t)2 * variant_pb(g.variant) * Np;
        int *d_so, *d_sj;
        cd *d_ws;
        if ((st = mupload(m, &d_so, soff)) || (st = mupload(m, &d_sj, sjob)) ||
            (st = mbuf(m, &d_ws, per_slot * g.slots)))
            return st;
        a.sched_off = d_so;
        a.sched_job = d_sj;
        a.ws = d_ws;
        m->groups.push_back(g);
    }
    HIPCHK(hipStreamSynchronize(m->s0));
    // one graph: fork from s0, one launch per Np on its own stream, join into s0
    HIPCHK(hipStreamBeginCapture(m->s0, hipStreamCaptureModeThreadLocal));
    bool cap_ok = hipEventRecord(m->fork, m->s0) == hipSuccess;
    for (int gi = 0; gi < ng && cap_ok; ++gi) {
        const auto &g = m->groups[gi];
        if (g.njobs == 0) continue;
        cap_ok = hipStreamWaitEvent(m->gs[gi], m->fork, 0) == hipSuccess;
        if (!cap_ok) break;
        launch_variant(g.variant, g.Np, g.slots, m->gs[gi], args[gi]);
        cap_ok = hipGetLastError() == hipSuccess && hipEventRecord(m->gev[gi], m->gs[gi]) == hipSuccess &&
                 hipStreamWaitEvent(m->s0, m->gev[gi], 0) == hipSuccess;
    }
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(m->s0, &graph);
    if (!cap_ok || ec != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        return fail(-12, "hipGraph capture of the uneven batch failed");
    }
    m->graph = graph;
    HIPCHK(hipGraphInstantiate(&m->exec, m->graph, nullptr, nullptr, 0));
    HIPCHK(hipEventRecord(m->ev0, m->s0));
    HIPCHK(hipGraphLaunch(m->exec, m->s0));
    HIPCHK(hipEventRecord(m->ev1, m->s0));
    return multi_collect(m, tau);
}

int negf_multi_ldos(negf_multi *m, const double *const *E, const int *nE, const double *const *kx,
                    const double *const *ky, const double *const *kw, const int *nk, int nsel,
                    const int *const *slices, double *const *ldos) {
    if (!m) return fail(-1, "null negf_multi");
    const int ndev = (int)m->ctxs.size();
    if (!E || !nE || !kx || !ky || !kw || !nk || !slices || !ldos || nsel < 1)
        return fail(-2, "null argument to negf_multi_ldos");
    for (int d = 0; d < ndev; ++d) {
        if (nE[d] < 0 || nk[d] < 1 || (nE[d] > 0 && (!E[d] || !ldos[d])) || !kx[d] || !ky[d] || !kw[d] || !slices[d])
            return fail(-2, "bad per-device arguments to negf_multi_ldos");
        for (int q = 0; q < nsel; ++q)
            if (slices[d][q] < 0 || slices[d][q] >= m->ctxs[d]->Nz) return fail(-2, "slice index out of range");
    }
    HIPCHK(hipSetDevice(m->device));
    multi_release_ldos(m);
    int st;
    auto lbuf = [&](auto **p, size_t count) {
        int r = dmalloc(p, count);
        if (r == 0) m->lbufs.push_back((void *)*p);
        return r;
    };
    auto lup = [&](auto **p, const auto &h) {
        int r = lbuf(p, h.size());
        if (r) return r;
        if (!h.empty())
            HIPCHK(hipMemcpyAsync(*p, h.data(), h.size() * sizeof(h[0]), hipMemcpyHostToDevice, m->s0));
        return 0;
    };
    std::vector<JobTables> tabs(ndev);
    std::vector<std::vector<int>> dje(ndev), djk(ndev);
    for (int d = 0; d < ndev; ++d) {
        negf_ctx *c = m->ctxs[d];
        const int nj = nE[d] * nk[d];
        std::vector<std::complex<double>> energies(nE[d]);
        std::vector<double> realE(E[d], E[d] + nE[d]);
        for (int i = 0; i < nE[d]; ++i) energies[i] = E[d][i];
        for (int j = 0; j < nj; ++j) {      // jobs in negf_ldos_batch's order
            dje[d].push_back(j % std::max(1, nE[d]));
            djk[d].push_back(j / std::max(1, nE[d]));
        }
        if ((st = job_tables(c, energies, realE, kx[d], ky[d], nk[d], dje[d], djk[d], false, tabs[d]))) return st;
    }
    // device outputs: N x N x nsel x nE complex, zeroed inside the graph
    std::vector<cd *> d_out(ndev, nullptr);
    std::vector<size_t> nout(ndev, 0);
    for (int d = 0; d < ndev; ++d) {
        nout[d] = (size_t)m->ctxs[d]->N * m->ctxs[d]->N * nsel * std::max(0, nE[d]);
        if ((st = lbuf(&d_out[d], std::max<size_t>(nout[d], 1)))) return st;
    }
    const int ng = (int)m->npv.size();
    std::vector<double> work(ng, 0.0);
    for (int d = 0; d < ndev; ++d) {
        const int g = (int)(std::find(m->npv.begin(), m->npv.end(), m->ctxs[d]->Np) - m->npv.begin());
        work[g] += (double)nE[d] * nk[d] * m->ctxs[d]->Nz * share_cost(m->ctxs[d]->Np);
    }
    double wsum = 0.0;
    for (double w : work) wsum += w;
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = 0;
    struct LGroup {
        int Np, slots, njobs;
        DiagArgs a;
        std::vector<int> devs;
    };
    std::vector<LGroup> groups;
    for (int gi = 0; gi < ng; ++gi) {
        LGroup g;
        g.Np = m->npv[gi];
        const int Np = g.Np;
        std::vector<SysView> svs;
        std::vector<cd> z, sl, sr, wg;
        std::vector<double> kinh, cost;
        std::vector<int> kidx, jdev, tgt, nsv, smap, smoff;
        int maxNz = 0;
        for (int d = 0; d < ndev; ++d) {
            negf_ctx *c = m->ctxs[d];
            if (c->Np != Np) continue;
            g.devs.push_back(d);
            const JobTables &t = tabs[d];
            const int dv = (int)svs.size(), krow = (int)(kinh.size() / Np);
            svs.push_back(sysview(c));
            maxNz = std::max(maxNz, c->Nz);
            kinh.insert(kinh.end(), t.kinh.begin(), t.kinh.end());
            nsv.push_back(nsel);
            smoff.push_back((int)smap.size());
            std::vector<int> mp(c->Nz, -1);
            for (int q = 0; q < nsel; ++q) mp[slices[d][q]] = q;
            smap.insert(smap.end(), mp.begin(), mp.end());
            const double delta = c->Lz / c->Nz, sc = 2 * delta * delta;
            for (size_t j = 0; j < t.z.size(); ++j) {
                jdev.push_back(dv);
                z.push_back(t.z[j]);
                kidx.push_back(krow + t.kidx[j]);
                sl.insert(sl.end(), t.sigL.begin() + j * Np, t.sigL.begin() + (j + 1) * Np);
                sr.insert(sr.end(), t.sigR.begin() + j * Np, t.sigR.begin() + (j + 1) * Np);
                wg.push_back(make_double2(-kw[d][djk[d][j]] / PI * sc, 0.0));   // src/negf.f90:230-231
                tgt.push_back(dje[d][j]);
                cost.push_back((double)c->Nz);
            }
        }
        g.njobs = (int)z.size();
        g.slots = 0;
        DiagArgs &a = g.a;
        a = DiagArgs{};
        if (g.njobs > 0) {
            a.sv = svs[0];
            a.sv.Nz = maxNz;
            a.njobs = g.njobs;
            // per-device output pointers, indexed by the group-local device index
            std::vector<cd *> go;
            for (int d : g.devs) go.push_back(d_out[d]);
            SysView *d_svs;
            cd *d_z, *d_sl, *d_sr, *d_w;
            cd **d_go;
            int *d_k, *d_jd, *d_t, *d_ns, *d_sm, *d_so2, *d_st;
            double *d_kin;
            if ((st = lup(&d_svs, svs)) || (st = lup(&d_z, z)) || (st = lup(&d_k, kidx)) || (st = lup(&d_jd, jdev)) ||
                (st = lup(&d_kin, kinh)) || (st = lup(&d_sl, sl)) || (st = lup(&d_sr, sr)) || (st = lup(&d_w, wg)) ||
                (st = lup(&d_t, tgt)) || (st = lup(&d_go, go)) || (st = lup(&d_ns, nsv)) || (st = lup(&d_sm, smap)) ||
                (st = lup(&d_so2, smoff)) || (st = lbuf(&d_st, g.njobs)))
                return st;
            a.svs = d_svs;
            a.jdev = d_jd;
            a.z = d_z;
            a.kidx = d_k;
            a.kinh = d_kin;
            a.sigL = d_sl;
            a.sigR = d_sr;
            a.wgt = d_w;
            a.target = d_t;
            a.outs = d_go;
            a.nsel = d_ns;
            a.smap = d_sm;
            a.smap_off = d_so2;
            a.status = d_st;
            const int pb = panel_width(Np);
            a.ws_stride = (size_t)(maxNz + 2) * Np * Np + (size_t)2 * pb * Np;
            const Shape shp = shape_for(Np);
            const void *kfn = shp == SHAPE_640 ? (const void *)k_diagonal<640, 32, true, true>
                            : shp == SHAPE_1024_TWO ? (const void *)k_diagonal<1024, 32, true, true>
                                                    : (const void *)k_diagonal<1024, 16, true, true>;
            int per_cu = 0;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kfn, threads_for_shape(Np), 0) != hipSuccess ||
                per_cu < 1)
                per_cu = 1;
            const double share = wsum > 0 ? work[gi] / wsum : 1.0;
            long slots = std::max(1L, std::lround(share * m->n_cu * per_cu));
            if (free_b > 0) {
                const long fit = (long)(share * 0.7 * (double)free_b / (double)(a.ws_stride * sizeof(cd)));
                slots = std::min(slots, std::max(1L, fit));
            }
            g.slots = (int)std::min<long>(slots, g.njobs);
            // longest job first onto the least-loaded workgroup
            std::vector<int> order(g.njobs);
            for (int j = 0; j < g.njobs; ++j) order[j] = j;
            std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return cost[x] > cost[y]; });
            typedef std::pair<double, int> Load;
            std::priority_queue<Load, std::vector<Load>, std::greater<Load>> heap;
            for (int b = 0; b < g.slots; ++b) heap.push(Load(0.0, b));
            std::vector<std::vector<int>> lists(g.slots);
            for (int j : order) {
                Load l = heap.top();
                heap.pop();
                lists[l.second].push_back(j);
                heap.push(Load(l.first + cost[j], l.second));
            }
            std::vector<int> soff(1, 0), sjob;
            for (auto &l : lists) {
                sjob.insert(sjob.end(), l.begin(), l.end());
                soff.push_back((int)sjob.size());
            }
            int *d_soff, *d_sjob;
            cd *d_ws;
            if ((st = lup(&d_soff, soff)) || (st = lup(&d_sjob, sjob)) || (st = lbuf(&d_ws, a.ws_stride * g.slots)))
                return st;
            a.sched_off = d_soff;
            a.sched_job = d_sjob;
            a.ws = d_ws;
        }
        groups.push_back(g);
    }
    HIPCHK(hipStreamSynchronize(m->s0));
    HIPCHK(hipStreamBeginCapture(m->s0, hipStreamCaptureModeThreadLocal));
    bool ok = hipEventRecord(m->fork, m->s0) == hipSuccess;
    for (int gi = 0; gi < ng && ok; ++gi) {
        const LGroup &g = groups[gi];
        if (g.njobs == 0) continue;
        hipStream_t gsm = m->gs[gi];
        ok = hipStreamWaitEvent(gsm, m->fork, 0) == hipSuccess;
        if (!ok) break;
        for (int d : g.devs)
            if (nout[d]) hipLaunchKernelGGL(k_zero_cd, dim3(256), dim3(256), 0, gsm, d_out[d], nout[d]);
        const int th = threads_for_shape(g.Np);
        switch (shape_for(g.Np)) {
        case SHAPE_640:
            hipLaunchKernelGGL((k_diagonal<640, 32, true, true>), dim3(g.slots), dim3(th), 0, gsm, g.a);
            break;
        case SHAPE_1024_TWO:
            hipLaunchKernelGGL((k_diagonal<1024, 32, true, true>), dim3(g.slots), dim3(th), 0, gsm, g.a);
            break;
        default:
            hipLaunchKernelGGL((k_diagonal<1024, 16, true, true>), dim3(g.slots), dim3(th), 0, gsm, g.a);
        }
        ok = hipGetLastError() == hipSuccess && hipEventRecord(m->gev[gi], gsm) == hipSuccess &&
             hipStreamWaitEvent(m->s0, m->gev[gi], 0) == hipSuccess;
    }
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(m->s0, &graph);
    if (!ok || ec != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        return fail(-12, "hipGraph capture of the uneven LDOS batch failed");
    }
    m->lgraph = graph;
    HIPCHK(hipGraphInstantiate(&m->lexec, m->lgraph, nullptr, nullptr, 0));
    HIPCHK(hipEventRecord(m->ev0, m->s0));
    HIPCHK(hipGraphLaunch(m->lexec, m->s0));
    HIPCHK(hipEventRecord(m->ev1, m->s0));
    HIPCHK(hipEventSynchronize(m->ev1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, m->ev0, m->ev1));
    m->last_s = ms * 1e-3;
    int status = 0;
    for (const LGroup &g : groups) {
        if (g.njobs == 0) continue;
        std::vector<int> h_st(g.njobs);
        HIPCHK(hipMemcpy(h_st.data(), g.a.status, g.njobs * sizeof(int), hipMemcpyDeviceToHost));
        for (int v : h_st)
            if (v > 0 && !status) status = v;
    }
    for (int d = 0; d < ndev; ++d)
        if (nout[d]) HIPCHK(hipMemcpy(ldos[d], d_out[d], nout[d] * sizeof(cd), hipMemcpyDeviceToHost));
    return status;
}

int negf_multi_replay(negf_multi *m, double *const *tau) {
    if (!m) return fail(-1, "null negf_multi");
    if (!m->exec) return fail(-2, "negf_multi_replay needs a previous negf_multi_transmission");
    if (!tau) return fail(-2, "tau is null");
    HIPCHK(hipSetDevice(m->device));
    HIPCHK(hipEventRecord(m->ev0, m->s0));
    HIPCHK(hipGraphLaunch(m->exec, m->s0));
    HIPCHK(hipEventRecord(m->ev1, m->s0));
    return multi_collect(m, tau);
}

double negf_multi_last_seconds(const negf_multi *m) { return m ? m->last_s : -1.0; }

int negf_multi_variant(int Np) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tuned.find(Np);
    return it == g_tuned.end() ? -1 : it->second;
}

}  // extern "C"
