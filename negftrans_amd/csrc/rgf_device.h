// rgf_device.h -- device building blocks of the RGF sweep (gfx950, wave64).
//
// Matrices are complex128 (double2), column-major, leading dimension Np (N padded to a
// multiple of 16; the padding block is the identity and stays inert through every
// operation).  One workgroup owns one matrix at a time; thread r owns row r (blockDim.x
// = Np rounded up to 64, at most 1024).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <utility>

#define NEGF_PB 16   // Gauss-Jordan panel width (columns factored per pass)
#define NEGF_CB 8    // column block of the trailing update

typedef double2 cd;

__device__ __forceinline__ cd cmk(double re, double im) { return make_double2(re, im); }
__device__ __forceinline__ cd cadd(cd a, cd b) { return cmk(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ cd csub(cd a, cd b) { return cmk(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ cd cmul(cd a, cd b) {
    return cmk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// c + a*b
__device__ __forceinline__ cd cfma(cd a, cd b, cd c) {
    c.x = fma(a.x, b.x, c.x);
    c.x = fma(-a.y, b.y, c.x);
    c.y = fma(a.x, b.y, c.y);
    c.y = fma(a.y, b.x, c.y);
    return c;
}
// c - a*b
__device__ __forceinline__ cd cfms(cd a, cd b, cd c) {
    c.x = fma(-a.x, b.x, c.x);
    c.x = fma(a.y, b.y, c.x);
    c.y = fma(-a.x, b.y, c.y);
    c.y = fma(-a.y, b.x, c.y);
    return c;
}
// 1/b with scaling (Smith's method), as Fortran complex division does.
__device__ __forceinline__ cd crecip(cd b) {
    if (fabs(b.x) >= fabs(b.y)) {
        double r = b.y / b.x, d = b.x + b.y * r;
        return cmk(1.0 / d, -r / d);
    } else {
        double r = b.x / b.y, d = b.x * r + b.y;
        return cmk(r / d, -1.0 / d);
    }
}
// |re| + |im|, the pivot measure of LAPACK izamax (dcabs1).
__device__ __forceinline__ double cabs1(cd a) { return fabs(a.x) + fabs(a.y); }

// Shared-memory scratch of one workgroup.  Sizes are bounded by NEGF_MAXNP.
#define NEGF_MAXNP 1024
struct RgfShared {
    int perm[NEGF_MAXNP];    // row permutation of the current panel (new row r <- old perm[r])
    int ipiv[NEGF_MAXNP];    // pivot row chosen for each column (LAPACK-style interchanges)
    int sigma[NEGF_MAXNP];   // column map of the finished inverse: inv[:, c] = S[:, sigma[c]]
    cd rowA[NEGF_PB];
    cd rowB[NEGF_PB];
    cd rowR[NEGF_PB];
    cd T[NEGF_PB * NEGF_CB];
    double rv[16];
    int ri[16];
    double red[16];
    int flag;
};

// Workgroup argmax of (v, idx): larger v wins, ties go to the smaller index (izamax).
__device__ __forceinline__ int block_argmax(double v, int idx, RgfShared &s) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        double ov = __shfl_xor(v, off, 64);
        int oi = __shfl_xor(idx, off, 64);
        if (ov > v || (ov == v && oi < idx)) { v = ov; idx = oi; }
    }
    if (lane == 0) { s.rv[w] = v; s.ri[w] = idx; }
    __syncthreads();
    double bv = s.rv[0];
    int bi = s.ri[0];
    for (int i = 1; i < nw; ++i) {
        double ov = s.rv[i];
        int oi = s.ri[i];
        if (ov > bv || (ov == bv && oi < bi)) { bv = ov; bi = oi; }
    }
    return bi;
}

__device__ __forceinline__ double block_sum(double v, RgfShared &s) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    __syncthreads();
    if (lane == 0) s.red[w] = v;
    __syncthreads();
    double t = 0.0;
    for (int i = 0; i < nw; ++i) t += s.red[i];
    __syncthreads();
    return t;
}

// One Gauss-Jordan step on the register-resident panel: pivot search over the rows not
// yet pivoted, row interchange, scaling of the pivot row, elimination in every other row.
// J is a template parameter so every access to P is a static register index.
template <int J>
__device__ __forceinline__ void gje_step(cd (&P)[NEGF_PB], const int k0, const int Np,
                                         const bool own, RgfShared &s) {
    const int r = threadIdx.x;
    const int c = k0 + J;
    // eligible rows: r >= c; a NaN entry ranks below every number but above the
    // ineligible rows, so the pivot is always a valid row index in [c, Np)
    double v = -1.0;
    if (own && r >= c) {
        v = cabs1(P[J]);
        if (!(v >= 0.0)) v = -0.5;
    }
    int p = block_argmax(v, own && r >= c ? r : 0x7fffffff, s);
    if (p < c || p >= Np) p = c;
    if (p != c) {
        if (r == p) {
#pragma unroll
            for (int q = 0; q < NEGF_PB; ++q) s.rowA[q] = P[q];
        }
        if (r == c) {
#pragma unroll
            for (int q = 0; q < NEGF_PB; ++q) s.rowB[q] = P[q];
        }
        __syncthreads();
        if (r == c) {
#pragma unroll
            for (int q = 0; q < NEGF_PB; ++q) P[q] = s.rowA[q];
        }
        if (r == p) {
#pragma unroll
            for (int q = 0; q < NEGF_PB; ++q) P[q] = s.rowB[q];
        }
        if (r == 0) {
            int t = s.perm[c];
            s.perm[c] = s.perm[p];
            s.perm[p] = t;
        }
    }
    if (r == 0) s.ipiv[c] = p;
    if (r == c) {
        if (P[J].x == 0.0 && P[J].y == 0.0 && s.flag == 0) s.flag = c + 1;
        const cd inv = crecip(P[J]);
#pragma unroll
        for (int q = 0; q < NEGF_PB; ++q) {
            P[q] = (q == J) ? inv : cmul(P[q], inv);
            s.rowR[q] = P[q];
        }
    }
    __syncthreads();
    if (own && r != c) {
        const cd f = P[J];
#pragma unroll
        for (int q = 0; q < NEGF_PB; ++q) {
            if (q == J) P[q] = cmul(cmk(-f.x, -f.y), s.rowR[J]);
            else P[q] = cfms(f, s.rowR[q], P[q]);
        }
    }
    __syncthreads();
}

template <int... Js>
__device__ __forceinline__ void gje_panel_steps(cd (&P)[NEGF_PB], const int k0, const int Np,
                                                const bool own, RgfShared &s,
                                                std::integer_sequence<int, Js...>) {
    (gje_step<Js>(P, k0, Np, own, s), ...);
}

// In-place inversion of the Np x Np matrix Z by blocked Gauss-Jordan elimination with
// partial (row) pivoting -- the same pivot rule as zgetrf (largest |re|+|im| among the
// rows not yet pivoted, first index on ties), and the same arithmetic count as
// zgetrf+zgetri (8 N^3 real flops).  Rows are interchanged physically, panel by panel;
// the column interchanges that undo them are NOT applied: on return
//     inverse[:, c] = Z[:, s.sigma[c]]
// Returns the 1-based column of the first exactly-zero pivot, or 0.
__device__ __forceinline__ int gje_invert(cd *__restrict__ Z, const int Np, RgfShared &s) {
    const int r = threadIdx.x;
    const bool own = r < Np;
    const size_t ld = (size_t)Np;
    int zero_piv = 0;
    for (int k0 = 0; k0 < Np; k0 += NEGF_PB) {
        cd P[NEGF_PB];
#pragma unroll
        for (int j = 0; j < NEGF_PB; ++j) P[j] = own ? Z[r + (k0 + j) * ld] : cmk(0.0, 0.0);
        for (int i = r; i < Np; i += blockDim.x) s.perm[i] = i;
        __syncthreads();
        gje_panel_steps(P, k0, Np, own, s, std::make_integer_sequence<int, NEGF_PB>{});
        // panel columns are final
        if (own) {
#pragma unroll
            for (int j = 0; j < NEGF_PB; ++j) Z[r + (k0 + j) * ld] = P[j];
        }
        // trailing update of every other column: rows permuted by this panel's
        // interchanges, pivot rows replaced, rank-PB update with the factored panel.
        const bool pivrow = (r >= k0 && r < k0 + NEGF_PB);
        const int pr = own ? s.perm[r] : 0;
        for (int cb = 0; cb < Np; cb += NEGF_CB) {
            if (cb >= k0 && cb < k0 + NEGF_PB) continue;
            for (int e = r; e < NEGF_PB * NEGF_CB; e += blockDim.x) {
                const int jj = e % NEGF_PB, cc = e / NEGF_PB;
                s.T[e] = Z[s.perm[k0 + jj] + (cb + cc) * ld];
            }
            cd a[NEGF_CB];
#pragma unroll
            for (int cc = 0; cc < NEGF_CB; ++cc)
                a[cc] = (own && !pivrow) ? Z[pr + (cb + cc) * ld] : cmk(0.0, 0.0);
            __syncthreads();
            if (own) {
#pragma unroll
                for (int cc = 0; cc < NEGF_CB; ++cc) {
                    cd acc = a[cc];
#pragma unroll
                    for (int jj = 0; jj < NEGF_PB; ++jj) acc = cfma(P[jj], s.T[jj + cc * NEGF_PB], acc);
                    Z[r + (cb + cc) * ld] = acc;
                }
            }
            __syncthreads();
        }
    }
    if (r == 0) {
        for (int c = 0; c < Np; ++c) s.sigma[c] = c;
        for (int c = Np - 1; c >= 0; --c) {
            const int p = s.ipiv[c];
            const int t = s.sigma[c];
            s.sigma[c] = s.sigma[p];
            s.sigma[p] = t;
        }
        zero_piv = s.flag;
    }
    __syncthreads();
    return zero_piv;
}

// C = A * B(:, sigB) (sigB == nullptr: identity), Np x Np, every wave computes 32x32
// tiles with a 4x4 register block per lane.  The epilogue functor receives
// (row, col, value) for every in-range element.
template <class Epi>
__device__ __forceinline__ void gemm_nn(const cd *__restrict__ A, const cd *__restrict__ B, const int *sigB,
                        const int Np, Epi epi) {
    const size_t ld = (size_t)Np;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const int tr = lane >> 3, tc = lane & 7;
    const int td = (Np + 31) / 32;
    for (int t = w; t < td * td; t += nw) {
        const int r0 = (t % td) * 32 + tr * 4, c0 = (t / td) * 32 + tc * 4;
        const cd *bcol[4];
        int rr[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int c = min(c0 + j, Np - 1);
            bcol[j] = B + (size_t)(sigB ? sigB[c] : c) * ld;
            rr[j] = min(r0 + j, Np - 1);
        }
        cd acc[4][4];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[i][j] = cmk(0.0, 0.0);
        for (int k = 0; k < Np; ++k) {
            cd a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = A[rr[i] + k * ld];
#pragma unroll
            for (int j = 0; j < 4; ++j) b[j] = bcol[j][k];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = cfma(a[i], b[j], acc[i][j]);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j)
                if (r0 + i < Np && c0 + j < Np) epi(r0 + i, c0 + j, acc[i][j]);
    }
}

// Per-job constants of one (energy, k-point) job.
struct JobView {
    cd z;                    // E + i eta
    const double *kinh;      // |G_i + k|^2 / 2, length Np (0 in the padding)
    const cd *sigL, *sigR;   // exp(i q^{L,R}_i Delta), length Np
};

// System constants shared by all jobs of a context.
struct SysView {
    int N, Nz, Np, NGX, NGY;
    double scale;            // 2 Delta^2
    const int *nx, *ny;      // basis, length N
    const cd *vrec;          // V_reciprocal_all, (2NGX+1)(2NGY+1)N_z
};

// M = D_slice - S(:, sig) (S == nullptr: M = D_slice); the padding block is the identity.
// D_k = 2 Delta^2 ((E + i eta) I - H_k) - 2 I (+ lead terms on the end slices), in the
// operation order of E_minus_H_construction (src/grid.f90:291-331) with
// H_k(i,j) = V(nx_i - nx_j, ny_i - ny_j, k) + delta_ij kin_i (src/grid.f90:248-259).
__device__ __forceinline__ void build_block(cd *__restrict__ M, const SysView &sv, const JobView &jv,
                            const int slice, const cd *__restrict__ S, const int *sig) {
    const int Np = sv.Np, N = sv.N;
    const int W = 2 * sv.NGX + 1;
    const cd *vs = sv.vrec + (size_t)slice * W * (2 * sv.NGY + 1);
    for (int col = 0; col < Np; ++col) {
        const cd *scol = S ? S + (size_t)(sig ? sig[col] : col) * Np : nullptr;
        for (int r = threadIdx.x; r < Np; r += blockDim.x) {
            cd v;
            if (r < N && col < N) {
                cd h = vs[(sv.nx[r] - sv.nx[col] + sv.NGX) + W * (sv.ny[r] - sv.ny[col] + sv.NGY)];
                if (r == col) h.x += jv.kinh[r];
                cd d = cmk(-h.x, -h.y);
                if (r == col) d = cadd(d, jv.z);
                d.x *= sv.scale;
                d.y *= sv.scale;
                if (r == col) {
                    d.x -= 2.0;
                    if (slice == 0) d = cadd(d, jv.sigL[r]);
                    if (slice == sv.Nz - 1) d = cadd(d, jv.sigR[r]);
                }
                if (scol) d = csub(d, scol[r]);
                v = d;
            } else {
                v = cmk(r == col ? 1.0 : 0.0, 0.0);
            }
            M[r + (size_t)col * Np] = v;
        }
    }
}

// dst = S(:, sig) (column gather), Np x Np.
__device__ __forceinline__ void copy_cols(cd *__restrict__ dst, const cd *__restrict__ S, const int *sig, int Np) {
    for (int col = 0; col < Np; ++col) {
        const cd *sc = S + (size_t)(sig ? sig[col] : col) * Np;
        for (int r = threadIdx.x; r < Np; r += blockDim.x) dst[r + (size_t)col * Np] = sc[r];
    }
}

// ====================================================================================
// Gauss-Jordan inversion, streaming version.
//
// Same pivot rule and arithmetic as gje_invert; reorganised so that the trailing
// update -- the O(N^3) part -- needs no workgroup barrier:
//  * each of the PB pivot steps of a panel takes ONE barrier: every wave publishes its
//    best candidate (value, row index, and the row itself) plus the old content of row c,
//    then every thread reads the winner from LDS; the LDS buffers alternate by step parity.
//  * after the panel, the old pivot rows (all columns) are copied to a per-job scratch
//    area T (PB x Np, column-major so a column block is contiguous) and the rows displaced
//    by the interchanges are moved in place, so the update reads rows in natural order.
//  * every wave then streams the column blocks on its own: it stages the T block in a
//    wave-private LDS area, prefetches the next block into registers while it computes
//    the current one, and writes its rows back.
// ====================================================================================
#ifndef NEGF_PB2
#define NEGF_PB2 32   // panel width of the streaming kernel
#endif
#define NEGF_CB2 4    // columns per streamed block
#define NEGF_MAXW 16  // waves per workgroup (1024 threads)

template <int PB>
struct GjeShared {
    int perm[NEGF_MAXNP];
    int ipiv[NEGF_MAXNP];
    int sigma[NEGF_MAXNP];
    int disp[2 * PB];
    int ndisp;
    int flag;
    double rv[2][NEGF_MAXW];
    int ri[2][NEGF_MAXW];
    cd cand[2][NEGF_MAXW][PB];
    cd rowC[2][PB];
    cd Tw[NEGF_MAXW][PB * NEGF_CB2];
    double red[NEGF_MAXW];
};

template <int PB, int J>
__device__ __forceinline__ void gje_step2(cd (&P)[PB], const int k0, const int Np, const bool own,
                                          GjeShared<PB> &s) {
    const int r = threadIdx.x, lane = r & 63, w = r >> 6, nw = blockDim.x >> 6;
    const int c = k0 + J, par = J & 1;
    double v = -1.0;
    int idx = 0x7fffffff;
    if (own && r >= c) {
        v = cabs1(P[J]);
        if (!(v >= 0.0)) v = -0.5;   // NaN ranks below numbers, above ineligible rows
        idx = r;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double ov = __shfl_xor(v, off, 64);
        const int oi = __shfl_xor(idx, off, 64);
        if (ov > v || (ov == v && oi < idx)) { v = ov; idx = oi; }
    }
    if (lane == 0) { s.rv[par][w] = v; s.ri[par][w] = idx; }
    if (r == idx) {
#pragma unroll
        for (int q = 0; q < PB; ++q) s.cand[par][w][q] = P[q];
    }
    if (r == c) {
#pragma unroll
        for (int q = 0; q < PB; ++q) s.rowC[par][q] = P[q];
    }
    __syncthreads();
    double bv = s.rv[par][0];
    int p = s.ri[par][0], wb = 0;
    for (int i = 1; i < nw; ++i) {
        const double ov = s.rv[par][i];
        const int oi = s.ri[par][i];
        if (ov > bv || (ov == bv && oi < p)) { bv = ov; p = oi; wb = i; }
    }
    if (p < c || p >= Np) p = c;
    const cd *prow = (p == c) ? s.rowC[par] : s.cand[par][wb];
    const cd inv = crecip(prow[J]);
    if (r == 0) {
        s.ipiv[c] = p;
        if (p != c) {
            const int t = s.perm[c];
            s.perm[c] = s.perm[p];
            s.perm[p] = t;
        }
        if (prow[J].x == 0.0 && prow[J].y == 0.0 && s.flag == 0) s.flag = c + 1;
    }
    if (r == c) {
#pragma unroll
        for (int q = 0; q < PB; ++q) P[q] = (q == J) ? inv : cmul(prow[q], inv);
    } else if (own) {
        if (r == p) {
#pragma unroll
            for (int q = 0; q < PB; ++q) P[q] = s.rowC[par][q];
        }
        const cd f = P[J];
#pragma unroll
        for (int q = 0; q < PB; ++q) {
            if (q == J) P[q] = cmul(cmk(-f.x, -f.y), inv);
            else P[q] = cfms(f, cmul(prow[q], inv), P[q]);
        }
    }
}

template <int PB, int... Js>
__device__ __forceinline__ void gje_steps2(cd (&P)[PB], const int k0, const int Np, const bool own,
                                           GjeShared<PB> &s, std::integer_sequence<int, Js...>) {
    (gje_step2<PB, Js>(P, k0, Np, own, s), ...);
}

__device__ __forceinline__ double block_sum2(double v, double *red) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    __syncthreads();
    if (lane == 0) red[w] = v;
    __syncthreads();
    double t = 0.0;
    for (int i = 0; i < nw; ++i) t += red[i];
    __syncthreads();
    return t;
}

// In-place inversion of Z (Np x Np, Np a multiple of PB); on return
// inverse[:, c] = Z[:, s.sigma[c]].  scratch: (2 PB) x Np complex per workgroup.
// Returns the 1-based column of the first exactly-zero pivot, or 0.
template <int PB>
__device__ __forceinline__ int gje_invert2(cd *__restrict__ Z, const int Np, GjeShared<PB> &s,
                                           cd *__restrict__ scratch) {
    const int r = threadIdx.x, lane = r & 63, w = r >> 6;
    const bool own = r < Np;
    const size_t ld = (size_t)Np;
    constexpr int CB = NEGF_CB2;
    constexpr int TPL = PB * CB / 64;      // T entries staged per lane per block
    cd *Tg = scratch;                      // old pivot rows: Tg[j + c PB]
    cd *Dg = scratch + (size_t)PB * Np;    // displaced rows: Dg[d + c PB]
    if (r == 0) s.flag = 0;
    for (int k0 = 0; k0 < Np; k0 += PB) {
        cd P[PB];
#pragma unroll
        for (int j = 0; j < PB; ++j) P[j] = own ? Z[r + (k0 + j) * ld] : cmk(0.0, 0.0);
        for (int i = r; i < Np; i += blockDim.x) s.perm[i] = i;
        if (r == 0) s.ndisp = 0;
        __syncthreads();
        gje_steps2<PB>(P, k0, Np, own, s, std::make_integer_sequence<int, PB>{});
        __syncthreads();
        if (own) {
#pragma unroll
            for (int j = 0; j < PB; ++j) Z[r + (k0 + j) * ld] = P[j];
            if (r >= k0 + PB && s.perm[r] != r) s.disp[atomicAdd(&s.ndisp, 1)] = r;
        }
        __syncthreads();
        // gather old pivot rows and displaced rows (every column but the panel's)
        const int nd = s.ndisp;
        const int rows = PB + nd;
        for (int e = r; e < rows * Np; e += blockDim.x) {
            const int q = e % rows, col = e / rows;
            if (col >= k0 && col < k0 + PB) continue;
            if (q < PB) Tg[q + (size_t)col * PB] = Z[s.perm[k0 + q] + col * ld];
            else Dg[(q - PB) + (size_t)col * PB] = Z[s.perm[s.disp[q - PB]] + col * ld];
        }
        __syncthreads();
        for (int e = r; e < nd * Np; e += blockDim.x) {
            const int q = e % nd, col = e / nd;
            if (col >= k0 && col < k0 + PB) continue;
            Z[s.disp[q] + col * ld] = Dg[q + (size_t)col * PB];
        }
        __syncthreads();
        // streamed rank-PB update of every other column, wave by wave
        const bool pivrow = (r >= k0 && r < k0 + PB);
        cd *tw = s.Tw[w];
        int cb = (k0 == 0) ? PB : 0;
        cd a[CB], tn[TPL];
        if (cb < Np) {
#pragma unroll
            for (int cc = 0; cc < CB; ++cc) a[cc] = (own && !pivrow) ? Z[r + (cb + cc) * ld] : cmk(0.0, 0.0);
#pragma unroll
            for (int t = 0; t < TPL; ++t) tn[t] = Tg[lane + 64 * t + (size_t)cb * PB];
        }
        while (cb < Np) {
            int cbn = cb + CB;
            if (cbn >= k0 && cbn < k0 + PB) cbn = k0 + PB;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int t = 0; t < TPL; ++t) tw[lane + 64 * t] = tn[t];
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
            __builtin_amdgcn_wave_barrier();
            cd an[CB];
            if (cbn < Np) {
#pragma unroll
                for (int cc = 0; cc < CB; ++cc)
                    an[cc] = (own && !pivrow) ? Z[r + (cbn + cc) * ld] : cmk(0.0, 0.0);
#pragma unroll
                for (int t = 0; t < TPL; ++t) tn[t] = Tg[lane + 64 * t + (size_t)cbn * PB];
            }
            if (own) {
#pragma unroll
                for (int cc = 0; cc < CB; ++cc) {
                    cd acc = a[cc];
#pragma unroll
                    for (int j = 0; j < PB; ++j) acc = cfma(P[j], tw[j + cc * PB], acc);
                    Z[r + (cb + cc) * ld] = acc;
                }
            }
#pragma unroll
            for (int cc = 0; cc < CB; ++cc) a[cc] = an[cc];
            cb = cbn;
        }
        __syncthreads();
    }
    if (r == 0) {
        for (int c = 0; c < Np; ++c) s.sigma[c] = c;
        for (int c = Np - 1; c >= 0; --c) {
            const int p = s.ipiv[c];
            const int t = s.sigma[c];
            s.sigma[c] = s.sigma[p];
            s.sigma[p] = t;
        }
    }
    __syncthreads();
    return s.flag;
}
