// rgf_device.h -- device building blocks of the RGF sweep (gfx950, wave64).
//
// Matrices are column-major with leading dimension Np: N padded to a multiple of 32 up to
// 512 (16 above); the padding block is the identity and stays inert through every
// operation.  One workgroup owns one matrix at a time: 2 Np threads (two lanes per row)
// for the matrix-core Gauss-Jordan inversion, Np rounded up to 64 (one thread per row)
// above Np = 512.  Complex128 throughout, except the FP32 sweep of the mixed path.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <utility>


typedef double2 cd;

__device__ __forceinline__ cd cmk(double re, double im) { return make_double2(re, im); }
// threadIdx.x through an empty asm: values derived from it are recomputed where they are
// used instead of being hoisted out of an enclosing loop (a job or slice loop) and kept
// live across it.
__device__ __forceinline__ int opaque_tid() {
    int x = threadIdx.x;
    asm volatile("" : "+v"(x));
    return x;
}

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
// 1/b via an exact power-of-two scaling, v_rcp_f64 and two Newton steps (about 1 ulp,
// no IEEE division sequences on the pivot's critical path).
__device__ __forceinline__ cd crecip_fast(cd b) {
    const double m = fmax(fabs(b.x), fabs(b.y));
    if (!(m > 0.0) || !(m < INFINITY)) return crecip(b);   // 0, inf, nan: reference path
    const int e = __builtin_amdgcn_frexp_exp(m);
    const double bx = __builtin_amdgcn_ldexp(b.x, -e), by = __builtin_amdgcn_ldexp(b.y, -e);
    const double d = fma(bx, bx, by * by);                  // in [1/4, 2)
    double r = __builtin_amdgcn_rcp(d);
    r = fma(r, fma(-d, r, 1.0), r);
    r = fma(r, fma(-d, r, 1.0), r);
    return cmk(__builtin_amdgcn_ldexp(bx * r, -e), __builtin_amdgcn_ldexp(-by * r, -e));
}

// |re| + |im|, the pivot measure of LAPACK izamax (dcabs1).
__device__ __forceinline__ double cabs1(cd a) { return fabs(a.x) + fabs(a.y); }

// complex64 (the single-precision sweep of the mixed path) and conversions
typedef float2 cfl;
__device__ __forceinline__ cfl cmk(float re, float im) { return make_float2(re, im); }
__device__ __forceinline__ cd to_cd(cd v) { return v; }
__device__ __forceinline__ cd to_cd(cfl v) { return cmk((double)v.x, (double)v.y); }
template <class C>
__device__ __forceinline__ C from_cd(cd v);
template <>
__device__ __forceinline__ cd from_cd<cd>(cd v) { return v; }
template <>
__device__ __forceinline__ cfl from_cd<cfl>(cd v) { return cmk((float)v.x, (float)v.y); }

#define NEGF_MAXNP 1024   // largest padded block (one workgroup per matrix)

typedef double negf_d4 __attribute__((ext_vector_type(4)));

// C = A * B(:, sigB) (sigB == nullptr: identity), Np x Np column-major (Np a multiple of
// 16), on the FP64 matrix cores.  Every wave computes 32 x 32 output blocks as 2 x 2
// v_mfma_f64_16x16x4 tiles with three real products per complex one (the 3M form:
// P1 = Ar Br, P2 = Ai Bi, P3 = (Ar + Ai)(Br + Bi); re = P1 - P2, im = P3 - P1 - P2).  Operand layout of the 16x16x4 FP64 MFMA: lane i + 16 k supplies A[i][k] and
// B[k][i]; output element (i, j) sits in lane j + 16 (i % 4), register i / 4.  Operands
// are read straight from memory in that layout (A: 16 consecutive rows of 4 columns, B:
// 4 consecutive rows of 16 columns).  Rows and columns past Np are clamped on load and
// skipped in the epilogue, which receives (row, col, value) for every in-range element.
template <class Epi>
__device__ __forceinline__ void gemm_nn(const cd *__restrict__ A, const cd *__restrict__ B, const int *sigB,
                                        const int Np, Epi epi) {
    const size_t ld = (size_t)Np;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int td = (Np + 31) / 32;
    for (int t = w; t < td * td; t += nw) {
        const int r0 = (t % td) * 32, c0 = (t / td) * 32;
        const int ra = min(r0 + li, Np - 1), rb = min(r0 + 16 + li, Np - 1);
        const int ca = min(c0 + li, Np - 1), cb = min(c0 + 16 + li, Np - 1);
        const cd *__restrict__ pa = A + ra, *__restrict__ pb = A + rb;
        const cd *__restrict__ qa = B + (size_t)(sigB ? sigB[ca] : ca) * ld;
        const cd *__restrict__ qb = B + (size_t)(sigB ? sigB[cb] : cb) * ld;
        negf_d4 p1[2][2], p2[2][2], p3[2][2];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                p1[a][b] = negf_d4{0.0, 0.0, 0.0, 0.0};
                p2[a][b] = negf_d4{0.0, 0.0, 0.0, 0.0};
                p3[a][b] = negf_d4{0.0, 0.0, 0.0, 0.0};
            }
        for (int k0 = 0; k0 < Np; k0 += 4) {
            const int k = k0 + lk;
            const cd av[2] = {pa[(size_t)k * ld], pb[(size_t)k * ld]};
            const cd bv[2] = {qa[k], qb[k]};
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b) {
                    p1[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a].x, bv[b].x, p1[a][b], 0, 0, 0);
                    p2[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a].y, bv[b].y, p2[a][b], 0, 0, 0);
                    p3[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a].x + av[a].y, bv[b].x + bv[b].y, p3[a][b],
                                                                     0, 0, 0);
                }
        }
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int r = r0 + 16 * a + 4 * q + lk, c = c0 + 16 * b + li;
                    const double re = p1[a][b][q] - p2[a][b][q], im = p3[a][b][q] - p1[a][b][q] - p2[a][b][q];
                    if (r < Np && c < Np) epi(r, c, cmk(re, im));
                }
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

// Element (r, col), r, col < N, of D_slice = 2 Delta^2 ((E + i eta) I - H_k) - 2 I (+ lead
// terms on the end slices), in the operation order of E_minus_H_construction
// (src/grid.f90:291-331) with H_k(i,j) = V(nx_i - nx_j, ny_i - ny_j, k) + delta_ij kin_i
// (src/grid.f90:248-259).  h = V(...), dg = the lead terms of row r.
__device__ __forceinline__ cd d_entry(cd h, bool diag, double kin, cd z, double scale, cd dg) {
    if (diag) h.x += kin;
    cd d = cmk(-h.x, -h.y);
    if (diag) d = cadd(d, z);
    d.x *= scale;
    d.y *= scale;
    if (diag) {
        d.x -= 2.0;
        d = cadd(d, dg);
    }
    return d;
}

// M = D_slice - S(:, sig) (S == nullptr: M = D_slice); the padding block is the identity.
// C = cd stores complex128, C = cfl rounds the FP64 result to complex64.
template <class T>
struct negf_id {
    typedef T type;
};
template <class C = cd>
__device__ __forceinline__ void build_block(C *__restrict__ M, const SysView &sv, const JobView &jv, const int slice,
                                            const typename negf_id<C>::type *__restrict__ S, const int *sig) {
    const int Np = sv.Np, N = sv.N;
    const int W = 2 * sv.NGX + 1;
    const cd *__restrict__ vs = sv.vrec + (size_t)slice * W * (2 * sv.NGY + 1);
    constexpr int CBB = 8;
    // every (row, block of CBB columns) pair is one work item, so all threads of the
    // workgroup take part (two per row at blockDim = 2 Np); rows stay contiguous per
    // column for coalesced stores
    const int nitems = Np * (Np / CBB);
    for (int it = threadIdx.x; it < nitems; it += blockDim.x) {
        const int r = it % Np, c0 = (it / Np) * CBB;
        const bool rin = r < N;
        const int nxr = rin ? sv.nx[r] : 0, nyr = rin ? sv.ny[r] : 0;
        const double kin = rin ? jv.kinh[r] : 0.0;
        cd dg = cmk(0.0, 0.0);           // diagonal additions after scaling
        if (rin) {
            if (slice == 0) dg = cadd(dg, jv.sigL[r]);
            if (slice == sv.Nz - 1) dg = cadd(dg, jv.sigR[r]);
        }
        {
            cd hv[CBB], sv_[CBB];
#pragma unroll
            for (int q = 0; q < CBB; ++q) {
                const int col = c0 + q;
                const bool in = rin && col < N;
                int ix = 0;
                if (in) ix = (nxr - sv.nx[col] + sv.NGX) + W * (nyr - sv.ny[col] + sv.NGY);
                hv[q] = in ? vs[ix] : cmk(0.0, 0.0);
                sv_[q] = S ? to_cd(S[r + (size_t)(sig ? sig[col] : col) * Np]) : cmk(0.0, 0.0);
            }
#pragma unroll
            for (int q = 0; q < CBB; ++q) {
                const int col = c0 + q;
                cd v;
                if (rin && col < N) {
                    cd d = d_entry(hv[q], r == col, kin, jv.z, sv.scale, dg);
                    if (S) d = csub(d, sv_[q]);
                    v = d;
                } else {
                    v = cmk(r == col ? 1.0 : 0.0, 0.0);
                }
                M[r + (size_t)col * Np] = from_cd<C>(v);
            }
        }
    }
}

// build_block for the mixed sweep (M = D_slice - S, FP64 arithmetic rounded to complex64),
// with the slice's V-hat window read through vs (LDS or global) and the basis keys
// ckey[c] = nx_c + (2 NGX + 1) ny_c (c < N) from LDS: the window index of (r, c) is
// ckey[r] - ckey[c] + NGX + (2 NGX + 1) NGY.
__device__ __forceinline__ void build_block_keys(cfl *__restrict__ M, const SysView &sv, const JobView &jv,
                                                 const int slice, const cfl *__restrict__ S, const cd *vs,
                                                 const int *ckey) {
    const int Np = sv.Np, N = sv.N;
    const int base0 = sv.NGX + (2 * sv.NGX + 1) * sv.NGY;
    constexpr int CBB = 8;
    const int nitems = Np * (Np / CBB);
#pragma unroll 2
    for (int it = opaque_tid(); it < nitems; it += blockDim.x) {
        const int r = it % Np, c0 = (it / Np) * CBB;
        const bool rin = r < N;
        const int rb = rin ? ckey[r] + base0 : 0;
        const double kin = rin ? jv.kinh[r] : 0.0;
        cd dg = cmk(0.0, 0.0);
        if (rin) {
            if (slice == 0) dg = cadd(dg, jv.sigL[r]);
            if (slice == sv.Nz - 1) dg = cadd(dg, jv.sigR[r]);
        }
        cfl sp[CBB];
        cd hv[CBB];
#pragma unroll
        for (int q = 0; q < CBB; ++q) {
            const int col = c0 + q;
            sp[q] = S ? S[r + (size_t)col * Np] : cmk(0.0f, 0.0f);
            hv[q] = (rin && col < N) ? vs[rb - ckey[col]] : cmk(0.0, 0.0);
        }
#pragma unroll
        for (int q = 0; q < CBB; ++q) {
            const int col = c0 + q;
            cd v;
            if (rin && col < N) v = csub(d_entry(hv[q], r == col, kin, jv.z, sv.scale, dg), to_cd(sp[q]));
            else v = cmk(r == col ? 1.0 : 0.0, 0.0);
            M[r + (size_t)col * Np] = from_cd<cfl>(v);
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
// Pivot rule of zgetrf (largest |re|+|im| among the rows not yet pivoted, first index on
// ties) and the arithmetic count of zgetrf+zgetri; rows are interchanged physically,
// panel by panel, and the column interchanges are left to the consumer (on return
// inverse[:, c] = Z[:, s.sigma[c]]).  One thread per row, for Np > 512.  The trailing
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
#define NEGF_TCH 128  // columns of the old pivot rows staged in LDS at a time

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

// Wave-wide max of a double with DPP (row-local shuffles and row broadcasts), result
// valid in every lane.
template <int CTRL, int ROWMASK>
__device__ __forceinline__ double dpp_max_step(double v) {
    const int lo = __double2loint(v), hi = __double2hiint(v);
    const int olo = __builtin_amdgcn_update_dpp(lo, lo, CTRL, ROWMASK, 0xf, false);
    const int ohi = __builtin_amdgcn_update_dpp(hi, hi, CTRL, ROWMASK, 0xf, false);
    const double o = __hiloint2double(ohi, olo);
    return o > v ? o : v;
}
__device__ __forceinline__ double wave_max(double v) {
    v = dpp_max_step<0xB1, 0xf>(v);    // quad_perm [1,0,3,2]
    v = dpp_max_step<0x4E, 0xf>(v);    // quad_perm [2,3,0,1]
    v = dpp_max_step<0x141, 0xf>(v);   // row_half_mirror
    v = dpp_max_step<0x140, 0xf>(v);   // row_mirror
    v = dpp_max_step<0x142, 0xa>(v);   // row_bcast:15
    v = dpp_max_step<0x143, 0xc>(v);   // row_bcast:31
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), 63);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), 63);
    return __hiloint2double(hi, lo);
}

// Value of lane L (compile-time) of a complex register, as a wave-uniform scalar.
__device__ __forceinline__ cd wave_bcast(cd v, int L) {
    const int xl = __builtin_amdgcn_readlane(__double2loint(v.x), L);
    const int xh = __builtin_amdgcn_readlane(__double2hiint(v.x), L);
    const int yl = __builtin_amdgcn_readlane(__double2loint(v.y), L);
    const int yh = __builtin_amdgcn_readlane(__double2hiint(v.y), L);
    return cmk(__hiloint2double(xh, xl), __hiloint2double(yh, yl));
}

// One Gauss-Jordan step (column c = k0 + J) of the register-resident panel.  J is a
// runtime (wave-uniform) index: the panel registers are addressed through unrolled
// selects, so the PB steps share one copy of the code instead of PB unrolled copies
// (the unrolled form overflows the instruction cache).
template <int PB>
__device__ __forceinline__ void gje_step2(cd (&P)[PB], const int J, const int k0, const int Np,
                                          const bool own, GjeShared<PB> &s) {
    const int r = threadIdx.x, lane = r & 63, w = r >> 6, nw = blockDim.x >> 6;
    const int c = k0 + J, par = J & 1;
    cd pj = cmk(0.0, 0.0);
#pragma unroll
    for (int q = 0; q < PB; ++q)
        if (q == J) pj = P[q];
    double v = -1.0;
    if (own && r >= c) {
        v = cabs1(pj);
        if (!(v >= 0.0)) v = -0.5;   // NaN ranks below numbers, above ineligible rows
    }
    // wave argmax: max value, then the lowest lane holding it (rows are lanes in order,
    // so the lowest lane is the lowest row index, as izamax picks the first maximum)
    const double vmax = wave_max(v);
    const unsigned long long hit = __ballot(v == vmax);
    const int idx = (vmax >= 0.0 || vmax == -0.5) ? (w << 6) + (__ffsll((long long)hit) - 1) : 0x7fffffff;
    if (lane == 0) { s.rv[par][w] = vmax; s.ri[par][w] = idx; }
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
    const cd pv = prow[J];
    const cd inv = crecip(pv);
    if (r == 0) {
        s.ipiv[c] = p;
        if (p != c) {
            const int t = s.perm[c];
            s.perm[c] = s.perm[p];
            s.perm[p] = t;
        }
        if (pv.x == 0.0 && pv.y == 0.0 && s.flag == 0) s.flag = c + 1;
    }
    // Row r <- row r - (P[r][J] / pivot) * pivot row; the pivot row itself becomes
    // pivot row / pivot (same form with base 0 and multiplier -1/pivot); column J holds
    // -multiplier.  One complex FMA per entry.
    if (own) {
        const bool piv = (r == c);
        const bool takeC = (r == p && r != c);
        cd mr = pj;
        if (takeC) {              // this row receives the old content of row c
            mr = s.rowC[par][J];
#pragma unroll
            for (int q = 0; q < PB; ++q) P[q] = s.rowC[par][q];
        }
        if (piv) {
#pragma unroll
            for (int q = 0; q < PB; ++q) P[q] = cmk(0.0, 0.0);
        }
        const cd f = piv ? cmk(-inv.x, -inv.y) : cmul(mr, inv);
#pragma unroll
        for (int q = 0; q < PB; ++q) {
            const cd nv = cfms(f, prow[q], P[q]);
            P[q] = (q == J) ? cmk(-f.x, -f.y) : nv;
        }
    }
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

__device__ __forceinline__ double block_max2(double v, double *red) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    __syncthreads();
    if (lane == 0) red[w] = v;
    __syncthreads();
    double t = 0.0;
    for (int i = 0; i < nw; ++i) t = fmax(t, red[i]);
    __syncthreads();
    return t;
}

// In-place inversion of Z (Np x Np, Np a multiple of PB); on return
// inverse[:, c] = Z[:, s.sigma[c]].  scratch: (2 PB) x Np complex per workgroup.
// Returns the 1-based column of the first exactly-zero pivot, or 0.
template <int PB>
__device__ __forceinline__ int gje_invert2(cd *__restrict__ Z, const int Np, GjeShared<PB> &s,
                                           cd *__restrict__ scratch) {
    static_assert(NEGF_CB2 * PB % 64 == 0, "T block must split evenly over a wave");
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
#pragma unroll 1
        for (int J = 0; J < PB; ++J) gje_step2<PB>(P, J, k0, Np, own, s);
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
        {
            constexpr int GB = 8;
            const int total = rows * Np;
            for (int e0 = r; e0 < total; e0 += GB * blockDim.x) {
                cd gv[GB];
#pragma unroll
                for (int g = 0; g < GB; ++g) {
                    const int e = e0 + g * blockDim.x;
                    const int q = e % rows, col = e / rows;
                    const bool ok = e < total && !(col >= k0 && col < k0 + PB);
                    const int src = ok ? (q < PB ? s.perm[k0 + q] : s.perm[s.disp[q - PB]]) : 0;
                    gv[g] = ok ? Z[src + col * ld] : cmk(0.0, 0.0);
                }
#pragma unroll
                for (int g = 0; g < GB; ++g) {
                    const int e = e0 + g * blockDim.x;
                    const int q = e % rows, col = e / rows;
                    if (e < total && !(col >= k0 && col < k0 + PB)) {
                        if (q < PB) Tg[q + (size_t)col * PB] = gv[g];
                        else Dg[(q - PB) + (size_t)col * PB] = gv[g];
                    }
                }
            }
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
            cd tc[TPL];
#pragma unroll
            for (int t = 0; t < TPL; ++t) tc[t] = tn[t];
            cd an[CB];
            if (cbn < Np) {
#pragma unroll
                for (int cc = 0; cc < CB; ++cc)
                    an[cc] = (own && !pivrow) ? Z[r + (cbn + cc) * ld] : cmk(0.0, 0.0);
#pragma unroll
                for (int t = 0; t < TPL; ++t) tn[t] = Tg[lane + 64 * t + (size_t)cbn * PB];
            }
            // T(j, cc) sits in lane (j + cc PB) % 64, register (j + cc PB) / 64 of tc;
            // readlane turns it into a scalar operand of the FMAs.
#pragma unroll
            for (int cc = 0; cc < CB; ++cc) {
                cd p4[4] = {a[cc], cmk(0.0, 0.0), cmk(0.0, 0.0), cmk(0.0, 0.0)};
#pragma unroll
                for (int j = 0; j < PB; ++j) {
                    const int e = j + cc * PB;
                    p4[j & 3] = cfma(P[j], wave_bcast(tc[e >> 6], e & 63), p4[j & 3]);
                }
                if (own) Z[r + (cb + cc) * ld] = cadd(cadd(p4[0], p4[1]), cadd(p4[2], p4[3]));
            }
#pragma unroll
            for (int cc = 0; cc < CB; ++cc) a[cc] = an[cc];
            cb = cbn;
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

// Value held by the partner lane (t ^ 1).
__device__ __forceinline__ double lane_swap(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), 0xB1, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), 0xB1, 0xf, 0xf, false);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ cd lane_swap(cd v) { return cmk(lane_swap(v.x), lane_swap(v.y)); }

// ====================================================================================
// Gauss-Jordan inversion with the trailing update on the matrix cores, in FP64
// (v_mfma_f64_16x16x4f64) or FP32 (v_mfma_f32_16x16x4f32).
//
// Np % 32 == 0, Np <= NEGF_MAXNP4, blockDim.x == 2 Np.  Wave w owns rows [32 w, 32 w + 32)
// as two 16-row tiles.  Lane l = j + 16 k (j = l & 15, k = l >> 4) holds, for tile t,
// the panel values P[t][q] = Z[32 w + 16 t + j][k0 + 4 q + k] -- exactly the B operand
// of the 16x16x4 MFMA (B[k][j] at lane j + 16 k, both precisions) for the products
//     Z[rows][ct .. ct+15] += P[rows][4q .. 4q+3] * T[4q .. 4q+3][ct .. ct+15]
// computed as D = A B with A[i][k] = T[4q + k][ct + i] (lane i + 16 k).  D[i][j] sits in
// lane j + 16 (i % 4), register i / 4 for FP64 and in lane j + 16 (i / 4), register i % 4
// for FP32, so the lane group q = l >> 4 and accumulator register r hold column
// ct + CQ q + CR r (GjePrec).  T (the old pivot rows) is staged in LDS.  During the pivot
// steps the panel registers rotate once per four steps, so the current column always
// sits in P[t][0] (no runtime register indexing).
// ====================================================================================
#define NEGF_MAXNP4 512

typedef float negf_f4 __attribute__((ext_vector_type(4)));
typedef int negf_i4 __attribute__((ext_vector_type(4)));
typedef int negf_i2 __attribute__((ext_vector_type(2)));

// complex64 arithmetic (same conventions as the complex128 helpers)
__device__ __forceinline__ cfl cmul(cfl a, cfl b) { return cmk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ __forceinline__ cfl cfms(cfl a, cfl b, cfl c) {
    c.x = fmaf(-a.x, b.x, c.x);
    c.x = fmaf(a.y, b.y, c.x);
    c.y = fmaf(-a.x, b.y, c.y);
    c.y = fmaf(-a.y, b.x, c.y);
    return c;
}
__device__ __forceinline__ float cabs1(cfl a) { return fabsf(a.x) + fabsf(a.y); }
// 1/b: exact power-of-two scaling, v_rcp_f32 and one Newton step; Smith's method for
// zero, infinite or NaN input.
__device__ __forceinline__ cfl crecip_fast(cfl b) {
    const float m = fmaxf(fabsf(b.x), fabsf(b.y));
    if (!(m > 0.0f) || !(m < INFINITY)) {
        if (fabsf(b.x) >= fabsf(b.y)) {
            const float r = b.y / b.x, d = b.x + b.y * r;
            return cmk(1.0f / d, -r / d);
        }
        const float r = b.x / b.y, d = b.x * r + b.y;
        return cmk(r / d, -1.0f / d);
    }
    const int e = __builtin_amdgcn_frexp_expf(m);
    const float bx = __builtin_amdgcn_ldexpf(b.x, -e), by = __builtin_amdgcn_ldexpf(b.y, -e);
    const float d = fmaf(bx, bx, by * by);
    float r = __builtin_amdgcn_rcpf(d);
    r = fmaf(r, fmaf(-d, r, 1.0f), r);
    return cmk(__builtin_amdgcn_ldexpf(bx * r, -e), __builtin_amdgcn_ldexpf(-by * r, -e));
}

template <int CTRL, int ROWMASK>
__device__ __forceinline__ float dpp_max_step(float v) {
    const int i = __float_as_int(v);
    const float o = __int_as_float(__builtin_amdgcn_update_dpp(i, i, CTRL, ROWMASK, 0xf, false));
    return o > v ? o : v;
}

// Max over the 16 lanes of DPP row `row`, returned wave-uniform.
__device__ __forceinline__ double row16_max(double v, int row) {
    v = dpp_max_step<0xB1, 0xf>(v);
    v = dpp_max_step<0x4E, 0xf>(v);
    v = dpp_max_step<0x141, 0xf>(v);
    v = dpp_max_step<0x140, 0xf>(v);
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), row << 4);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), row << 4);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ float row16_max(float v, int row) {
    v = dpp_max_step<0xB1, 0xf>(v);
    v = dpp_max_step<0x4E, 0xf>(v);
    v = dpp_max_step<0x141, 0xf>(v);
    v = dpp_max_step<0x140, 0xf>(v);
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), row << 4));
}

// Raw buffer access: one 32-bit VGPR offset per lane plus a wave-uniform SGPR offset,
// instead of a 64-bit address per access.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t negf_rsrc(void *p, unsigned bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(p, (short)0, (int)bytes, 0x00020000);
}

template <class R>
struct GjePrec;
template <>
struct GjePrec<double> {
    typedef cd C;
    typedef negf_d4 Acc;
    static constexpr unsigned ESZ = 16;       // bytes per complex element
    static constexpr int CQ = 1, CR = 4;      // accumulator column = CQ q + CR r
    __device__ static __forceinline__ Acc mfma(double a, double b, Acc c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    __device__ static __forceinline__ C load(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
        const negf_i4 v = __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0);
        return cmk(__hiloint2double(v.y, v.x), __hiloint2double(v.w, v.z));
    }
    __device__ static __forceinline__ void store(C x, __amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
        negf_i4 v;
        v.x = __double2loint(x.x);
        v.y = __double2hiint(x.x);
        v.z = __double2loint(x.y);
        v.w = __double2hiint(x.y);
        __builtin_amdgcn_raw_buffer_store_b128(v, r, voff, soff, 0);
    }
};
template <>
struct GjePrec<float> {
    typedef cfl C;
    typedef negf_f4 Acc;
    static constexpr unsigned ESZ = 8;
    static constexpr int CQ = 4, CR = 1;
    __device__ static __forceinline__ Acc mfma(float a, float b, Acc c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    __device__ static __forceinline__ C load(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
        const negf_i2 v = __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0);
        return cmk(__int_as_float(v.x), __int_as_float(v.y));
    }
    __device__ static __forceinline__ void store(C x, __amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
        negf_i2 v;
        v.x = __float_as_int(x.x);
        v.y = __float_as_int(x.y);
        __builtin_amdgcn_raw_buffer_store_b64(v, r, voff, soff, 0);
    }
};

// MAXNP: largest padded size the workgroup serves (blockDim = 2 Np <= 2 MAXNP); TCH:
// columns of the old pivot rows staged in LDS at a time.  Smaller values shrink the LDS
// footprint so that several workgroups share a CU at small Np; the arithmetic does not
// depend on either.
template <int PB, class R = double, int MAXNP = NEGF_MAXNP4, int TCH = NEGF_TCH>
struct Gje4Shared {
    typedef typename GjePrec<R>::C C;
    static constexpr int MAXW = 2 * MAXNP / 64;   // waves per workgroup at most
    static constexpr int CH = TCH, LDT = TCH + 1;  // LDS row stride of the staged rows (padded)
    int perm[MAXNP];
    int ipiv[MAXNP];
    int sigma[MAXNP];
    int disp[2 * PB];
    int ndisp;
    int flag;
    R rv[2][MAXW];
    int ri[2][MAXW];
    C cand[2][MAXW][PB];        // each wave's best candidate row (panel columns)
    C rowC[2][PB];              // row c (the pivot position) before the swap
    C colJ[2][MAXNP];           // panel column J of every row
    C Tl[PB * LDT];             // old pivot rows of one column chunk: Tl[p * LDT + (c - c0)]
    double red[MAXW];
};

// One pivot step, column J = 4 g + KJ of the panel; P[.][q] holds panel column
// 4 ((q + g) % (PB / 4)) + k.
template <int PB, int KJ, class R, class SH>
__device__ __forceinline__ void gje4_step(typename GjePrec<R>::C (&P)[2][PB / 4], const int g, const int k0,
                                          const int Np, const int w, const int j, const int k, SH &s) {
    typedef typename GjePrec<R>::C C;
    constexpr int KK = PB / 4;
    const int t = threadIdx.x, lane = t & 63, nw = blockDim.x >> 6;
    const int J = 4 * g + KJ, c = k0 + J, par = J & 1;
    const int r0 = (w << 5) + j, r1 = r0 + 16;
    const bool mine = (k == KJ);
    R v0 = R(-1), v1 = R(-1);
    if (mine) {
        if (r0 >= c) {
            v0 = cabs1(P[0][0]);
            if (!(v0 >= R(0))) v0 = R(-0.5);
        }
        if (r1 >= c) {
            v1 = cabs1(P[1][0]);
            if (!(v1 >= R(0))) v1 = R(-0.5);
        }
        s.colJ[par][r0] = P[0][0];
        s.colJ[par][r1] = P[1][0];
    }
    // izamax order: largest |re| + |im|, lowest row on ties (tile 0 rows precede tile 1)
    const R vmax = row16_max(v0 > v1 ? v0 : v1, KJ);
    int idx = 0x7fffffff;
    if (vmax > R(-1)) {
        const unsigned long long h0 = __ballot(v0 == vmax), h1 = __ballot(v1 == vmax);
        idx = h0 ? (w << 5) + ((__ffsll((long long)h0) - 1) & 15)
                 : (w << 5) + 16 + ((__ffsll((long long)h1) - 1) & 15);
    }
    if (lane == 0) {
        s.rv[par][w] = vmax;
        s.ri[par][w] = idx;
    }
    // (separate stores per tile: a select between P[0] and P[1] would become a select
    // of addresses and push P out of registers)
    const bool cnd = idx != 0x7fffffff && j == (idx & 15);
    const bool rc = (c >> 5) == w && j == (c & 15);
    if (cnd && !((idx >> 4) & 1)) {
#pragma unroll
        for (int q = 0; q < KK; ++q) s.cand[par][w][4 * ((q + g) % KK) + k] = P[0][q];
    }
    if (cnd && ((idx >> 4) & 1)) {
#pragma unroll
        for (int q = 0; q < KK; ++q) s.cand[par][w][4 * ((q + g) % KK) + k] = P[1][q];
    }
    if (rc && !((c >> 4) & 1)) {
#pragma unroll
        for (int q = 0; q < KK; ++q) s.rowC[par][4 * ((q + g) % KK) + k] = P[0][q];
    }
    if (rc && ((c >> 4) & 1)) {
#pragma unroll
        for (int q = 0; q < KK; ++q) s.rowC[par][4 * ((q + g) % KK) + k] = P[1][q];
    }
    __syncthreads();
    // winner over the waves (at most 16, all in DPP row 0): first wave holding the max
    const R wv = lane < nw ? s.rv[par][lane] : R(-2);
    const int wi = lane < nw ? s.ri[par][lane] : 0x7fffffff;
    const R bmax = row16_max(wv, 0);
    const int wb = __ffsll((long long)__ballot(wv == bmax)) - 1;
    int p = __builtin_amdgcn_readlane(wi, wb);
    if (p < c || p >= Np) p = c;
    const C *prow = (p == c) ? s.rowC[par] : s.cand[par][wb];
    const C inv = crecip_fast(prow[J]);
    if (t == 0) {
        s.ipiv[c] = p;
        if (p != c) {
            const int tt = s.perm[c];
            s.perm[c] = s.perm[p];
            s.perm[p] = tt;
        }
        if (prow[J].x == R(0) && prow[J].y == R(0) && s.flag == 0) s.flag = c + 1;
    }
    // eliminate column J from both rows of this lane: the pivot row position takes the
    // scaled pivot row, the row that supplied the pivot takes old row c
    C m0 = s.colJ[par][r0], m1 = s.colJ[par][r1];
    const bool piv0 = (r0 == c), piv1 = (r1 == c);
    if (r0 == p && p != c) {
        m0 = s.rowC[par][J];
#pragma unroll
        for (int q = 0; q < KK; ++q) P[0][q] = s.rowC[par][4 * ((q + g) % KK) + k];
    }
    if (r1 == p && p != c) {
        m1 = s.rowC[par][J];
#pragma unroll
        for (int q = 0; q < KK; ++q) P[1][q] = s.rowC[par][4 * ((q + g) % KK) + k];
    }
    if (piv0) {
#pragma unroll
        for (int q = 0; q < KK; ++q) P[0][q] = cmk(R(0), R(0));
    }
    if (piv1) {
#pragma unroll
        for (int q = 0; q < KK; ++q) P[1][q] = cmk(R(0), R(0));
    }
    const C f0 = piv0 ? cmk(-inv.x, -inv.y) : cmul(m0, inv);
    const C f1 = piv1 ? cmk(-inv.x, -inv.y) : cmul(m1, inv);
#pragma unroll
    for (int q = 0; q < KK; ++q) {
        const C pq = prow[4 * ((q + g) % KK) + k];
        const C n0 = cfms(f0, pq, P[0][q]), n1 = cfms(f1, pq, P[1][q]);
        P[0][q] = (q == 0 && mine) ? cmk(-f0.x, -f0.y) : n0;
        P[1][q] = (q == 0 && mine) ? cmk(-f1.x, -f1.y) : n1;
    }
}

// In-place inversion of Z (Np x Np) by a workgroup of 2 Np threads; on return
// inverse[:, c] = Z[:, s.sigma[c]].  scratch: (2 PB) x Np elements per workgroup.
// Returns the 1-based column of the first exactly-zero pivot, or 0.
template <int PB, class R = double, class SH = Gje4Shared<PB, R>>
__device__ __forceinline__ int gje_invert4(typename GjePrec<R>::C *__restrict__ Z, const int Np, SH &s,
                                           typename GjePrec<R>::C *__restrict__ scratch) {
    static_assert(PB == 32, "two 16-row tiles per wave hold one 32-column panel");
    typedef GjePrec<R> G;
    typedef typename G::C C;
    typedef typename G::Acc Acc;
    constexpr int KK = PB / 4;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int j = lane & 15, k = lane >> 4;
    const int R0 = w << 5;
    const size_t ld = (size_t)Np;
    C *Tg = scratch;                      // old pivot rows: Tg[p + c PB]
    C *Dg = scratch + (size_t)PB * Np;    // displaced rows: Dg[d + c PB]
    if (t == 0) s.flag = 0;
    for (int k0 = 0; k0 < Np; k0 += PB) {
        C P[2][KK];
#pragma unroll
        for (int q = 0; q < KK; ++q) {
            const C *zc = Z + (R0 + j) + (size_t)(k0 + 4 * q + k) * ld;
            P[0][q] = zc[0];
            P[1][q] = zc[16];
        }
        for (int i = t; i < Np; i += blockDim.x) s.perm[i] = i;
        if (t == 0) s.ndisp = 0;
        __syncthreads();
#pragma unroll 1
        for (int g = 0; g < KK; ++g) {
            gje4_step<PB, 0, R>(P, g, k0, Np, w, j, k, s);
            gje4_step<PB, 1, R>(P, g, k0, Np, w, j, k, s);
            gje4_step<PB, 2, R>(P, g, k0, Np, w, j, k, s);
            gje4_step<PB, 3, R>(P, g, k0, Np, w, j, k, s);
            const C x0 = P[0][0], x1 = P[1][0];
#pragma unroll
            for (int q = 0; q < KK - 1; ++q) {
                P[0][q] = P[0][q + 1];
                P[1][q] = P[1][q + 1];
            }
            P[0][KK - 1] = x0;
            P[1][KK - 1] = x1;
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < KK; ++q) {
            C *zc = Z + (R0 + j) + (size_t)(k0 + 4 * q + k) * ld;
            zc[0] = P[0][q];
            zc[16] = P[1][q];
        }
        if (k == 0) {
#pragma unroll
            for (int tt = 0; tt < 2; ++tt) {
                const int r = R0 + 16 * tt + j;
                if (r >= k0 + PB && s.perm[r] != r) s.disp[atomicAdd(&s.ndisp, 1)] = r;
            }
        }
        __syncthreads();
        const int nd = s.ndisp;
        const int rows = PB + nd;
        {
            constexpr int GB = 4;
            const int total = rows * Np;
            for (int e0 = t; e0 < total; e0 += GB * blockDim.x) {
                C gv[GB];
#pragma unroll
                for (int gg = 0; gg < GB; ++gg) {
                    const int e = e0 + gg * blockDim.x;
                    const int q = e % rows, col = e / rows;
                    const bool ok = e < total && !(col >= k0 && col < k0 + PB);
                    const int src = ok ? (q < PB ? s.perm[k0 + q] : s.perm[s.disp[q - PB]]) : 0;
                    gv[gg] = ok ? Z[src + col * ld] : cmk(R(0), R(0));
                }
#pragma unroll
                for (int gg = 0; gg < GB; ++gg) {
                    const int e = e0 + gg * blockDim.x;
                    const int q = e % rows, col = e / rows;
                    if (e < total && !(col >= k0 && col < k0 + PB)) {
                        if (q < PB) Tg[q + (size_t)col * PB] = gv[gg];
                        else Dg[(q - PB) + (size_t)col * PB] = gv[gg];
                    }
                }
            }
        }
        __syncthreads();
        for (int e = t; e < nd * Np; e += blockDim.x) {
            const int q = e % nd, col = e / nd;
            if (col >= k0 && col < k0 + PB) continue;
            Z[s.disp[q] + col * ld] = Dg[q + (size_t)col * PB];
        }
        __syncthreads();
        // trailing update on the matrix cores, one chunk of SH::CH columns at a time;
        // the pivot rows themselves (this wave's rows when R0 == k0) start from zero.
        // Column tiles run in increasing order (panel tiles skipped) and the base values
        // of the next tile are loaded while the current one is on the matrix cores.
        const bool zb = (R0 == k0);
        const __amdgpu_buffer_rsrc_t zr = negf_rsrc(Z, (unsigned)(Np * Np) * G::ESZ);
        const unsigned cstr = (unsigned)G::CR * (unsigned)Np * G::ESZ;   // register r -> r + 1
        auto next_tile = [&](int ct) {                      // last tile: reload itself (unused)
            const int n = (ct + 16 == k0) ? k0 + PB : ct + 16;
            return n < Np ? n : ct;
        };
        auto voff = [&](int ct) { return (unsigned)((R0 + j) + (ct + G::CQ * k) * Np) * G::ESZ; };
        C nb0[4], nb1[4];
        auto load_base = [&](int ct) {
            const unsigned vo = voff(ct);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                nb0[q] = G::load(zr, vo, q * cstr);
                nb1[q] = G::load(zr, vo + 16u * G::ESZ, q * cstr);
            }
        };
        load_base(k0 == 0 ? PB : 0);
        for (int c0 = 0; c0 < Np; c0 += SH::CH) {
            const int c1 = min(c0 + SH::CH, Np);
            if (c0 >= k0 && c1 <= k0 + PB) continue;       // chunk is all panel
            for (int e = t; e < (c1 - c0) * PB; e += blockDim.x) {
                const int pp = e % PB, cl = e / PB;
                const int col = c0 + cl;
                if (!(col >= k0 && col < k0 + PB)) s.Tl[pp * SH::LDT + cl] = Tg[e + (size_t)c0 * PB];
            }
            __syncthreads();
            for (int ct = c0; ct < c1; ct += 16) {
                if (ct >= k0 && ct < k0 + PB) continue;
                Acc ar0, ai0, ar1, ai1;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    ar0[q] = zb ? R(0) : nb0[q].x;
                    ai0[q] = zb ? R(0) : nb0[q].y;
                    ar1[q] = zb ? R(0) : nb1[q].x;
                    ai1[q] = zb ? R(0) : nb1[q].y;
                }
                load_base(next_tile(ct));
                __builtin_amdgcn_sched_barrier(0);          // keep the prefetch ahead of the MFMAs
                int lx = threadIdx.x & 63;
                asm volatile("" : "+v"(lx));                 // recompute, do not keep live
                const C *tl = s.Tl + (lx >> 4) * SH::LDT + (ct - c0 + (lx & 15));
#pragma unroll
                for (int q = 0; q < KK; ++q) {
                    const C tv = tl[4 * q * SH::LDT];
                    const R ntim = -tv.y;
                    ar0 = G::mfma(tv.x, P[0][q].x, ar0);
                    ar1 = G::mfma(tv.x, P[1][q].x, ar1);
                    ai0 = G::mfma(tv.y, P[0][q].x, ai0);
                    ai1 = G::mfma(tv.y, P[1][q].x, ai1);
                    ar0 = G::mfma(ntim, P[0][q].y, ar0);
                    ar1 = G::mfma(ntim, P[1][q].y, ar1);
                    ai0 = G::mfma(tv.x, P[0][q].y, ai0);
                    ai1 = G::mfma(tv.x, P[1][q].y, ai1);
                }
                const unsigned vo = voff(ct);
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    G::store(cmk(ar0[q], ai0[q]), zr, vo, q * cstr);
                    G::store(cmk(ar1[q], ai1[q]), zr, vo + 16u * G::ESZ, q * cstr);
                }
            }
            __syncthreads();
        }
    }
    if (t == 0) {
        for (int c = 0; c < Np; ++c) s.sigma[c] = c;
        for (int c = Np - 1; c >= 0; --c) {
            const int p = s.ipiv[c];
            const int tt = s.sigma[c];
            s.sigma[c] = s.sigma[p];
            s.sigma[p] = tt;
        }
    }
    __syncthreads();
    return s.flag;
}

// ====================================================================================
// Block Gauss-Jordan without row interchanges, FP32 on the matrix cores (the sweep of the
// mixed-precision path).  Panels of 64 columns K = [k0, k0 + 64) (the last one 32 wide when
// Np % 64 == 32) are factored as four 16-column sub-blocks K_s:
//     P = M_{K_s K_s}^{-1}                          (FP64, 16 x 16, by the owner wave)
//     M_{I K_s} <- -M_{I K_s} P (I != K_s),  M_{K_s K_s} <- P
//     M_{I J} <- M_{I J} + M_{I K_s} M_{K_s J},  M_{K_s J} <- P M_{K_s J}   (J in K \ K_s)
// all in registers (the panel) with the 16 pivot rows exchanged through LDS; the panel then
// holds the Gauss-Jordan transform W of the whole 64-block, and the columns outside it get
// the one rank-64 update  Z_{I J} <- [I not in K] Z_{I J} + W_I Z_{K J}  per panel: 5 read +
// write passes over the block at Np = 320 instead of 10 with 32-column panels.  This is the
// in-place inversion of zgetri without pivoting; its stability rests on the RGF blocks
// D_m - g_{m+1}, and the call reports a failed growth check (P not finite, |P| above
// NEGF_BGJE_PMAX, a multiplier |M_IK P| above NEGF_BGJE_MMAX) so that the caller can
// compute the job on the pivoted FP64 path instead.  No column map: inverse = Z.
//
// Layout: wave w owns rows [32 w, 32 w + 32) as two 16-row tiles; lane l = j + 16 k holds
// the panel values Pre/Pim[t][g][r] = Z[32 w + 16 t + j][k0 + 16 g + 4 r + k] (real and
// imaginary parts).  That is the B operand of v_mfma_f32_16x16x4f32 for the K-step
// q = 4 g + r (B[k][j] at lane j + 16 k), and also the D layout of a product whose A
// operand rows are taken in the order c(i) = 4 (i % 4) + i / 4 (D[i][j] lands in lane
// j + 16 (i / 4), register i % 4), so the sub-block products accumulate straight into
// the panel registers.
// ====================================================================================
#define NEGF_BGJE_PMAX 1.0e6
#define NEGF_BGJE_MMAX 1.0e3

template <int MAXNP = NEGF_MAXNP4>
struct BgjeShared {
    // staged old pivot rows of one chunk of CH columns, column-major: Tl[cl * LDT + p]
    // (p < 64 the pivot row in the panel); the stride is padded by one complex value
    static constexpr int CH = 64, LDT = 64 + 1;
    static constexpr int LDP = 17, LDR = 65;        // 16 x 16 diagonal block, 16 pivot rows of the panel
    static constexpr int MAXW = 2 * MAXNP / 64;
    static constexpr int TSZ = CH * LDT;            // complex values in one staging buffer
    static constexpr int VHCAP = (TSZ * 8 + 2 * 16 * LDP * 16 + 2 * 16 * LDR * 8) / 16;
    union {
        struct {
            cfl Tl[TSZ];
            cd Pd[2][16 * LDP];  // M_{K_s K_s}, inverted in place in FP64 (alternating by sub-block)
            cfl Pr[2][16 * LDR]; // the sub-block's pivot rows over the panel columns, before the step
        };
        cd vh[VHCAP];            // outside the inversion: one slice's V-hat window (the kernel's)
    };
    static constexpr int WLSZ = 2 * 16 * 64;        // one donor wave's panel rows W, in LDS
    int sigma[MAXNP];    // identity (the consumers read inverse[:, c] = Z[:, sigma[c]])
    int bad;
    int job;             // the workgroup's current job (kernels that take jobs in turn)
    double red[MAXW];
    int simd[MAXW];      // the SIMD each wave runs on
    int help[MAXW];      // trailing-update roles (bgje_roles): -1 none, -2 donor, >= 0 the donor helped
    int wslot[MAXW];     // a donor's W area in the second staging buffer
};

// The trailing update's roles for balance across SIMDs (bgje_panel), from the SIMD each
// wave runs on (HW_ID bits 5:4): on each SIMD with the most waves the last wave becomes a
// donor, paired with the first wave of a SIMD with the fewest (at least one); every call
// of bgje_invert64 in the workgroup then uses them.  Call once, by every thread; ends with
// a barrier.  The result depends only on the wave placement, never on the data.
template <class SH>
__device__ __forceinline__ void bgje_roles(SH &s) {
    const int t = threadIdx.x, w = t >> 6, nw = blockDim.x >> 6;
    if ((t & 63) == 0) s.simd[w] = (__builtin_amdgcn_s_getreg((31 << 11) | (0 << 6) | 4) >> 4) & 3;
    __syncthreads();
    if (t == 0) {
        int cnt[4] = {0, 0, 0, 0}, first[4] = {-1, -1, -1, -1}, last[4] = {-1, -1, -1, -1};
        for (int v = 0; v < nw; ++v) {
            const int c = s.simd[v];
            ++cnt[c];
            if (first[c] < 0) first[c] = v;
            last[c] = v;
            s.help[v] = -1;
            s.wslot[v] = 0;
        }
        int mx = 0, mn = 1 << 30;
        for (int c = 0; c < 4; ++c) {
            mx = max(mx, cnt[c]);
            mn = min(mn, cnt[c]);
        }
        if (mn >= 1 && mx > mn) {
            int h = 0, l = 0, slot = 0;
            while (slot < 2) {
                while (h < 4 && cnt[h] != mx) ++h;
                while (l < 4 && cnt[l] != mn) ++l;
                if (h >= 4 || l >= 4) break;
                s.help[last[h]] = -2;
                s.wslot[last[h]] = slot++;
                s.help[first[l]] = last[h];
                ++h;
                ++l;
            }
        }
    }
    __syncthreads();
}

// In-place FP64 Gauss-Jordan inverse (no pivoting) of the 16 x 16 block A[r * 17 + c] in
// LDS by one wave: lane r + 16 h owns row r, columns [4 h, 4 h + 4).  Every lane reads its
// step operands before any lane of the wave writes (LDS operations of one wave are done in
// program order), so the steps need no barrier.  bad: a pivot reciprocal beyond
// NEGF_BGJE_PMAX (or not finite).
__device__ __forceinline__ void bgje_diag16(cd *A, bool &bad) {
    const int lane = opaque_tid() & 63, r = lane & 15, h = lane >> 4;
    cd *row = A + r * 17 + 4 * h;
#pragma unroll 1
    for (int J = 0; J < 16; ++J) {
        const cd *prow = A + J * 17 + 4 * h;
        const cd d = A[J * 17 + J];
        const cd m = A[r * 17 + J];
        cd pv[4], x[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            pv[c] = prow[c];
            x[c] = row[c];
        }
        const cd inv = crecip_fast(d);
        if (!(cabs1(inv) <= NEGF_BGJE_PMAX)) bad = true;
        const bool piv = (r == J);
        const cd f = piv ? cmk(-inv.x, -inv.y) : cmul(m, inv);
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const cd n = cfms(f, pv[c], piv ? cmk(0.0, 0.0) : x[c]);
            row[c] = (4 * h + c == J) ? cmk(-f.x, -f.y) : n;
        }
    }
}

// One sub-block step SB (static) of bgje_invert64's panel: see the comment above.
// after_sync() runs right after the step's barrier.  The lane indices are derived here
// from an opaque copy of the thread index: kept live across the panel instead, they and
// the LDS addresses built from them were spilled, and every step waited for the reloads.
template <int SB, class SH, class F>
__device__ __forceinline__ void bgje_substep(negf_f4 (&Pre)[2][4], negf_f4 (&Pim)[2][4], const int k0, const int nsb,
                                             const int w, SH &s, bool &bad, F &&after_sync) {
    constexpr int TS = SB & 1, BUF = SB & 1;
    typedef GjePrec<float> G;
    int lane = threadIdx.x & 63;
    asm volatile("" : "+v"(lane));
    const int j = lane & 15, k = lane >> 4;
    const int ci = ((lane & 3) << 2) | ((lane >> 2) & 3);      // A row order c(i), i = lane & 15
    const bool own = (w == (k0 >> 5) + (SB >> 1));      // wave-uniform
    if (own) {
        // publish the diagonal block in FP64 and the pivot rows over the panel, then invert
#pragma unroll
        for (int r = 0; r < 4; ++r) s.Pd[BUF][j * SH::LDP + 4 * r + k] = cmk((double)Pre[TS][SB][r], (double)Pim[TS][SB][r]);
#pragma unroll
        for (int g = 0; g < 4; ++g)
#pragma unroll
            for (int r = 0; r < 4; ++r) s.Pr[BUF][j * SH::LDR + 16 * g + 4 * r + k] = cmk(Pre[TS][g][r], Pim[TS][g][r]);
        bgje_diag16(s.Pd[BUF], bad);
    }
    __syncthreads();
    after_sync();
    // W = -M_{I K_s} P for the tiles outside K_s, W = P for the owner tile
#pragma unroll
    for (int tt = 0; tt < 2; ++tt) {
        if (own && tt == TS) {
            float pm = 0.0f;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const cd x = s.Pd[BUF][j * SH::LDP + 4 * r + k];
                Pre[tt][SB][r] = (float)x.x;
                Pim[tt][SB][r] = (float)x.y;
                pm = fmaxf(pm, fabsf(Pre[tt][SB][r]) + fabsf(Pim[tt][SB][r]));
            }
            if (!(pm <= (float)NEGF_BGJE_PMAX)) bad = true;
        } else {
            // 3M: c1 += Ar Br, c2 += Ai Bi, c3 += (Ar + Ai)(Br + Bi); re = c1 - c2, im = c3 - c1 - c2
            negf_f4 c1 = {0.f, 0.f, 0.f, 0.f}, c2 = {0.f, 0.f, 0.f, 0.f}, c3 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int qq = 0; qq < 4; ++qq) {
                const cd pd = s.Pd[BUF][(4 * qq + k) * SH::LDP + ci];
                const float ax = (float)-pd.x, ay = (float)-pd.y;
                const float bx = Pre[tt][SB][qq], by = Pim[tt][SB][qq];
                c1 = G::mfma(ax, bx, c1);
                c2 = G::mfma(ay, by, c2);
                c3 = G::mfma(ax + ay, bx + by, c3);
            }
            const negf_f4 ar = c1 - c2, ai = c3 - c1 - c2;
            float vm = 0.0f;
#pragma unroll
            for (int r = 0; r < 4; ++r) vm = fmaxf(vm, fabsf(ar[r]) + fabsf(ai[r]));
            if (!(vm <= (float)NEGF_BGJE_MMAX)) bad = true;
            Pre[tt][SB] = ar;
            Pim[tt][SB] = ai;
        }
    }
    // the other sub-blocks of the panel: M_{I J} += W_I M_{K_s J}; the owner tile starts from 0
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        if (g == SB || g >= nsb) continue;
#pragma unroll
        for (int tt = 0; tt < 2; ++tt) {
            const bool zero = own && tt == TS;
            const negf_f4 z4 = {0.f, 0.f, 0.f, 0.f};
            // 3M, from c1 = base re, c2 = 0, c3 = base re + base im (zero for the owner tile)
            negf_f4 c1 = zero ? z4 : Pre[tt][g];
            negf_f4 c3 = zero ? z4 : Pre[tt][g] + Pim[tt][g];
            negf_f4 c2 = z4;
#pragma unroll
            for (int qq = 0; qq < 4; ++qq) {
                const cfl av = s.Pr[BUF][(4 * qq + k) * SH::LDR + 16 * g + ci];
                const float bx = Pre[tt][SB][qq], by = Pim[tt][SB][qq];
                c1 = G::mfma(av.x, bx, c1);
                c2 = G::mfma(av.y, by, c2);
                c3 = G::mfma(av.x + av.y, bx + by, c3);
            }
            Pre[tt][g] = c1 - c2;
            Pim[tt][g] = c3 - c1 - c2;
        }
    }
}

// One panel of bgje_invert64: NSB = 4 sub-blocks (64 columns from k0), or 2 for the last
// panel when Np % 64 == 32.  The panel width is a template argument so that every panel
// register is indexed statically.
template <int NSB, class SH>
__device__ __forceinline__ void bgje_panel(cfl *__restrict__ Z, const int Np, const int k0, SH &s, cfl *T2,
                                           bool &bad) {
    typedef GjePrec<float> G;
    typedef negf_f4 Acc;
    constexpr int KW = 16 * NSB;
    const int t = opaque_tid(), lane = t & 63, nw = blockDim.x >> 6;
    const int w = __builtin_amdgcn_readfirstlane(t >> 6);
    const int j = lane & 15, k = lane >> 4;
    const int R0 = w << 5;
    const __amdgpu_buffer_rsrc_t zr = negf_rsrc(Z, (unsigned)(Np * Np) * 8u);
    const unsigned cstr = (unsigned)Np * 8u;                   // one column
    auto voff = [&](int row, int col) { return (unsigned)(row + col * Np) * 8u; };
    negf_f4 Pre[2][4], Pim[2][4];
#pragma unroll
    for (int tt = 0; tt < 2; ++tt)
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            Pre[tt][g] = negf_f4{0.f, 0.f, 0.f, 0.f};
            Pim[tt][g] = negf_f4{0.f, 0.f, 0.f, 0.f};
        }
    // panel loads and stores: one per-lane offset (from an opaque lane index, so that it is
    // recomputed rather than kept live) plus wave-uniform scalar offsets, instead of 32
    // per-lane addresses, which were spilled
    auto panel_voff = [&]() {
        int lx = threadIdx.x & 63;
        asm volatile("" : "+v"(lx));
        return voff(R0 + (lx & 15), k0 + (lx >> 4));
    };
    auto panel_soff = [&](int tt, int g, int r) { return (unsigned)(16 * tt + (16 * g + 4 * r) * Np) * 8u; };
    {
        const unsigned pv = panel_voff();
#pragma unroll
        for (int tt = 0; tt < 2; ++tt)
#pragma unroll
            for (int g = 0; g < NSB; ++g)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const cfl v = G::load(zr, pv, panel_soff(tt, g, r));
                    Pre[tt][g][r] = v.x;
                    Pim[tt][g][r] = v.y;
                }
    }
    // The trailing update's column chunks, skipping the panel's own.  The old pivot rows
    // Z_{K J} of a chunk go into LDS by DMA (buffer_load ... lds: one wave instruction
    // moves 64 floats, 32 pivot rows of one column, without passing through registers), so
    // a chunk's copy runs behind the panel or the previous chunk's products.  Those rows
    // keep their old values until the chunk's own update: the panel steps touch only the
    // panel columns.
    const int cfirst = (k0 == 0) ? SH::CH : 0;
    auto next_chunk = [&](int c0) { return (c0 + SH::CH == k0) ? k0 + SH::CH : c0 + SH::CH; };
    auto stage = [&](cfl *buf, const int c0) {
        constexpr int nh = KW / 32;
        const int ninst = (min(c0 + SH::CH, Np) - c0) * nh;
        for (int n = w; n < ninst; n += nw) {
            const int cl = n / nh, h = n - cl * nh;      // wave-uniform
            __builtin_amdgcn_raw_ptr_buffer_load_lds(zr, buf + cl * SH::LDT + 32 * h, 4,
                                                     voff(k0 + 32 * h, c0 + cl) + 4u * (unsigned)lane, 0, 0, 0);
        }
    };
    // the first chunk's DMA goes out after the last sub-block's barrier: issued earlier,
    // it would hold up the sub-block barriers (a barrier waits for the wave's memory
    // operations)
    auto none = [] {};
    auto first = [&] {
        if (cfirst < Np) stage(s.Tl, cfirst);
    };
    if constexpr (NSB == 4) {
        bgje_substep<0>(Pre, Pim, k0, NSB, w, s, bad, none);
        bgje_substep<1>(Pre, Pim, k0, NSB, w, s, bad, none);
        bgje_substep<2>(Pre, Pim, k0, NSB, w, s, bad, none);
        bgje_substep<3>(Pre, Pim, k0, NSB, w, s, bad, first);
    } else {
        bgje_substep<0>(Pre, Pim, k0, NSB, w, s, bad, none);
        bgje_substep<1>(Pre, Pim, k0, NSB, w, s, bad, first);
    }
    // the panel columns take their final values
    {
        const unsigned pv = panel_voff();
#pragma unroll
        for (int tt = 0; tt < 2; ++tt)
#pragma unroll
            for (int g = 0; g < NSB; ++g)
#pragma unroll
                for (int r = 0; r < 4; ++r) G::store(cmk(Pre[tt][g][r], Pim[tt][g][r]), zr, pv, panel_soff(tt, g, r));
    }
    // Trailing update, one chunk of CH columns at a time from the staged old pivot rows;
    // rows in K start from zero.  Balance across SIMDs: the waves of the workgroup sit on
    // the four SIMDs unevenly (ten waves: three, three, two, two), and the SIMDs with
    // three waves set the pace of every chunk.  So a donor wave there (s.help[w] == -2)
    // leaves the last column tile of each chunk to a helper wave on a SIMD with fewer
    // waves (s.help[w] = the donor), which computes it for the donor's rows after its own
    // tiles, with the donor's panel rows W taken from LDS (the donor writes them after the
    // panel store, ahead of the first chunk's barrier).  One tile of four measured faster
    // than two (9.50 s against 9.65 s per 2048-energy launch at Np = 320): the helper's
    // tiles read W from LDS, and its SIMD then became the slower one.
    const int role = __builtin_amdgcn_readfirstlane(s.help[w]);
    cfl *wlb = T2 + SH::TSZ + (role >= 0 ? s.wslot[role] : role == -2 ? s.wslot[w] : 0) * SH::WLSZ;
    if (role == -2) {
#pragma unroll
        for (int tt = 0; tt < 2; ++tt)
#pragma unroll
            for (int g = 0; g < NSB; ++g)
#pragma unroll
                for (int r = 0; r < 4; ++r) wlb[((16 * tt + 4 * g + r) << 6) + lane] = cmk(Pre[tt][g][r], Pim[tt][g][r]);
    }
    const int Rd = role >= 0 ? role << 5 : 0;
    const bool zb = (R0 >= k0 && R0 < k0 + KW), zbd = (Rd >= k0 && Rd < k0 + KW);
    Acc nb0r, nb0i, nb1r, nb1i;
    auto load_base = [&](int row0, int ct) {
        const unsigned vo = voff(row0 + j, ct + 4 * k);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const cfl a0 = G::load(zr, vo, r * cstr), a1 = G::load(zr, vo + 128u, r * cstr);
            nb0r[r] = a0.x;
            nb0i[r] = a0.y;
            nb1r[r] = a1.x;
            nb1i[r] = a1.y;
        }
    };
    load_base(R0, cfirst < Np ? cfirst : 0);
    // One 32-row x 16-column tile at (row0, ct) of the chunk in tb: its base was prefetched
    // into nb, the next item's base (nrow, ncol) is prefetched here, and with dma the wave's
    // share of the next chunk's DMA goes out after that prefetch (the memory counter is in
    // order: issued earlier, it would hold up the waits for the tile bases in between).
    // LB: the panel rows W come from LDS (wl) instead of this wave's registers.
    auto do_tile = [&](auto lds_b, const cfl *tb, cfl *tn, const int c0, const int cn, const int ct,
                       const int row0, const bool zbt, const cfl *wl, const int nrow, const int ncol,
                       const bool dma) {
        constexpr bool LB = decltype(lds_b)::value;
        // 3M complex products: with A = T (staged pivot rows) and B = W (panel),
        //   c1 += Ar Br,  c2 += Ai Bi,  c3 += (Ar + Ai)(Br + Bi),
        //   re = c1 - c2,  im = c3 - c1 - c2,
        // from c1 = base re, c2 = 0, c3 = base re + base im: three matrix-core products per
        // complex K-step instead of four
        Acc c10, c20, c30, c11, c21, c31;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            c10[r] = zbt ? 0.f : nb0r[r];
            c30[r] = zbt ? 0.f : nb0r[r] + nb0i[r];
            c11[r] = zbt ? 0.f : nb1r[r];
            c31[r] = zbt ? 0.f : nb1r[r] + nb1i[r];
            c20[r] = 0.f;
            c21[r] = 0.f;
        }
        load_base(nrow, ncol);
        if (dma) stage(tn, cn);
        __builtin_amdgcn_sched_barrier(0);          // keep the prefetch ahead of the MFMAs
        int lx = threadIdx.x & 63;
        asm volatile("" : "+v"(lx));                 // recompute, do not keep live
        const cfl *tl = tb + (ct - c0 + (lx & 15)) * SH::LDT + (lx >> 4);
#pragma unroll
        for (int q = 0; q < 4 * NSB; ++q) {
            const cfl tv = tl[4 * q];
            const float ts = tv.x + tv.y;
            float p0r, p0i, p1r, p1i;
            if constexpr (LB) {
                const cfl b0 = wl[(q << 6) + lx], b1 = wl[((16 + q) << 6) + lx];
                p0r = b0.x;
                p0i = b0.y;
                p1r = b1.x;
                p1i = b1.y;
            } else {
                p0r = Pre[0][q >> 2][q & 3];
                p0i = Pim[0][q >> 2][q & 3];
                p1r = Pre[1][q >> 2][q & 3];
                p1i = Pim[1][q >> 2][q & 3];
            }
            c10 = G::mfma(tv.x, p0r, c10);
            c11 = G::mfma(tv.x, p1r, c11);
            c20 = G::mfma(tv.y, p0i, c20);
            c21 = G::mfma(tv.y, p1i, c21);
            c30 = G::mfma(ts, p0r + p0i, c30);
            c31 = G::mfma(ts, p1r + p1i, c31);
        }
        const unsigned vo = voff(row0 + j, ct + 4 * k);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            G::store(cmk(c10[r] - c20[r], c30[r] - c10[r] - c20[r]), zr, vo, r * cstr);
            G::store(cmk(c11[r] - c21[r], c31[r] - c11[r] - c21[r]), zr, vo + 128u, r * cstr);
        }
    };
    // One chunk from the buffer tb while the next one (from cn, if any) is staged into tn.
    // The barrier first: it waits for this wave's DMA of the chunk (a barrier waits for the
    // wave's memory operations) and for every wave to be done with tn (the chunk before).
    // The two buffers alternate in a loop unrolled by two, so that each LDS read names its
    // buffer statically.
    auto run_chunk = [&](const cfl *tb, cfl *tn, const int c0, const int cn) {
        __builtin_amdgcn_s_waitcnt(0x0f70);             // vmcnt(0): this wave's DMA has landed
        __syncthreads();
        const int c1 = min(c0 + SH::CH, Np);
        const int ch = c1 - 16;                          // the tile the helper takes
        const int own_end = role == -2 ? ch : c1;
        const int nxt = cn < Np ? cn : c0;               // first tile of the next chunk (in bounds)
        for (int ct = c0; ct < own_end; ct += 16) {
            const bool last = ct + 16 >= own_end;
            const int nrow = (last && role >= 0) ? Rd : R0;
            const int ncol = !last ? ct + 16 : role >= 0 ? ch : nxt;
            do_tile(std::false_type{}, tb, tn, c0, cn, ct, R0, zb, wlb, nrow, ncol, last && cn < Np);
        }
        if (role >= 0) {
            for (int ct = ch; ct < c1; ct += 16) {
                const bool last = ct + 16 >= c1;
                do_tile(std::true_type{}, tb, tn, c0, cn, ct, Rd, zbd, wlb, last ? R0 : Rd, last ? nxt : ct + 16,
                        false);
            }
        }
    };
    for (int c0 = cfirst; c0 < Np;) {
        int cn = next_chunk(c0);
        run_chunk(s.Tl, T2, c0, cn);
        c0 = cn;
        if (c0 >= Np) break;
        cn = next_chunk(c0);
        run_chunk(T2, s.Tl, c0, cn);
        c0 = cn;
    }
    __syncthreads();      // the chunks' writes before the next panel; the buffers are free
}

// In-place inversion of Z (Np x Np, Np % 32 == 0, blockDim.x == 2 Np); on return
// inverse = Z and s.sigma is the identity.  Returns 0, or 1 if the growth check failed.
// T2: a second LDS staging buffer of SH::TSZ + 2 SH::WLSZ complex values, free for the
// call; s.help and s.wslot set by bgje_roles.
template <class SH>
__device__ __forceinline__ int bgje_invert64(cfl *__restrict__ Z, const int Np, SH &s, cfl *T2) {
    const int t = opaque_tid(), lane = t & 63;
    for (int i = t; i < Np; i += blockDim.x) s.sigma[i] = i;
    if (t == 0) s.bad = 0;
    bool bad = false;
    for (int k0 = 0; k0 < Np; k0 += 64) {
        if (k0 + 64 <= Np) bgje_panel<4>(Z, Np, k0, s, T2, bad);
        else bgje_panel<2>(Z, Np, k0, s, T2, bad);
    }
    if (__ballot(bad) && lane == 0) atomicOr(&s.bad, 1);
    __syncthreads();
    return __builtin_amdgcn_readfirstlane(s.bad);   // uniform: callers branch on it around barriers
}

// ====================================================================================
// The same matrix-core Gauss-Jordan for 512 < Np <= 1024: every wave owns NT = 4 row
// tiles (64 rows; ceil(Np / 64) <= 16 waves, so at most 1024 threads) and the panel is
// PB = 16 columns wide, which keeps the panel registers at NT * PB / 4 = 16 complex values
// per lane, as the two-tile, 32-column form has.  Lane l = j + 16 k holds
// P[t][q] = Z[R0 + 16 t + j][k0 + 4 q + k]; tiles at or past Np are skipped.  Pivot rule,
// interchanges, column map and the trailing update's MFMA layout are those of
// gje_invert4; the update runs one row tile at a time.
// ====================================================================================
template <int PB, int NT, int KJ, class R, class SH>
__device__ __forceinline__ void gje4n_step(typename GjePrec<R>::C (&P)[NT][PB / 4], const int g, const int k0,
                                           const int Np, const int w, const int j, const int k, SH &s,
                                           const int R0, const int ntv) {
    typedef typename GjePrec<R>::C C;
    constexpr int KK = PB / 4;
    const int t = threadIdx.x, lane = t & 63, nw = blockDim.x >> 6;
    const int J = 4 * g + KJ, c = k0 + J, par = J & 1;
    const bool mine = (k == KJ);
    R v[NT];
#pragma unroll
    for (int u = 0; u < NT; ++u) v[u] = R(-1);
    if (mine) {
#pragma unroll
        for (int u = 0; u < NT; ++u) {
            if (u < ntv) {
                const int r = R0 + 16 * u + j;
                if (r >= c) {
                    v[u] = cabs1(P[u][0]);
                    if (!(v[u] >= R(0))) v[u] = R(-0.5);
                }
                s.colJ[par][r] = P[u][0];
            }
        }
    }
    // izamax order: largest |re| + |im|, lowest row on ties (tiles in row order)
    R vl = v[0];
#pragma unroll
    for (int u = 1; u < NT; ++u) vl = v[u] > vl ? v[u] : vl;
    const R vmax = row16_max(vl, KJ);
    int idx = 0x7fffffff;
    if (vmax > R(-1)) {
#pragma unroll
        for (int u = NT - 1; u >= 0; --u) {
            const unsigned long long h = __ballot(v[u] == vmax);
            if (h) idx = R0 + 16 * u + ((__ffsll((long long)h) - 1) & 15);
        }
    }
    if (lane == 0) {
        s.rv[par][w] = vmax;
        s.ri[par][w] = idx;
    }
    const bool cnd = idx != 0x7fffffff && j == (idx & 15);
    const int ti = (idx - R0) >> 4;
    const bool rc = c >= R0 && c < R0 + 16 * ntv && j == (c & 15);
    const int tc = (c - R0) >> 4;
#pragma unroll
    for (int u = 0; u < NT; ++u) {
        if (cnd && ti == u) {
#pragma unroll
            for (int q = 0; q < KK; ++q) s.cand[par][w][4 * ((q + g) % KK) + k] = P[u][q];
        }
        if (rc && tc == u) {
#pragma unroll
            for (int q = 0; q < KK; ++q) s.rowC[par][4 * ((q + g) % KK) + k] = P[u][q];
        }
    }
    __syncthreads();
    // winner over the waves (at most 16, all in DPP row 0): first wave holding the max
    const R wv = lane < nw ? s.rv[par][lane] : R(-2);
    const int wi = lane < nw ? s.ri[par][lane] : 0x7fffffff;
    const R bmax = row16_max(wv, 0);
    const int wb = __ffsll((long long)__ballot(wv == bmax)) - 1;
    int p = __builtin_amdgcn_readlane(wi, wb);
    if (p < c || p >= Np) p = c;
    const C *prow = (p == c) ? s.rowC[par] : s.cand[par][wb];
    const C inv = crecip_fast(prow[J]);
    if (t == 0) {
        s.ipiv[c] = p;
        if (p != c) {
            const int tt = s.perm[c];
            s.perm[c] = s.perm[p];
            s.perm[p] = tt;
        }
        if (prow[J].x == R(0) && prow[J].y == R(0) && s.flag == 0) s.flag = c + 1;
    }
    // eliminate column J from the rows of this lane: the pivot row position takes the
    // scaled pivot row, the row that supplied the pivot takes old row c
#pragma unroll
    for (int u = 0; u < NT; ++u) {
        if (u < ntv) {
            const int r = R0 + 16 * u + j;
            C m = s.colJ[par][r];
            const bool piv = (r == c);
            if (r == p && p != c) {
                m = s.rowC[par][J];
#pragma unroll
                for (int q = 0; q < KK; ++q) P[u][q] = s.rowC[par][4 * ((q + g) % KK) + k];
            }
            if (piv) {
#pragma unroll
                for (int q = 0; q < KK; ++q) P[u][q] = cmk(R(0), R(0));
            }
            const C f = piv ? cmk(-inv.x, -inv.y) : cmul(m, inv);
#pragma unroll
            for (int q = 0; q < KK; ++q) {
                const C pq = prow[4 * ((q + g) % KK) + k];
                const C n = cfms(f, pq, P[u][q]);
                P[u][q] = (q == 0 && mine) ? cmk(-f.x, -f.y) : n;
            }
        }
    }
}

// In-place inversion of Z (Np x Np, 512 < Np <= 1024, Np % 32 == 0) by ceil(Np / 64)
// waves; on return inverse[:, c] = Z[:, s.sigma[c]].  scratch: (2 PB) x Np elements.
// Returns the 1-based column of the first exactly-zero pivot, or 0.
template <int PB, int NT, class R = double, class SH = Gje4Shared<PB, R>>
__device__ __forceinline__ int gje_invert4n(typename GjePrec<R>::C *__restrict__ Z, const int Np, SH &s,
                                            typename GjePrec<R>::C *__restrict__ scratch) {
    static_assert(PB == 16, "four 16-row tiles per wave hold one 16-column panel");
    typedef GjePrec<R> G;
    typedef typename G::C C;
    typedef typename G::Acc Acc;
    constexpr int KK = PB / 4;
    const int t = threadIdx.x, lane = t & 63;
    const int w = __builtin_amdgcn_readfirstlane(t >> 6);
    const int j = lane & 15, k = lane >> 4;
    const int R0 = w * 16 * NT;
    const int ntv = max(0, min(NT, (Np - R0) >> 4));      // tiles of this wave inside Np
    const size_t ld = (size_t)Np;
    C *Tg = scratch;                      // old pivot rows: Tg[p + c PB]
    C *Dg = scratch + (size_t)PB * Np;    // displaced rows: Dg[d + c PB]
    if (t == 0) s.flag = 0;
    for (int k0 = 0; k0 < Np; k0 += PB) {
        C P[NT][KK];
#pragma unroll
        for (int u = 0; u < NT; ++u)
#pragma unroll
            for (int q = 0; q < KK; ++q)
                P[u][q] = u < ntv ? Z[(R0 + 16 * u + j) + (size_t)(k0 + 4 * q + k) * ld] : cmk(R(0), R(0));
        for (int i = t; i < Np; i += blockDim.x) s.perm[i] = i;
        if (t == 0) s.ndisp = 0;
        __syncthreads();
#pragma unroll 1
        for (int g = 0; g < KK; ++g) {
            gje4n_step<PB, NT, 0, R>(P, g, k0, Np, w, j, k, s, R0, ntv);
            gje4n_step<PB, NT, 1, R>(P, g, k0, Np, w, j, k, s, R0, ntv);
            gje4n_step<PB, NT, 2, R>(P, g, k0, Np, w, j, k, s, R0, ntv);
            gje4n_step<PB, NT, 3, R>(P, g, k0, Np, w, j, k, s, R0, ntv);
#pragma unroll
            for (int u = 0; u < NT; ++u) {
                const C x = P[u][0];
#pragma unroll
                for (int q = 0; q < KK - 1; ++q) P[u][q] = P[u][q + 1];
                P[u][KK - 1] = x;
            }
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < NT; ++u)
            if (u < ntv) {
#pragma unroll
                for (int q = 0; q < KK; ++q) Z[(R0 + 16 * u + j) + (size_t)(k0 + 4 * q + k) * ld] = P[u][q];
            }
        if (k == 0) {
#pragma unroll
            for (int u = 0; u < NT; ++u) {
                const int r = R0 + 16 * u + j;
                if (u < ntv && r >= k0 + PB && s.perm[r] != r) s.disp[atomicAdd(&s.ndisp, 1)] = r;
            }
        }
        __syncthreads();
        const int nd = s.ndisp;
        const int rows = PB + nd;
        {
            constexpr int GB = 4;
            const int total = rows * Np;
            for (int e0 = t; e0 < total; e0 += GB * blockDim.x) {
                C gv[GB];
#pragma unroll
                for (int gg = 0; gg < GB; ++gg) {
                    const int e = e0 + gg * blockDim.x;
                    const int q = e % rows, col = e / rows;
                    const bool ok = e < total && !(col >= k0 && col < k0 + PB);
                    const int sr = ok ? (q < PB ? s.perm[k0 + q] : s.perm[s.disp[q - PB]]) : 0;
                    gv[gg] = ok ? Z[sr + col * ld] : cmk(R(0), R(0));
                }
#pragma unroll
                for (int gg = 0; gg < GB; ++gg) {
                    const int e = e0 + gg * blockDim.x;
                    const int q = e % rows, col = e / rows;
                    if (e < total && !(col >= k0 && col < k0 + PB)) {
                        if (q < PB) Tg[q + (size_t)col * PB] = gv[gg];
                        else Dg[(q - PB) + (size_t)col * PB] = gv[gg];
                    }
                }
            }
        }
        __syncthreads();
        for (int e = t; e < nd * Np; e += blockDim.x) {
            const int q = e % nd, col = e / nd;
            if (col >= k0 && col < k0 + PB) continue;
            Z[s.disp[q] + col * ld] = Dg[q + (size_t)col * PB];
        }
        __syncthreads();
        // trailing update on the matrix cores, one chunk of SH::CH columns at a time and
        // one row tile at a time; the pivot rows themselves start from zero
        for (int c0 = 0; c0 < Np; c0 += SH::CH) {
            const int c1 = min(c0 + SH::CH, Np);
            if (c0 >= k0 && c1 <= k0 + PB) continue;       // chunk is all panel
            for (int e = t; e < (c1 - c0) * PB; e += blockDim.x) {
                const int pp = e % PB, cl = e / PB;
                const int col = c0 + cl;
                if (!(col >= k0 && col < k0 + PB)) s.Tl[pp * SH::LDT + cl] = Tg[e + (size_t)c0 * PB];
            }
            __syncthreads();
            for (int ct = c0; ct < c1; ct += 16) {
                if (ct >= k0 && ct < k0 + PB) continue;
                const C *tl = s.Tl + k * SH::LDT + (ct - c0 + j);
                C tv[KK];
#pragma unroll
                for (int q = 0; q < KK; ++q) tv[q] = tl[4 * q * SH::LDT];
#pragma unroll
                for (int u = 0; u < NT; ++u) {
                    if (u < ntv) {
                        const int rt = R0 + 16 * u;
                        const bool zb = rt >= k0 && rt < k0 + PB;
                        C *zc = Z + (rt + j) + (size_t)(ct + G::CQ * k) * ld;
                        Acc ar, ai;
#pragma unroll
                        for (int q = 0; q < 4; ++q) {
                            const C b = zb ? cmk(R(0), R(0)) : zc[(size_t)(G::CR * q) * ld];
                            ar[q] = b.x;
                            ai[q] = b.y;
                        }
#pragma unroll
                        for (int q = 0; q < KK; ++q) {
                            ar = G::mfma(tv[q].x, P[u][q].x, ar);
                            ai = G::mfma(tv[q].y, P[u][q].x, ai);
                            ar = G::mfma(-tv[q].y, P[u][q].y, ar);
                            ai = G::mfma(tv[q].x, P[u][q].y, ai);
                        }
#pragma unroll
                        for (int q = 0; q < 4; ++q) zc[(size_t)(G::CR * q) * ld] = cmk(ar[q], ai[q]);
                    }
                }
            }
            __syncthreads();
        }
    }
    if (t == 0) {
        for (int c = 0; c < Np; ++c) s.sigma[c] = c;
        for (int c = Np - 1; c >= 0; --c) {
            const int p = s.ipiv[c];
            const int tt = s.sigma[c];
            s.sigma[c] = s.sigma[p];
            s.sigma[p] = tt;
        }
    }
    __syncthreads();
    return s.flag;
}
