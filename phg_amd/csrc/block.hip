// Many right-hand sides: Jacobi-PCG over a block of up to KB = 16 vectors.
//
// PHG's PCG solves one vector at a time (solver-phg.c:495 asserts
// nvec == 1); here the KB recurrences advance together so that every
// iteration reads the matrix once for all of them: the SpMV becomes an SpMM
// over multivectors stored row-major, [n][KB], so the KB values of one row
// are one contiguous 64 / 128-byte segment.  Each column keeps its own
// alpha, beta and stopping test (solver-phg.c:535-580 per column); a column
// that has converged stops changing.
//
// Matrix: the solver copy (Morton numbering).  For vector problems (3
// components per node) the scalar CSR is regrouped into 3x3 blocks (BSR3),
// one column index per 9 values; scalar problems use the CSR itself.
//
// Mixed precision: FP32 SpMM and vectors with reliable updates -- when a
// column's FP32 residual has dropped by delta below its largest value since
// the previous replacement, or reaches the tolerance, the FP32 increments of
// all columns are folded into x (FP64) and every residual is replaced by
// the FP64 residual b - A x from an FP64 SpMM; convergence is only declared
// on FP64 residuals.
//
// One iteration slot launches
//   k_blk_ctrl   (1 work-group) column totals -> status, beta, slot kind
//   k_blk_p      p = z + beta p  (or, in an update slot, x += x32)
//   k_blk_spmm   Ap = A p, partial (p, Ap) per column
//   k_blk_alpha  (1 work-group) alpha per column
//   k_blk_update x32 += alpha p, r -= alpha Ap, partial (r,r), (r,z)
//   k_blk_resid  update slots only: r = b - A x in FP64, partials
#include "phg_hip_internal.h"

#include <math.h>
#include <string.h>
#include <algorithm>
#include <type_traits>
#include <vector>

namespace {

constexpr int KB = 16;            // vectors per block
constexpr int TB = 256;           // work-group size
constexpr int NXCD = 8;
constexpr int MAXG = 1024;        // partial-sum slots per value and column
constexpr int BSR_MAXB = 128;     // blocks per block row (builder limit)

__device__ inline int64_t xcd_block()
{
    int64_t nb = gridDim.x, b = blockIdx.x;
    int64_t x = b % NXCD, k = b / NXCD;
    int64_t q = nb / NXCD, rem = nb % NXCD;
    return x * q + (x < rem ? x : rem) + k;
}


//------------------------------------------------------------- BSR3 build
// block row R = scalar rows 3R..3R+2; blocks in the order their columns
// first appear scanning the three rows in storage order
__global__ void k_bsr_count(int64_t nb, const int32_t *rp, const int32_t *ci,
                            int32_t *cnt, int32_t *overflow)
{
    int64_t R = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (R >= nb)
        return;
    int32_t seen[BSR_MAXB];
    int k = 0;
    for (int p = 0; p < 3; p++)
        for (int32_t j = rp[3 * R + p]; j < rp[3 * R + p + 1]; j++) {
            int32_t C = ci[j] / 3;
            int i = 0;
            while (i < k && seen[i] != C)
                i++;
            if (i == k) {
                if (k == BSR_MAXB) {
                    *overflow = 1;
                    cnt[R] = 0;
                    return;
                }
                seen[k++] = C;
            }
        }
    cnt[R] = k;
}

template <typename T>
__global__ void k_bsr_fill(int64_t nb, const int32_t *rp, const int32_t *ci,
                           const double *va, const int64_t *bptr,
                           int32_t *bcol, T *bval)
{
    int64_t R = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (R >= nb)
        return;
    const int64_t b0 = bptr[R];
    int k = 0;
    for (int64_t i = b0; i < bptr[R + 1]; i++)
        for (int t = 0; t < 9; t++)
            bval[i * 9 + t] = (T)0;
    for (int p = 0; p < 3; p++)
        for (int32_t j = rp[3 * R + p]; j < rp[3 * R + p + 1]; j++) {
            int32_t C = ci[j] / 3, q = ci[j] % 3;
            int i = 0;
            while (i < k && bcol[b0 + i] != C)
                i++;
            if (i == k)
                bcol[b0 + k++] = C;
            bval[(b0 + i) * 9 + p * 3 + q] = (T)va[j];
        }
}

// S_R = sum of the blocks of block row R, summed in FP64, stored in FP32
template <int B>
__global__ void k_bsr_rsum(int64_t nb, const int32_t *bptr, const double *va, float *out)
{
    int64_t R = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (R >= nb)
        return;
    double s[B * B];
    for (int t = 0; t < B * B; t++)
        s[t] = 0.0;
    for (int32_t j = bptr[R]; j < bptr[R + 1]; j++)
        for (int t = 0; t < B * B; t++)
            s[t] += va[(int64_t)j * (B * B) + t];
    for (int t = 0; t < B * B; t++)
        out[R * (B * B) + t] = (float)s[t];
}

__global__ void k_i64_to_i32(int64_t n, const int64_t *a, int32_t *b)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i <= n)
        b[i] = (int32_t)a[i];
}

//------------------------------------------------------------------ SpMM
// Block CSR with B x B blocks (B = 1: the CSR itself).  A lane works on CPL
// consecutive columns of the multivector (CPL = 1, or 4 with 16-byte
// loads); NCG = KB / CPL column groups.  Lane mapping inside a wavefront:
// lane = (row * S + s) * NCG + cg, 64 / (S * NCG) block rows per wavefront.
// Slot s takes the row's blocks s, s + S, ...; the S partial sums are
// combined with shuffles.
struct BsrView {
    const int32_t *bptr, *bcol;
    int64_t nb;          // block rows
};

template <typename T, int CPL> struct Cols { T v[CPL]; };

template <typename T, int CPL>
__device__ inline Cols<T, CPL> load_cols(const T *p)
{
    Cols<T, CPL> r;
    if (CPL == 1)
        r.v[0] = p[0];
    else if (sizeof(T) == 4) {
        const float4 a = *(const float4 *)p;
        r.v[0] = a.x; r.v[1] = a.y; r.v[2] = a.z; r.v[3] = a.w;
    }
    else {
        const double2 a = ((const double2 *)p)[0], b = ((const double2 *)p)[1];
        r.v[0] = a.x; r.v[1] = a.y; r.v[2] = b.x; r.v[3] = b.y;
    }
    return r;
}

// value of lane k (0..3) of each quad of lanes, by DPP quad_perm
template <int K>
__device__ inline int quad_bcast_i(int v)
{
    return __builtin_amdgcn_mov_dpp(v, K | K << 2 | K << 4 | K << 6, 0xf, 0xf, false);
}
__device__ inline int quad_bcast_i(int v, int k)
{
    switch (k) {
    case 0: return quad_bcast_i<0>(v);
    case 1: return quad_bcast_i<1>(v);
    case 2: return quad_bcast_i<2>(v);
    default: return quad_bcast_i<3>(v);
    }
}
__device__ inline float quad_bcast(float v, int k)
{
    return __int_as_float(quad_bcast_i(__float_as_int(v), k));
}
__device__ inline double quad_bcast(double v, int k)
{
    const int64_t b = __double_as_longlong(v);
    const int lo = quad_bcast_i((int)(b & 0xffffffff), k);
    const int hi = quad_bcast_i((int)(b >> 32), k);
    return __longlong_as_double((int64_t)(uint32_t)lo | ((int64_t)hi << 32));
}

// block rows one work-group step covers for a shape
template <int S, int CPL>
constexpr int rows_per_group() { return (64 / (S * (KB / CPL))) * (TB / 64); }

// REL: the products act on X_C - X_R (R the block row, C the block column)
// and the row-block sum S_R = sum_C A_RC adds S_R X_R at the end.  Exact
// algebra; in FP32 the products then round relative to the variation of X
// across the row's stencil rather than to X itself.  The near-kernel modes
// of the operator (constants, rigid translations), which the Krylov method
// resolves last, lose most of their FP32 error that way.
template <typename T, int B, int S, int CPL, typename Epi, bool REL = false>
__device__ inline void bsr_spmm(const BsrView &m, const T *__restrict__ va,
                                const T *__restrict__ X, Epi epi,
                                const T *__restrict__ rsum = nullptr)
{
    constexpr int NCG = KB / CPL;
    constexpr int LPR = S * NCG;         // lanes per block row
    constexpr int RPW = 64 / LPR;        // block rows per wavefront
    constexpr int RPG = RPW * (TB / 64); // block rows per work-group step
    static_assert(LPR <= 64 && 64 % LPR == 0, "wavefront shape");
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int cg = lane % NCG, s = (lane / NCG) % S, rw = lane / LPR;
    int64_t g0, g1, gstep;
    phg::xcd_sweep((m.nb + RPG - 1) / RPG, g0, g1, gstep);
    for (int64_t gi = g0; gi < g1; gi += gstep) {
        const int64_t R = gi * RPG + wv * RPW + rw;
        T acc[B][CPL];
#pragma unroll
        for (int p = 0; p < B; p++)
#pragma unroll
            for (int i = 0; i < CPL; i++)
                acc[p][i] = 0;
        Cols<T, CPL> xr[B];
        if (REL && R < m.nb)
#pragma unroll
            for (int q = 0; q < B; q++)
                xr[q] = load_cols<T, CPL>(X + (R * B + q) * KB + cg * CPL);
        if (R < m.nb) {
            // the row's block columns, loaded once (coalesced) by the
            // row's lanes and handed out by shuffles, so the X gathers of
            // consecutive blocks do not wait on a column load each
            const int32_t j0 = m.bptr[R], nblk = m.bptr[R + 1] - j0;
            const int base = lane - lane % LPR, lr = lane % LPR;
            constexpr int NT = BSR_MAXB / LPR;
            int32_t colt[NT];
#pragma unroll
            for (int t = 0; t < NT; t++)
                colt[t] = t * LPR + lr < nblk ? m.bcol[j0 + t * LPR + lr] : 0;
            // slot s takes blocks s, s + S, ... (t * LPR + s + S v)
#pragma unroll
            for (int t = 0; t < NT; t++) {
                if (t * LPR >= nblk)
                    break;
#pragma unroll
                for (int v = 0; v < LPR / S; v++) {
                    const int jj = t * LPR + s + S * v;
                    const int64_t C = __shfl(colt[t], base + s + S * v, 64);
                    if (jj < nblk) {
                        Cols<T, CPL> xq[B];
#pragma unroll
                        for (int q = 0; q < B; q++) {
                            xq[q] = load_cols<T, CPL>(X + (C * B + q) * KB + cg * CPL);
                            if (REL)
#pragma unroll
                                for (int i = 0; i < CPL; i++)
                                    xq[q].v[i] -= xr[q].v[i];
                        }
                        const T *a = va + (int64_t)(j0 + jj) * (B * B);
                        if constexpr (NCG == 4 && B == 3) {
                            // the four column-group lanes of a quad share
                            // the block: lane cg loads entries cg, cg + 4
                            // and 8 (cg = 0; the others repeat entry cg),
                            // and the entries go round the quad by DPP, so
                            // the block is fetched once, not four times
                            const T ld[3] = {a[cg], a[cg + 4], a[cg == 0 ? 8 : cg]};
#pragma unroll
                            for (int p = 0; p < B; p++)
#pragma unroll
                                for (int q = 0; q < B; q++) {
                                    const int e = p * B + q;
                                    const T apq = quad_bcast(ld[e >> 2], e & 3);
#pragma unroll
                                    for (int i = 0; i < CPL; i++)
                                        acc[p][i] += apq * xq[q].v[i];
                                }
                        }
                        else {
#pragma unroll
                            for (int p = 0; p < B; p++)
#pragma unroll
                                for (int q = 0; q < B; q++) {
                                    const T apq = a[p * B + q];
#pragma unroll
                                    for (int i = 0; i < CPL; i++)
                                        acc[p][i] += apq * xq[q].v[i];
                                }
                        }
                    }
                }
            }
        }
#pragma unroll
        for (int o = NCG; o < LPR; o <<= 1)
#pragma unroll
            for (int p = 0; p < B; p++)
#pragma unroll
                for (int i = 0; i < CPL; i++)
                    acc[p][i] += __shfl_xor(acc[p][i], o, 64);
        if (R < m.nb && s == 0) {
            if (REL)
#pragma unroll
                for (int p = 0; p < B; p++)
#pragma unroll
                    for (int q = 0; q < B; q++) {
                        const T spq = rsum[R * (B * B) + p * B + q];
#pragma unroll
                        for (int i = 0; i < CPL; i++)
                            acc[p][i] += spq * xr[q].v[i];
                    }
#pragma unroll
            for (int p = 0; p < B; p++)
#pragma unroll
                for (int i = 0; i < CPL; i++)
                    epi(R * B + p, cg * CPL + i, i, acc[p][i]);
        }
    }
}

// per-column sums of a work-group: v[k][i] held by the lanes of column
// group cg (columns cg * CPL + i, cg = lane % (KB / CPL)); written to
// part[(k * KB + c) * G + block]
template <int NV, int CPL>
__device__ inline void write_col_partials(double (&v)[NV][CPL], double *part)
{
    constexpr int NCG = KB / CPL;
    __shared__ double sh[NV][TB / 64][KB];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; k++)
#pragma unroll
        for (int i = 0; i < CPL; i++) {
            double t = v[k][i];
#pragma unroll
            for (int o = NCG; o < 64; o <<= 1)
                t += __shfl_xor(t, o, 64);
            if (lane < NCG)
                sh[k][wv][lane * CPL + i] = t;
        }
    __syncthreads();
    if (threadIdx.x < NV * KB) {
        const int k = threadIdx.x / KB, c = threadIdx.x % KB;
        double t = 0.0;
#pragma unroll
        for (int w = 0; w < TB / 64; w++)
            t += sh[k][w][c];
        part[((int64_t)k * KB + c) * gridDim.x + blockIdx.x] = t;
    }
}

//--------------------------------------------------------------- state
constexpr int ST_ITER = 1, ST_UPDATE = 2, ST_START = 0;
struct BState {
    double rz[KB], res[KB], tol[KB], rmax[KB], alpha[KB], beta[KB], r0[KB];
    double rtol, btol, atol, delta;
    int32_t active[KB], nits[KB], first[KB];
    int32_t zero_x[KB];   // b == 0: x = 0 (solver-phg.c:511-518)
    int32_t kind, maxit, mixed, updates, all_done, slots;
};

// column totals of NV values from `np` partials each
template <int NV>
__device__ inline void col_totals(const double *part, int np, double (&t)[NV][KB])
{
    __shared__ double sh[NV * KB][TB / 16];
    // 16 threads per (value, column) pair would need NV*KB*16 threads; use
    // a strided loop: thread t handles pair t % (NV*KB), chunk t / (NV*KB)
    const int pairs = NV * KB;
    const int per = TB / pairs;           // threads per pair (>= 1)
    const int pr = threadIdx.x % pairs, ch = threadIdx.x / pairs;
    double a = 0.0;
    if (ch < per)
        for (int i = ch; i < np; i += per)
            a += part[(int64_t)pr * np + i];
    if (ch < per)
        sh[pr][ch] = a;
    __syncthreads();
    if (threadIdx.x < pairs) {
        double s = 0.0;
        for (int k = 0; k < per; k++)
            s += sh[threadIdx.x][k];
        sh[threadIdx.x][0] = s;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NV; k++)
#pragma unroll
        for (int c = 0; c < KB; c++)
            t[k][c] = sh[k * KB + c][0];
    __syncthreads();
}

// Reads the column totals left by the previous slot -- (r,r), (r,z),
// #zero diagonals, (b,b) -- from the iteration kernel (after an iteration
// slot) or from the FP64 residual kernel (after the prologue or an update
// slot), advances every column's recurrence and picks the next slot kind.
__global__ void __launch_bounds__(TB) k_blk_ctrl(const double *part_it, int np_it,
                                                const double *part_up, int np_up,
                                                BState *st)
{
    const int last = st->kind;
    const bool from_update = last != ST_ITER;
    double t[4][KB];
    if (from_update)
        col_totals<4>(part_up, np_up, t);
    else
        col_totals<4>(part_it, np_it, t);
    if (threadIdx.x != 0)
        return;
    BState &S = *st;
    if (S.all_done)
        return;
    bool want_update = false;
    int alive = 0;
    double nbeta[KB], nrz[KB];
    int nfirst[KB];
    for (int c = 0; c < KB; c++) {
        if (!S.active[c])
            continue;
        const double res = sqrt(t[0][c]), rz = t[1][c];
        if (last == ST_ITER)
            S.nits[c]++;              // the slot just run iterated column c
        if (last == ST_START) {
            // get_tol (solver-phg.c:104) per column on the initial residual
            const double rb = sqrt(t[3][c]);
            double tol = res * S.rtol;
            if (tol < rb * S.btol)
                tol = rb * S.btol;
            if (tol < S.atol)
                tol = S.atol;
            S.tol[c] = tol;
            S.r0[c] = res;
            S.res[c] = res;
            S.rmax[c] = res;
            if (rb < 10.0 * 2.2250738585072014e-308) {
                S.zero_x[c] = 1;
                S.res[c] = 0.0;
            }
            if (S.zero_x[c] || !(res > tol) || S.maxit <= 0 || t[2][c] != 0.0) {
                S.active[c] = 0;
                continue;
            }
            nbeta[c] = 0.0;
            nrz[c] = rz;
            nfirst[c] = 1;
            alive++;
            continue;
        }
        S.res[c] = res;
        const bool fp64_res = !S.mixed || from_update;
        if (fp64_res && !(res > S.tol[c])) {
            S.active[c] = 0;          // converged on an FP64 residual
            continue;
        }
        if (S.nits[c] >= S.maxit) {
            if (S.mixed && !from_update)
                want_update = true;   // finish on an FP64 residual
            else {
                S.active[c] = 0;
                continue;
            }
        }
        if (S.mixed && !from_update &&
            (!(res > S.tol[c]) || res < S.delta * S.rmax[c]))
            want_update = true;
        if (from_update || res > S.rmax[c])
            S.rmax[c] = res;
        if (S.rz[c] == 0.0) {       // solver-phg.c:547 breakdown
            S.active[c] = 0;
            continue;
        }
        nbeta[c] = rz / S.rz[c];
        nrz[c] = rz;
        nfirst[c] = 0;
        alive++;
    }
    if (alive == 0) {
        S.all_done = 1;
        S.kind = ST_ITER;
        return;
    }
    S.kind = want_update ? ST_UPDATE : ST_ITER;
    if (S.kind == ST_UPDATE) {
        // the FP32 totals are discarded: after the FP64 residual the next
        // slot takes beta from it against the rz that built the current p
        S.updates++;
        return;
    }
    for (int c = 0; c < KB; c++)
        if (S.active[c]) {
            S.beta[c] = nbeta[c];
            S.rz[c] = nrz[c];
            S.first[c] = nfirst[c];
        }
    S.slots++;
}

template <typename T>
__global__ void __launch_bounds__(TB) k_blk_p(int64_t n, const T *r,
                                             const T *diag, T *p, double *x,
                                             T *x32, const BState *st)
{
    const BState &S = *st;
    if (S.all_done)
        return;
    const int c = threadIdx.x % KB;
    const int64_t total = n * KB;
    if (S.kind == ST_UPDATE) {
        for (int64_t i = (int64_t)blockIdx.x * TB + threadIdx.x; i < total;
             i += (int64_t)gridDim.x * TB) {
            x[i] += (double)x32[i];
            x32[i] = (T)0;
        }
        return;
    }
    const bool act = S.active[c] != 0, first = S.first[c] != 0;
    const T beta = (T)S.beta[c];
    if (!act)
        return;
    for (int64_t i = (int64_t)blockIdx.x * TB + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * TB) {
        const T z = r[i] / diag[i / KB];
        p[i] = first ? z : beta * p[i] + z;
    }
}

// the FP32 SpMM of the mixed solver runs in the relative form (rsum: the
// FP32 row-block sums)
template <typename T, int B, int S, int CPL>
__global__ void __launch_bounds__(TB) k_blk_spmm(BsrView m, const T *va,
                                                const T *p, T *Ap,
                                                double *part, const BState *st,
                                                const T *rsum)
{
    if (st->all_done || st->kind != ST_ITER)
        return;
    double v[1][CPL] = {};
    auto epi = [&](int64_t row, int c, int i, T s) {
        Ap[row * KB + c] = s;
        v[0][i] += (double)p[row * KB + c] * (double)s;
    };
    bsr_spmm<T, B, S, CPL, decltype(epi), sizeof(T) == 4>(m, va, p, epi, rsum);
    write_col_partials<1, CPL>(v, part);
}

__global__ void __launch_bounds__(TB) k_blk_alpha(const double *part, int np,
                                                 BState *st)
{
    if (st->all_done || st->kind != ST_ITER)
        return;
    double t[1][KB];
    col_totals<1>(part, np, t);
    if (threadIdx.x != 0)
        return;
    for (int c = 0; c < KB; c++) {
        if (!st->active[c])
            continue;
        if (t[0][c] == 0.0) {           // solver-phg.c:561
            st->active[c] = 0;
            st->alpha[c] = 0.0;
            continue;
        }
        st->alpha[c] = st->rz[c] / t[0][c];
    }
}

template <typename T>
__global__ void __launch_bounds__(TB) k_blk_update(int64_t n, T *x, const T *p,
                                                  T *r, const T *Ap,
                                                  const T *diag, double *part,
                                                  const BState *st)
{
    if (st->all_done || st->kind != ST_ITER)
        return;
    const int c = threadIdx.x % KB;
    const bool act = st->active[c] != 0;
    const T alpha = (T)st->alpha[c];
    double v[4][1] = {};
    const int64_t total = n * KB;
    for (int64_t i = (int64_t)blockIdx.x * TB + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * TB) {
        T ri = r[i];
        if (act) {
            x[i] += alpha * p[i];
            ri = ri - alpha * Ap[i];
            r[i] = ri;
        }
        const T z = ri / diag[i / KB];
        v[0][0] += (double)ri * (double)ri;
        v[1][0] += (double)ri * (double)z;
    }
    write_col_partials<4, 1>(v, part);
}

// FP64 residual r = b - A x for every column; r_out (T) receives it;
// partials (r,r), (r,z), #zero diagonals, (b,b)
template <typename T, int B, int S, int CPL>
__global__ void __launch_bounds__(TB) k_blk_resid(BsrView m, const double *va,
                                                 const double *diag,
                                                 const double *b,
                                                 const double *x, T *r_out,
                                                 T *x32_zero, double *part,
                                                 const BState *st, int force)
{
    if (!force && (st->all_done || st->kind != ST_UPDATE))
        return;
    double v[4][CPL] = {};
    bsr_spmm<double, B, S, CPL>(m, va, x, [&](int64_t row, int c, int k, double s) {
        const int64_t i = row * KB + c;
        const double bi = b[i], ri = bi - s, d = diag[row];
        r_out[i] = (T)ri;
        if (x32_zero)
            x32_zero[i] = (T)0;
        v[0][k] += ri * ri;
        if (d == 0.0)
            v[2][k] += 1.0;
        else
            v[1][k] += ri * (ri / d);
        v[3][k] += bi * bi;
    });
    write_col_partials<4, CPL>(v, part);
}

// pack / unpack between KB separate vectors (PHG numbering) and the
// row-major multivector in the solver numbering
struct PtrTab {
    const double *src[KB];
    double *dst[KB];
};

__global__ void k_blk_pack(int64_t n, int nrhs, const int32_t *perm,
                           PtrTab t, double *out)
{
    const int64_t total = n * KB;
    for (int64_t i = (int64_t)blockIdx.x * TB + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * TB) {
        const int64_t row = i / KB;
        const int c = (int)(i % KB);
        const int64_t src = perm ? perm[row] : row;
        out[i] = c < nrhs ? t.src[c][src] : 0.0;
    }
}

// zero_mask bit c: column c is written as zero
__global__ void k_blk_unpack(int64_t n, int nrhs, const int32_t *perm,
                             const double *in, PtrTab t, int zero_mask)
{
    const int64_t total = n * KB;
    for (int64_t i = (int64_t)blockIdx.x * TB + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * TB) {
        const int64_t row = i / KB;
        const int c = (int)(i % KB);
        if (c < nrhs)
            t.dst[c][perm ? perm[row] : row] = (zero_mask >> c) & 1 ? 0.0 : in[i];
    }
}

__global__ void k_blk_fold(int64_t total, double *x, const float *x32)
{
    for (int64_t i = (int64_t)blockIdx.x * TB + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * TB)
        x[i] += (double)x32[i];
}

inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

int grid_vec(int64_t total)
{
    int cus = phg::ctx()->num_cus > 0 ? phg::ctx()->num_cus : 256;
    return (int)std::max<int64_t>(1, std::min<int64_t>((total + TB - 1) / TB,
                                                       std::min(cus * 4, MAXG)));
}

template <typename K>
int grid_res(const void *kern, int64_t units)
{
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, TB, 0) !=
            hipSuccess ||
        per_cu <= 0)
        per_cu = 4;
    int cus = phg::ctx()->num_cus > 0 ? phg::ctx()->num_cus : 256;
    return (int)std::max<int64_t>(
        1, std::min<int64_t>(units, std::min(per_cu * cus, MAXG)));
}

}  // namespace

namespace phg {

void free_bsr(phg_hip_mat *A)
{
    if (A->b_owns) {
        dfree(A->b_ptr);
        dfree(A->b_col);
    }
    dfree(A->b_val64);
    dfree(A->b_val32);
    dfree(A->b_rsum32);
    A->b_rsum32 = nullptr;
    A->b_ptr = A->b_col = nullptr;
    A->b_val64 = nullptr;
    A->b_val32 = nullptr;
    A->b_owns = 0;
    A->b_nb = A->b_nblk = 0;
    A->b_dim = 0;
}

// block copy of the solver matrix: 3x3 blocks for 3-component problems,
// the CSR itself (1x1 blocks) otherwise
int build_bsr(phg_hip_mat *A)
{
    int rc;
    if ((rc = ensure_solver_copy(A)))
        return rc;
    if (A->b_dim != 0)
        return PHG_HIP_OK;
    hipStream_t s = ctx()->stream;
    if ((rc = ensure_fp32_copy(A)))
        return rc;
    if (A->dim != 3 || A->n % 3 != 0) {
        A->b_ptr = A->q_rowptr;
        A->b_col = A->q_cols;
        A->b_nb = A->n;
        A->b_nblk = A->nnz;
        A->b_owns = 0;
        A->b_dim = 1;
        if ((rc = dmalloc((void **)&A->b_rsum32, (size_t)A->n * 4))) {
            free_bsr(A);
            return rc;
        }
        hipLaunchKernelGGL(k_bsr_rsum<1>, dim3(nblk(A->n, 128)), dim3(128), 0, s, A->n,
                           (const int32_t *)A->q_rowptr, (const double *)A->q_vals,
                           A->b_rsum32);
        PHG_HIP_CHECK(hipGetLastError());
        PHG_HIP_CHECK(hipStreamSynchronize(s));
        return PHG_HIP_OK;
    }
    const int64_t nb = A->n / 3;
    int32_t *cnt = nullptr, *ovf = nullptr;
    int64_t *ptr64 = nullptr;
    void *tmp = nullptr;
    size_t tb = 0;
    int32_t h_ovf = 0;
    int64_t nbk = 0;
    auto cleanup = [&]() {
        dfree(cnt); dfree(ovf); dfree(ptr64); dfree(tmp);
    };
    if ((rc = dmalloc((void **)&cnt, (size_t)(nb + 1) * 4)) ||
        (rc = dmalloc((void **)&ovf, 4)) ||
        (rc = dmalloc((void **)&ptr64, (size_t)(nb + 1) * 8))) {
        cleanup();
        return rc;
    }
    PHG_HIP_CHECK(hipMemsetAsync(ovf, 0, 4, s));
    PHG_HIP_CHECK(hipMemsetAsync(cnt + nb, 0, 4, s));
    hipLaunchKernelGGL(k_bsr_count, dim3(nblk(nb, 128)), dim3(128), 0, s, nb,
                       (const int32_t *)A->q_rowptr, (const int32_t *)A->q_cols,
                       cnt, ovf);
    if ((rc = download(&h_ovf, ovf, 4))) {
        cleanup();
        return rc;
    }
    if (h_ovf) {
        cleanup();
        return set_error(PHG_HIP_E_ARG, "block row with more than %d blocks",
                         BSR_MAXB);
    }
    // exclusive scan of the counts on the host (nb + 1 values)
    {
        std::vector<int32_t> hc((size_t)nb);
        std::vector<int64_t> hp((size_t)nb + 1);
        if ((rc = download(hc.data(), cnt, (size_t)nb * 4))) {
            cleanup();
            return rc;
        }
        hp[0] = 0;
        for (int64_t i = 0; i < nb; i++)
            hp[(size_t)i + 1] = hp[(size_t)i] + hc[(size_t)i];
        nbk = hp[(size_t)nb];
        if (nbk >= INT32_MAX / 9) {
            cleanup();
            return set_error(PHG_HIP_E_ARG, "too many blocks");
        }
        if ((rc = upload(ptr64, hp.data(), hp.size() * 8))) {
            cleanup();
            return rc;
        }
    }
    if ((rc = dmalloc((void **)&A->b_ptr, (size_t)(nb + 1) * 4)) ||
        (rc = dmalloc((void **)&A->b_col, (size_t)(nbk + 1) * 4)) ||
        (rc = dmalloc((void **)&A->b_val64, (size_t)(nbk + 1) * 9 * 8)) ||
        (rc = dmalloc((void **)&A->b_val32, (size_t)(nbk + 1) * 9 * 4))) {
        A->b_owns = 1;
        free_bsr(A);
        cleanup();
        return rc;
    }
    A->b_owns = 1;
    hipLaunchKernelGGL(k_bsr_fill<double>, dim3(nblk(nb, 128)), dim3(128), 0, s,
                       nb, (const int32_t *)A->q_rowptr,
                       (const int32_t *)A->q_cols, (const double *)A->q_vals,
                       (const int64_t *)ptr64, A->b_col, A->b_val64);
    hipLaunchKernelGGL(k_bsr_fill<float>, dim3(nblk(nb, 128)), dim3(128), 0, s,
                       nb, (const int32_t *)A->q_rowptr,
                       (const int32_t *)A->q_cols, (const double *)A->q_vals,
                       (const int64_t *)ptr64, A->b_col, A->b_val32);
    hipLaunchKernelGGL(k_i64_to_i32, dim3(nblk(nb + 1, 256)), dim3(256), 0, s,
                       nb, (const int64_t *)ptr64, A->b_ptr);
    if ((rc = dmalloc((void **)&A->b_rsum32, (size_t)nb * 9 * 4))) {
        free_bsr(A);
        cleanup();
        return rc;
    }
    hipLaunchKernelGGL(k_bsr_rsum<3>, dim3(nblk(nb, 128)), dim3(128), 0, s, nb,
                       (const int32_t *)A->b_ptr, (const double *)A->b_val64, A->b_rsum32);
    PHG_HIP_CHECK(hipGetLastError());
    PHG_HIP_CHECK(hipStreamSynchronize(s));
    cleanup();
    A->b_nb = nb;
    A->b_nblk = nbk;
    A->b_dim = 3;
    return PHG_HIP_OK;
}

}  // namespace phg

namespace {

struct BlkWork {
    int64_t n = 0;
    double *b = nullptr, *x = nullptr, *r64 = nullptr;     // [n][KB]
    void *r = nullptr, *p = nullptr, *Ap = nullptr, *x32 = nullptr;
    double *part_it = nullptr, *part_up = nullptr, *part_s = nullptr;
    BState *st = nullptr;
    void release()
    {
        phg::dfree(b); phg::dfree(x); phg::dfree(r64); phg::dfree(r);
        phg::dfree(p); phg::dfree(Ap); phg::dfree(x32); phg::dfree(part_it);
        phg::dfree(part_up); phg::dfree(part_s); phg::dfree(st);
        *this = BlkWork();
    }
};

BlkWork g_bw;

int get_work(int64_t n)
{
    if (g_bw.n == n)
        return PHG_HIP_OK;
    g_bw.release();
    const size_t nk = (size_t)n * KB;
    int rc;
    if ((rc = phg::dmalloc((void **)&g_bw.b, nk * 8)) ||
        (rc = phg::dmalloc((void **)&g_bw.x, nk * 8)) ||
        (rc = phg::dmalloc(&g_bw.r, nk * 8)) ||
        (rc = phg::dmalloc(&g_bw.p, nk * 8)) ||
        (rc = phg::dmalloc(&g_bw.Ap, nk * 8)) ||
        (rc = phg::dmalloc(&g_bw.x32, nk * 8)) ||
        (rc = phg::dmalloc((void **)&g_bw.part_it, (size_t)4 * KB * MAXG * 8)) ||
        (rc = phg::dmalloc((void **)&g_bw.part_up, (size_t)4 * KB * MAXG * 8)) ||
        (rc = phg::dmalloc((void **)&g_bw.part_s, (size_t)KB * MAXG * 8)) ||
        (rc = phg::dmalloc((void **)&g_bw.st, sizeof(BState)))) {
        g_bw.release();
        return rc;
    }
    g_bw.n = n;
    return PHG_HIP_OK;
}

// row-block sums for the relative FP32 SpMM (null for FP64)
template <typename T>
const T *rsum_of(const phg_hip_mat *A)
{
    return sizeof(T) == 4 ? (const T *)(const void *)A->b_rsum32 : nullptr;
}

// block values of precision T: the CSR copies for 1x1 blocks
template <typename T, int B>
const void *mat_vals(const phg_hip_mat *A)
{
    if (B == 1)
        return sizeof(T) == 4 ? (const void *)A->q_vals32 : (const void *)A->q_vals;
    return sizeof(T) == 4 ? (const void *)A->b_val32 : (const void *)A->b_val64;
}

// one solve with element type T for the iteration and SpMM shape (B, S)
template <typename T, int B, int S, int CPL>
int block_solve(phg_hip_mat *A, int mixed, int32_t maxit, double rtol,
                double btol, double atol, double delta, BState *hfin)
{
    hipStream_t s = phg::ctx()->stream;
    const int64_t n = A->n, total = n * KB;
    const BsrView m{A->b_ptr, A->b_col, A->b_nb};
    const T *va = (const T *)mat_vals<T, B>(A);
    const double *va64 = B == 1 ? A->q_vals : A->b_val64;
    const T *diag = (const T *)(sizeof(T) == 4 ? (const void *)A->q_diag32
                                               : (const void *)A->q_diag);
    constexpr int RPG = rows_per_group<S, CPL>();
    static_assert(RPG > 0, "shape");
    const int64_t units = (A->b_nb + RPG - 1) / RPG;
    const int gs = grid_res<T>((const void *)k_blk_spmm<T, B, S, CPL>, units);
    const int gr = grid_res<T>((const void *)k_blk_resid<T, B, S, CPL>, units);
    const int gv = grid_vec(total);
    T *r = (T *)g_bw.r, *p = (T *)g_bw.p, *Ap = (T *)g_bw.Ap;
    // FP64 mode iterates on x itself; mixed mode on the FP32 increment
    T *xi = mixed ? (T *)g_bw.x32 : (T *)(void *)g_bw.x;
    BState h;
    memset(&h, 0, sizeof(h));
    h.rtol = rtol;
    h.btol = btol;
    h.atol = atol;
    h.delta = delta;
    h.maxit = maxit;
    h.mixed = mixed;
    h.kind = ST_START;
    for (int c = 0; c < KB; c++)
        h.active[c] = 1;
    PHG_HIP_CHECK(hipMemcpyAsync(g_bw.st, &h, sizeof(h), hipMemcpyHostToDevice, s));
    // prologue: r = b - A x in FP64 (x32 = 0 in mixed mode)
    hipLaunchKernelGGL((k_blk_resid<T, B, S, CPL>), dim3(gr), dim3(TB), 0, s, m,
                       va64, (const double *)A->q_diag, (const double *)g_bw.b,
                       (const double *)g_bw.x, r, mixed ? (T *)g_bw.x32 : (T *)nullptr,
                       g_bw.part_up, (const BState *)g_bw.st, 1);
    {
        hipError_t e = hipGetLastError();
        if (e != hipSuccess)
            return phg::set_error(PHG_HIP_E_HIP,
                                  "block PCG prologue launch (grid %d / %d / %d, "
                                  "%lld block rows): %s",
                                  gr, gs, gv, (long long)A->b_nb,
                                  hipGetErrorString(e));
    }
    BState *hm = (BState *)phg::ctx()->host_mailbox;
    static_assert(sizeof(BState) <= 4096, "mailbox size");
    int chunk = 8;
    int64_t slot = 0;
    for (;;) {
        for (int i = 0; i < chunk; i++, slot++) {
            hipLaunchKernelGGL(k_blk_ctrl, dim3(1), dim3(TB), 0, s,
                               (const double *)g_bw.part_it, gv,
                               (const double *)g_bw.part_up, gr, g_bw.st);
            hipLaunchKernelGGL(k_blk_p<T>, dim3(gv), dim3(TB), 0, s, n,
                               (const T *)r, diag, p, g_bw.x, (T *)g_bw.x32,
                               (const BState *)g_bw.st);
            hipLaunchKernelGGL((k_blk_spmm<T, B, S, CPL>), dim3(gs), dim3(TB), 0, s,
                               m, va, (const T *)p, Ap, g_bw.part_s,
                               (const BState *)g_bw.st, rsum_of<T>(A));
            hipLaunchKernelGGL(k_blk_alpha, dim3(1), dim3(TB), 0, s,
                               (const double *)g_bw.part_s, gs, g_bw.st);
            hipLaunchKernelGGL(k_blk_update<T>, dim3(gv), dim3(TB), 0, s, n, xi,
                               (const T *)p, r, (const T *)Ap, diag,
                               g_bw.part_it, (const BState *)g_bw.st);
            if (mixed)
                hipLaunchKernelGGL((k_blk_resid<T, B, S, CPL>), dim3(gr), dim3(TB), 0,
                                   s, m, va64, (const double *)A->q_diag,
                                   (const double *)g_bw.b,
                                   (const double *)g_bw.x, r, (T *)nullptr,
                                   g_bw.part_up, (const BState *)g_bw.st, 0);
        }
        PHG_HIP_CHECK(hipGetLastError());
        PHG_HIP_CHECK(hipMemcpyAsync(hm, g_bw.st, sizeof(BState),
                                     hipMemcpyDeviceToHost, s));
        PHG_HIP_CHECK(hipStreamSynchronize(s));
        if (hm->all_done)
            break;
        if (slot > 2 * (int64_t)maxit + 4096)
            return phg::set_error(PHG_HIP_E_HIP, "block PCG did not finish");
        chunk = std::min(chunk * 2, 64);
    }
    *hfin = *hm;
    if (mixed)
        hipLaunchKernelGGL(k_blk_fold, dim3(gv), dim3(TB), 0, s, total, g_bw.x,
                           (const float *)g_bw.x32);
    PHG_HIP_CHECK(hipGetLastError());
    return PHG_HIP_OK;
}

// SpMM shapes the tuner chooses from: (slots per row, columns per lane)
constexpr int NSHAPE = 6;
static_assert(NSHAPE <= (int)(sizeof(phg_hip_mat::spmm_us) / sizeof(phg_hip_mat::spmm_us[0])),
              "SpMM timing table too small");
const int kShapeS[NSHAPE] = {1, 2, 4, 4, 8, 16};
const int kShapeC[NSHAPE] = {1, 1, 1, 4, 4, 4};

template <int V> using IC = std::integral_constant<int, V>;

template <typename Fn>
int with_shape(int k, Fn &&f)
{
    switch (k) {
    case 0: return f(IC<1>{}, IC<1>{});
    case 1: return f(IC<2>{}, IC<1>{});
    case 2: return f(IC<4>{}, IC<1>{});
    case 3: return f(IC<4>{}, IC<4>{});
    case 4: return f(IC<8>{}, IC<4>{});
    default: return f(IC<16>{}, IC<4>{});
    }
}

template <typename T, int B>
int block_solve_s(phg_hip_mat *A, int shape, int mixed, int32_t maxit, double rtol,
                  double btol, double atol, double delta, BState *hfin)
{
    return with_shape(shape, [&](auto S, auto C) {
        return block_solve<T, B, decltype(S)::value, decltype(C)::value>(
            A, mixed, maxit, rtol, btol, atol, delta, hfin);
    });
}

// time `reps` SpMM launches of one shape (all columns active)
template <typename T, int B, int S, int CPL>
int time_spmm(phg_hip_mat *A, int reps, double *us)
{
    hipStream_t s = phg::ctx()->stream;
    const int64_t n = A->n;
    const BsrView m{A->b_ptr, A->b_col, A->b_nb};
    const T *va = (const T *)mat_vals<T, B>(A);
    constexpr int RPG = rows_per_group<S, CPL>();
    static_assert(RPG > 0, "shape");
    const int64_t units = (A->b_nb + RPG - 1) / RPG;
    const int gs = grid_res<T>((const void *)k_blk_spmm<T, B, S, CPL>, units);
    BState h;
    memset(&h, 0, sizeof(h));
    h.kind = ST_ITER;
    PHG_HIP_CHECK(hipMemcpyAsync(g_bw.st, &h, sizeof(h), hipMemcpyHostToDevice, s));
    PHG_HIP_CHECK(hipMemsetAsync(g_bw.p, 0, (size_t)n * KB * sizeof(T), s));
    hipEvent_t e0, e1;
    PHG_HIP_CHECK(hipEventCreate(&e0));
    PHG_HIP_CHECK(hipEventCreate(&e1));
    for (int i = 0; i < reps + 2; i++) {
        if (i == 2)
            PHG_HIP_CHECK(hipEventRecord(e0, s));
        hipLaunchKernelGGL((k_blk_spmm<T, B, S, CPL>), dim3(gs), dim3(TB), 0, s, m,
                           va, (const T *)g_bw.p, (T *)g_bw.Ap, g_bw.part_s,
                           (const BState *)g_bw.st, rsum_of<T>(A));
    }
    PHG_HIP_CHECK(hipEventRecord(e1, s));
    PHG_HIP_CHECK(hipEventSynchronize(e1));
    float ms = 0.f;
    PHG_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *us = ms * 1e3 / reps;
    return PHG_HIP_OK;
}

template <typename T, int B>
int time_spmm_s(phg_hip_mat *A, int shape, int reps, double *us)
{
    return with_shape(shape, [&](auto S, auto C) {
        return time_spmm<T, B, decltype(S)::value, decltype(C)::value>(A, reps, us);
    });
}

}  // namespace

//================================================================= C ABI

extern "C" int phg_hip_pcg_block(const phg_hip_mat *A_, int32_t nrhs,
                                 const phg_hip_vec *const *b,
                                 phg_hip_vec *const *x, int32_t maxit,
                                 double rtol, double btol, double atol,
                                 int32_t mixed, double delta, int32_t *nits,
                                 double *residual)
{
    int rc;
    if ((rc = phg::require_ctx()))
        return rc;
    phg_hip_mat *A = const_cast<phg_hip_mat *>(A_);
    if (A == nullptr || b == nullptr || x == nullptr || nrhs < 1 || nrhs > KB)
        return phg::set_error(PHG_HIP_E_ARG,
                              "phg_hip_pcg_block: 1 <= nrhs <= %d required", KB);
    if (!A->eliminated || A->d_diag == nullptr)
        return phg::set_error(PHG_HIP_E_ARG,
                              "matrix not assembled: call phg_hip_mat_eliminate");
    for (int c = 0; c < nrhs; c++)
        if (b[c] == nullptr || x[c] == nullptr || b[c]->n != A->n ||
            x[c]->n != A->n)
            return phg::set_error(PHG_HIP_E_ARG, "phg_hip_pcg_block: vector %d", c);
    if (mixed && !(delta > 0.0 && delta < 1.0))
        return phg::set_error(PHG_HIP_E_ARG, "delta must lie in (0, 1)");
    if ((rc = phg::build_bsr(A)) || (rc = get_work(A->n)))
        return rc;
    hipStream_t s = phg::ctx()->stream;
    const int64_t n = A->n, total = n * KB;
    const int gv = grid_vec(total);
    PtrTab t;
    memset(&t, 0, sizeof(t));
    for (int c = 0; c < nrhs; c++) {
        t.src[c] = b[c]->d;
        t.dst[c] = x[c]->d;
    }
    hipLaunchKernelGGL(k_blk_pack, dim3(gv), dim3(TB), 0, s, n, nrhs,
                       (const int32_t *)A->q_perm, t, g_bw.b);
    for (int c = 0; c < nrhs; c++)
        t.src[c] = x[c]->d;
    hipLaunchKernelGGL(k_blk_pack, dim3(gv), dim3(TB), 0, s, n, nrhs,
                       (const int32_t *)A->q_perm, t, g_bw.x);
    PHG_HIP_CHECK(hipGetLastError());
    BState fin;
    const int S = A->spmm_shape[mixed ? 1 : 0];
    if (A->b_dim == 3)
        rc = mixed ? block_solve_s<float, 3>(A, S, 1, maxit, rtol, btol, atol, delta, &fin)
                   : block_solve_s<double, 3>(A, S, 0, maxit, rtol, btol, atol, delta, &fin);
    else
        rc = mixed ? block_solve_s<float, 1>(A, S, 1, maxit, rtol, btol, atol, delta, &fin)
                   : block_solve_s<double, 1>(A, S, 0, maxit, rtol, btol, atol, delta, &fin);
    if (rc)
        return rc;
    int zmask = 0;
    for (int c = 0; c < nrhs; c++)
        if (fin.zero_x[c])
            zmask |= 1 << c;
    hipLaunchKernelGGL(k_blk_unpack, dim3(gv), dim3(TB), 0, s, n, nrhs,
                       (const int32_t *)A->q_perm, (const double *)g_bw.x, t,
                       zmask);
    PHG_HIP_CHECK(hipGetLastError());
    PHG_HIP_CHECK(hipStreamSynchronize(s));
    int32_t mx = 0;
    for (int c = 0; c < nrhs; c++) {
        const int32_t it = fin.nits[c];
        if (nits)
            nits[c] = it;
        if (residual)
            residual[c] = fin.res[c];
        mx = std::max(mx, it);
    }
    return mx;
}

extern "C" int phg_hip_spmm_tune(phg_hip_mat *A, int32_t reps)
{
    int rc;
    if ((rc = phg::require_ctx()))
        return rc;
    if (A == nullptr || !A->eliminated)
        return phg::set_error(PHG_HIP_E_ARG, "phg_hip_spmm_tune: matrix not assembled");
    if ((rc = phg::build_bsr(A)) || (rc = get_work(A->n)))
        return rc;
    if (reps <= 0)
        reps = 10;
    for (int prec = 0; prec < 2; prec++) {
        int best = 0;
        for (int k = 0; k < NSHAPE; k++) {
            double us = 0.0;
            const int S = k;
            if (A->b_dim == 3)
                rc = prec ? time_spmm_s<float, 3>(A, S, reps, &us)
                          : time_spmm_s<double, 3>(A, S, reps, &us);
            else
                rc = prec ? time_spmm_s<float, 1>(A, S, reps, &us)
                          : time_spmm_s<double, 1>(A, S, reps, &us);
            if (rc)
                return rc;
            A->spmm_us[k][prec] = us;
            if (us < A->spmm_us[best][prec])
                best = k;
        }
        A->spmm_shape[prec] = best;
    }
    return PHG_HIP_OK;
}

extern "C" int phg_hip_spmm_tune_time(const phg_hip_mat *A, int32_t shape,
                                      int32_t fp32, double *us)
{
    if (A == nullptr || us == nullptr || shape < 0 || shape >= NSHAPE)
        return phg::set_error(PHG_HIP_E_ARG, "phg_hip_spmm_tune_time: bad args");
    *us = A->spmm_us[shape][fp32 ? 1 : 0];
    return PHG_HIP_OK;
}

extern "C" int phg_hip_spmm_bytes(const phg_hip_mat *A, int32_t fp32,
                                  int64_t *bytes)
{
    if (A == nullptr || bytes == nullptr || A->b_dim == 0)
        return phg::set_error(PHG_HIP_E_STATE, "block layout not built");
    const int64_t sv = fp32 ? 4 : 8, B = A->b_dim;
    // blocks (values + column) + block row pointers + one read of X and
    // one write of Y, KB columns each
    *bytes = A->b_nblk * (B * B * sv + 4) + (A->b_nb + 1) * 4 +
             2 * A->n * KB * sv;
    return PHG_HIP_OK;
}

extern "C" int phg_hip_spmm_set_shape(phg_hip_mat *A, int32_t shape64,
                                      int32_t shape32)
{
    if (A == nullptr || shape64 < 0 || shape64 >= NSHAPE || shape32 < 0 ||
        shape32 >= NSHAPE)
        return phg::set_error(PHG_HIP_E_ARG, "phg_hip_spmm_set_shape: bad args");
    A->spmm_shape[0] = shape64;
    A->spmm_shape[1] = shape32;
    return PHG_HIP_OK;
}
